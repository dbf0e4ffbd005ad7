// Workgroup-level dense complex linear algebra used by every NEGF kernel.
//
//  * wg_gemm       C(m x n) = opA * opB on the matrix cores (16x16x4 tiles),
//                  results held in registers until every operand read is
//                  done, so C may alias A or B (in-place products).
//  * wg_gj_inverse Gauss-Jordan inverse with partial pivoting, matrix held in
//                  registers (each thread owns one column stripe), pivot row
//                  and column exchanged through LDS; no physical row swaps.
//
// Operands are passed as accessor functors (i, k) -> cx<R> so that the same
// code reads LDS tiles, global Hamiltonian blocks or "H - z S" on the fly.
#pragma once
#include "common.hpp"

namespace negf {

// C = sum_k A(i,k) B(k,j) for i<m, j<n, k<kd.  Every thread of the workgroup
// must call it.  `epi(i, j, value)` is invoked once per output element after a
// workgroup barrier that follows the last operand read.
template <typename R, int NW, int MAXT, class FA, class FB, class EPI>
__device__ __forceinline__ void wg_gemm(int m, int n, int kd, FA ga, FB gb, EPI epi) {
  typedef typename mfma<R>::v4 v4;
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = threadIdx.x / WAVE;
  const int tn = (n + 15) >> 4;
  const int nt = ((m + 15) >> 4) * tn;
  v4 cr[MAXT], ci[MAXT];
#pragma unroll
  for (int s = 0; s < MAXT; ++s) {
    cr[s] = v4{0, 0, 0, 0};
    ci[s] = v4{0, 0, 0, 0};
  }
  const int li = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < kd; k0 += 4) {
    const int kk = k0 + lk;
#pragma unroll
    for (int s = 0; s < MAXT; ++s) {
      const int t = wave + NW * s;
      if (t < nt) {  // wave-uniform
        const int i = (t / tn) * 16 + li;
        const int j = (t % tn) * 16 + li;
        cx<R> a = {0, 0}, b = {0, 0};
        if (i < m && kk < kd) a = ga(i, kk);
        if (kk < kd && j < n) b = gb(kk, j);
        cr[s] = mfma<R>::op(a.re, b.re, cr[s]);
        cr[s] = mfma<R>::op(-a.im, b.im, cr[s]);
        ci[s] = mfma<R>::op(a.re, b.im, ci[s]);
        ci[s] = mfma<R>::op(a.im, b.re, ci[s]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < MAXT; ++s) {
    const int t = wave + NW * s;
    if (t < nt) {
      const int j = (t % tn) * 16 + li;
      const int ib = (t / tn) * 16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = ib + mfma<R>::row(lk, r);
        if (i < m && j < n) epi(i, j, cx<R>{cr[s][r], ci[s][r]});
      }
    }
  }
}

template <typename R, int NMAX>
struct GJScratch {
  cx<R> col[2][NMAX];
  cx<R> row[NMAX];
  int piv[NMAX];
  int pinv[NMAX];
  int used[NMAX];
  int singular;
};

// In-place style inverse: reads the matrix through load(i,k), writes the
// inverse through store(i,k,v).  Both may address the same storage.  NT is
// the workgroup size; a thread owns column k = t % n and rows g + G*r with
// G = NT / n, so n <= NT and ceil(n / G) <= MAXR are required.  Returns
// nonzero (uniformly) if an exactly-zero pivot was met.
//
// Pivot choice follows LAPACK's izamax: largest |re| + |im| among the rows
// not yet used, first index on ties.  After the sweep the registers hold Y
// with A^{-1}[j][piv[j']] = Y[piv[j]][j'], undone while storing.
template <typename R, int NT, int NMAX, int MAXR, class FL, class FS>
__device__ __forceinline__ int wg_gj_inverse(int n, FL load, FS store, GJScratch<R, NMAX>* sh) {
  const int t = threadIdx.x;
  const int lane = t & (WAVE - 1);
  const int G = NT / n;
  const int k = t % n, g = t / n;
  const bool act = g < G;
  cx<R> v[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    const int i = g + G * r;
    v[r] = (act && i < n) ? load(i, k) : cx<R>{0, 0};
  }
  for (int i = t; i < n; i += NT) sh->used[i] = 0;
  if (t == 0) sh->singular = 0;
  if (act && k == 0) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int i = g + G * r;
      if (i < n) sh->col[0][i] = v[r];
    }
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const cx<R>* cb = sh->col[j & 1];
    // pivot search, redundantly in every wave
    R best = R(-1);
    int bi = n;
    for (int i = lane; i < n; i += WAVE) {
      if (!sh->used[i]) {
        const R m_ = abs1(cb[i]);
        if (m_ > best) { best = m_; bi = i; }
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const R ob = __shfl_xor(best, o, WAVE);
      const int oi = __shfl_xor(bi, o, WAVE);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const int p = bi;
    if (act) {
#pragma unroll
      for (int r = 0; r < MAXR; ++r)
        if (g + G * r == p) sh->row[k] = v[r];
    }
    __syncthreads();
    const cx<R> d = cb[p];
    if (t == 0) {
      sh->used[p] = 1;
      sh->piv[j] = p;
      if (d.re == R(0) && d.im == R(0)) sh->singular = 1;
    }
    const cx<R> dinv = recip(d);
    const cx<R> rs = sh->row[k] * dinv;
    cx<R>* nb = sh->col[(j + 1) & 1];
    if (act) {
#pragma unroll
      for (int r = 0; r < MAXR; ++r) {
        const int i = g + G * r;
        if (i < n) {
          if (i == p) {
            v[r] = (k == j) ? dinv : rs;
          } else {
            const cx<R> f = cb[i];
            v[r] = (k == j) ? -(f * dinv) : v[r] - f * rs;
          }
          if (k == j + 1) nb[i] = v[r];
        }
      }
    }
    __syncthreads();
  }
  for (int j = t; j < n; j += NT) sh->pinv[sh->piv[j]] = j;
  __syncthreads();
  const int sing = sh->singular;
  if (act) {
    const int pc = sh->piv[k];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int i = g + G * r;
      if (i < n) store(sh->pinv[i], pc, v[r]);
    }
  }
  return sing;
}

}  // namespace negf
