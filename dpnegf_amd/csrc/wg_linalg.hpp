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
// workgroup barrier that follows the last operand read.  SYNC = false drops
// that barrier, for epilogues that write nothing any tile reads.  Tile t goes
// to wave (t - ROT) mod NW, so ROT = 1 leaves wave 0 with the fewest tiles.
template <typename R, int NW, int MAXT, bool SYNC = true, int ROT = 0, class FA, class FB, class EPI>
__device__ __forceinline__ void wg_gemm(int m, int n, int kd, FA ga, FB gb, EPI epi) {
  typedef typename mfma<R>::v4 v4;
  const int lane = threadIdx.x & (WAVE - 1);
  const int wave = (threadIdx.x / WAVE + NW - ROT) % NW;
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
  if (SYNC) __syncthreads();
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

// wg_gemm for compile-time shapes (M, N multiples of 16, K a multiple of 4):
// no bounds checks, and tiles past the end are clamped to a valid one (their
// results are dropped), so the k-loop has no branches.  PFA / PFB: the A / B
// operand of every tile for step k + 1 is loaded before the matrix-core
// instructions of step k are issued, so a long-latency operand (a Hamiltonian
// block read through L2) overlaps the MFMAs; LDS operands need no prefetch.
// Same contract as wg_gemm otherwise.
template <typename R, int NW, int M, int N, int K, bool PFA, bool PFB, bool SYNC = true,
          int ROT = 0, class FA, class FB, class EPI>
__device__ __forceinline__ void wg_gemm_fixed(FA ga, FB gb, EPI epi) {
  typedef typename mfma<R>::v4 v4;
  static_assert(M % 16 == 0 && N % 16 == 0 && K % 4 == 0, "wg_gemm_fixed shape");
  constexpr int TN = N / 16;
  constexpr int NT = (M / 16) * TN;
  constexpr int MAXT = (NT + NW - 1) / NW;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = (threadIdx.x / WAVE + NW - ROT) % NW;
  const int li = lane & 15, lk = lane >> 4;
  int ti[MAXT], tj[MAXT];
#pragma unroll
  for (int s = 0; s < MAXT; ++s) {
    const int t = w + NW * s;
    const int tc = t < NT ? t : NT - 1;
    ti[s] = (tc / TN) * 16 + li;
    tj[s] = (tc % TN) * 16 + li;
  }
  v4 cr[MAXT], ci[MAXT];
  cx<R> pa[PFA ? MAXT : 1], pb[PFB ? MAXT : 1];
#pragma unroll
  for (int s = 0; s < MAXT; ++s) {
    cr[s] = v4{0, 0, 0, 0};
    ci[s] = v4{0, 0, 0, 0};
    if constexpr (PFA) pa[s] = ga(ti[s], lk);
    if constexpr (PFB) pb[s] = gb(lk, tj[s]);
  }
  for (int k0 = 0; k0 < K; k0 += 4) {
    cx<R> na[PFA ? MAXT : 1], nb[PFB ? MAXT : 1];
    if (k0 + 4 < K) {
#pragma unroll
      for (int s = 0; s < MAXT; ++s) {
        if constexpr (PFA) na[s] = ga(ti[s], k0 + 4 + lk);
        if constexpr (PFB) nb[s] = gb(k0 + 4 + lk, tj[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < MAXT; ++s) {
      cx<R> a, b;
      if constexpr (PFA) a = pa[s]; else a = ga(ti[s], k0 + lk);
      if constexpr (PFB) b = pb[s]; else b = gb(k0 + lk, tj[s]);
      cr[s] = mfma<R>::op(a.re, b.re, cr[s]);
      cr[s] = mfma<R>::op(-a.im, b.im, cr[s]);
      ci[s] = mfma<R>::op(a.re, b.im, ci[s]);
      ci[s] = mfma<R>::op(a.im, b.re, ci[s]);
    }
    if (k0 + 4 < K) {
#pragma unroll
      for (int s = 0; s < MAXT; ++s) {
        if constexpr (PFA) pa[s] = na[s];
        if constexpr (PFB) pb[s] = nb[s];
      }
    }
  }
  if (SYNC) __syncthreads();
#pragma unroll
  for (int s = 0; s < MAXT; ++s) {
    if (w + NW * s < NT) {
      const int ib = ti[s] - li;
#pragma unroll
      for (int r = 0; r < 4; ++r) epi(ib + mfma<R>::row(lk, r), tj[s], cx<R>{cr[s][r], ci[s][r]});
    }
  }
}

template <typename R, int NMAX>
struct GJScratch {
  cx<R> col[2][NMAX];   // current pivot column, double-buffered across steps
  cx<R> row[NMAX];      // current pivot row
  R pval[2][NMAX];      // per-row-group partial pivot magnitude of the next column
  int pidx[2][NMAX];    // ... and its row
  int piv[NMAX];
  int pinv[NMAX];
  int singular;
};

// 1/d through one division: conj(d) / |d|^2 (block pivots are far from the
// overflow range, so Smith's scaling is not needed here).
template <typename R>
__device__ __forceinline__ cx<R> recip_fast(cx<R> d) {
  const R s = R(1) / (d.re * d.re + d.im * d.im);
  return {d.re * s, -d.im * s};
}

// 1/d = conj(d) / |d|^2, with 1/|d|^2 from the hardware reciprocal refined by
// Newton steps (two in FP64, one in FP32: a few ulp, and a shorter dependency
// chain than the IEEE division sequence).  Pivots are far from the overflow /
// underflow range of |d|^2.
__device__ __forceinline__ cx<double> recip_nr(cx<double> d) {
  const double m = d.re * d.re + d.im * d.im;
  double r = __builtin_amdgcn_rcp(m);
  double e = fma(-m, r, 1.0);
  r = fma(r, e, r);
  e = fma(-m, r, 1.0);
  r = fma(r, e, r);
  return {d.re * r, -d.im * r};
}
__device__ __forceinline__ cx<float> recip_nr(cx<float> d) {
  const float m = d.re * d.re + d.im * d.im;
  float r = __builtin_amdgcn_rcpf(m);
  r = fmaf(r, fmaf(-m, r, 1.0f), r);
  return {d.re * r, -d.im * r};
}

// Pivot key: LAPACK izamax's |re| + |im|; NaN counts as 0 so that a row is
// always selectable (the selection below never returns an out-of-range row).
template <typename R>
__device__ __forceinline__ R pivot_key(cx<R> a) {
  // max(m, 0) is m, or 0 for a NaN m (maxNum): one instruction on the chain
  // instead of a compare and two selects
  return __builtin_fmax(abs1(a), R(0));
}

// In-place style inverse: reads the matrix through load(i,k), writes the
// inverse through store(i,k,v).  Both may address the same storage.  NT is
// the workgroup size; a thread owns column k = t % n and rows g + G*r with
// G = NT / n, so n <= NT and ceil(n / G) <= MAXR are required.  Returns
// nonzero (uniformly) if an exactly-zero pivot was met.
//
// Pivot choice follows LAPACK's izamax: largest |re| + |im| among the rows
// not yet used, first index on ties.  The owners of the next column publish
// one (magnitude, row) partial per row group while they update it, so each
// step reduces at most min(G, n) partials from LDS instead of shuffling.
// After the sweep the registers hold Y with A^{-1}[j][piv[j']] = Y[piv[j]][j'],
// undone while storing.
template <typename R, int NT, int NMAX, int MAXR, class FL, class FS>
__device__ __forceinline__ int wg_gj_inverse(int n, FL load, FS store, GJScratch<R, NMAX>* sh) {
  static_assert(MAXR <= 64, "row mask is 64 bits");
  const int t = threadIdx.x;
  const int G = NT / n;
  const int k = t % n, g = t / n;
  const bool act = g < G;
  const int ng = G < n ? G : n;  // row groups that own rows
  uint64_t used = 0;             // bit r: row g + G*r already a pivot row
  cx<R> v[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) {
    const int i = g + G * r;
    v[r] = (act && i < n) ? load(i, k) : cx<R>{0, 0};
  }
  if (t == 0) sh->singular = 0;
  // publish column 0 and its partial pivots
  if (act && k == 0) {
    R best = R(-1);
    int bi = n;
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      const int i = g + G * r;
      if (i < n) {
        sh->col[0][i] = v[r];
        const R m = pivot_key(v[r]);
        if (m > best) { best = m; bi = i; }
      }
    }
    if (g < ng) { sh->pval[0][g] = best; sh->pidx[0][g] = bi; }
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const int b = j & 1;
    R best = R(-1);
    int bi = n;
    for (int q = 0; q < ng; ++q) {
      const R m = sh->pval[b][q];
      const int i = sh->pidx[b][q];
      if (m > best || (m == best && i < bi)) { best = m; bi = i; }
    }
    const int p = bi < n ? bi : n - 1;
    if (act && p % G == g) used |= 1ull << (p / G);
    if (act) {
#pragma unroll
      for (int r = 0; r < MAXR; ++r)
        if (g + G * r == p) sh->row[k] = v[r];
    }
    __syncthreads();
    const cx<R> d = sh->col[b][p];
    if (t == 0) {
      sh->piv[j] = p;
      if (d.re == R(0) && d.im == R(0)) sh->singular = 1;
    }
    const cx<R> dinv = recip_fast(d);
    const cx<R> rs = sh->row[k] * dinv;
    const cx<R>* cb = sh->col[b];
    const bool colj = (k == j);
    const cx<R> prow = colj ? dinv : rs;
    if constexpr (NT <= 256 && MAXR <= 16) {
      // every column read first (rows past n read a valid dummy), then selects
      cx<R> f[MAXR];
#pragma unroll
      for (int r = 0; r < MAXR; ++r) {
        const int i = g + G * r;
        f[r] = cb[i < n ? i : 0];
      }
#pragma unroll
      for (int r = 0; r < MAXR; ++r) {
        const cx<R> upd = colj ? -(f[r] * dinv) : v[r] - f[r] * rs;
        v[r] = (g + G * r == p) ? prow : upd;
      }
    } else {
      // many rows per thread (1024-thread inverse): keep register pressure low
#pragma unroll
      for (int r = 0; r < MAXR; ++r) {
        const int i = g + G * r;
        if (i < n) {
          if (i == p) {
            v[r] = prow;
          } else {
            const cx<R> f = cb[i];
            v[r] = colj ? -(f * dinv) : v[r] - f * rs;
          }
        }
      }
    }
    if (act) {
      const bool nxt = (k == j + 1);
      R nbest = R(-1);
      int nbi = n;
#pragma unroll
      for (int r = 0; r < MAXR; ++r) {
        const int i = g + G * r;
        if (i < n) {
          if (nxt) {
            sh->col[b ^ 1][i] = v[r];
            if (!((used >> r) & 1)) {
              const R m = pivot_key(v[r]);
              if (m > nbest) { nbest = m; nbi = i; }
            }
          }
        }
      }
      if (nxt && g < ng) { sh->pval[b ^ 1][g] = nbest; sh->pidx[b ^ 1][g] = nbi; }
    }
    __syncthreads();
  }
  // pinv starts as the identity so every store index stays in [0, n) even if a
  // pivot row were ever selected twice
  for (int j = t; j < n; j += NT) sh->pinv[j] = j;
  __syncthreads();
  for (int j = t; j < n; j += NT) {
    const int pj = sh->piv[j];
    if (pj >= 0 && pj < n) sh->pinv[pj] = j;
  }
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
  __syncthreads();
  return sing;
}



// ---- blocked Gauss-Jordan (matrix in LDS) -------------------------------------
template <typename R, int N>
struct GJBScratch {
  cx<R> F[N][4];  // rank-4 multipliers of the current panel
  int pr[4];      // pivot rows of the current panel
  short piv[N];   // piv[j]: row pivoting column j (16 bits: the 48-orbital
  short pinv[N];  // transmission sweep fits four workgroups' LDS in one CU)
  int singular;
};

__device__ __forceinline__ double bcast_lane(double x, int src) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast_lane(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
template <typename R>
__device__ __forceinline__ cx<R> bcast_lane(cx<R> x, int src) {
  return {bcast_lane(x.re, src), bcast_lane(x.im, src)};
}

// Blocked Gauss-Jordan in-place inverse of the N x N matrix A in LDS (leading
// dimension LDA), N <= 64 and a multiple of 4, workgroup of NT threads; rows
// are never swapped (piv[j] is the row that pivots column j).
//
// Panels of BW = 4 columns.  Wave 0 sweeps a panel one row per lane (pivot =
// largest |re| + |im| among unused rows, compared in single precision to 17
// mantissa bits, lowest row on ties, found with a DPP wave maximum; pivot row
// broadcast with readlane), recording each row's multiplier F[i][k] (its
// column-k value at step k).  Each step eliminates with the standard
// multiplier, row -= (f / d) prow, the panel column taking 1 in the pivot
// row and -f / d elsewhere, and leaves the pivot row unscaled; the pivot rows
// are scaled by 1 / d once at the end of the panel (an unscaled row stays d
// times the scaled one through every later step), and so are their
// multipliers of later steps.  For the other columns C the same steps are,
// exactly as in sequential order:
//   pivot rows  u_k = (Z_k - sum_{k'<k} F[p_k][k'] u_k') / d_k   (Z: rows before
//               the panel; forward substitution by wave 0, written over Z)
//   all rows    A[i, C] -= sum_k Fm[i][k] u_k   (rank-4, matrix cores)
// with Fm = F except that a pivot row p_k drops steps <= k (already in u_k).
// This is the blocked-LU form of the update; applying the panel's combined
// transform to the unreduced pivot rows instead (an inverse of the 4 x 4 pivot
// block times Z) costs about two digits at ill-conditioned points.
// Returns nonzero (uniformly) if an exactly-zero pivot was met.
template <typename R, int NT, int N, int LDA>
__device__ __forceinline__ int wg_gj_blocked_core(cx<R>* A, GJBScratch<R, N>* sh) {
  constexpr int BW = 4;
  constexpr int NW = NT / WAVE;
  constexpr int NTILE = ((N + 15) / 16) * ((N + 15) / 16);
  constexpr int MAXT = (NTILE + NW - 1) / NW;
  static_assert(N <= WAVE && N % BW == 0, "blocked Gauss-Jordan: N <= 64, multiple of 4");
  const int t = threadIdx.x;
  const int lane = t & (WAVE - 1);
  const int wave = t / WAVE;
  if (t == 0) sh->singular = 0;
  bool used = false;
  for (int j0 = 0; j0 < N; j0 += BW) {
    if (wave == 0) {
      const bool row_ok = lane < N;
      cx<R> v[BW], F[BW], dinv[BW], L[BW][BW];
      int pr[BW];
      int pstep = BW;  // step at which this lane's row pivots (BW: none)
#pragma unroll
      for (int c = 0; c < BW; ++c) v[c] = row_ok ? A[lane * LDA + j0 + c] : cx<R>{0, 0};
#pragma unroll
      for (int jj = 0; jj < BW; ++jj) {
        unsigned key = 0;
        if (row_ok && !used) {
          const float kf = (float)pivot_key(v[jj]);
          key = (((__float_as_uint(kf) >> 6) + 1u) << 6) | (unsigned)(WAVE - 1 - lane);
        }
        key = wave_max_u32(key);
        int p = WAVE - 1 - (int)(key & 63u);
        p = p < N ? p : N - 1;
        pr[jj] = p;
        const bool isp = (lane == p);
        if (isp) { used = true; pstep = jj; }
        F[jj] = v[jj];
#pragma unroll
        for (int k2 = 0; k2 < jj; ++k2) L[jj][k2] = bcast_lane(F[k2], p);
        const cx<R> d = bcast_lane(F[jj], p);
        dinv[jj] = recip_nr(d);
        if (lane == 0) {
          sh->piv[j0 + jj] = p;
          sh->pr[jj] = p;
          if (d.re == R(0) && d.im == R(0)) sh->singular = 1;
        }
        const cx<R> m0 = F[jj] * dinv[jj];
        const cx<R> m = {isp ? R(0) : m0.re, isp ? R(0) : m0.im};
#pragma unroll
        for (int c = 0; c < BW; ++c) {
          if (c == jj) continue;
          const cx<R> pc = bcast_lane(v[c], p);
          v[c].re = fma(-m.re, pc.re, fma(m.im, pc.im, v[c].re));
          v[c].im = fma(-m.re, pc.im, fma(-m.im, pc.re, v[c].im));
        }
        v[jj] = cx<R>{isp ? R(1) : -m.re, isp ? R(0) : -m.im};
      }
      cx<R> s = {R(1), R(0)};
#pragma unroll
      for (int k = 0; k < BW; ++k)
        if (pstep == k) s = dinv[k];
      if (row_ok) {
#pragma unroll
        for (int c = 0; c < BW; ++c) {
          A[lane * LDA + j0 + c] = v[c] * s;
          sh->F[lane][c] = (pstep < BW && c <= pstep) ? cx<R>{0, 0} : F[c] * s;
        }
      }
      // pivot rows of the other columns: forward substitution in step order
      for (int c = lane; c < N; c += WAVE) {
        if (c >= j0 && c < j0 + BW) continue;
        cx<R> u[BW];
#pragma unroll
        for (int k = 0; k < BW; ++k) {
          cx<R> acc = A[pr[k] * LDA + c];
#pragma unroll
          for (int k2 = 0; k2 < k; ++k2) acc = acc - L[k][k2] * u[k2];
          u[k] = acc * dinv[k];
        }
#pragma unroll
        for (int k = 0; k < BW; ++k) A[pr[k] * LDA + c] = u[k];
      }
    }
    __syncthreads();
    const int p0 = sh->pr[0], p1 = sh->pr[1], p2 = sh->pr[2], p3 = sh->pr[3];
    wg_gemm<R, NW, MAXT, true, 1>(
        N, N, BW, [&](int i, int k) { return sh->F[i][k]; },
        [&](int k, int c) {
          const int pk = k == 0 ? p0 : k == 1 ? p1 : k == 2 ? p2 : p3;
          return A[pk * LDA + c];
        },
        [&](int i, int c, cx<R> val) {
          if (c < j0 || c >= j0 + BW) A[i * LDA + c] = A[i * LDA + c] - val;
        });
    __syncthreads();
  }
  return sh->singular;
}

// The same blocked Gauss-Jordan for 64 < N <= 128: the sweeping wave holds two
// rows per lane (rows lane and lane + 64), and the pivot key carries the row in
// its low 7 bits (16 mantissa bits of |re| + |im| compared, lowest row on
// ties).  Steps, multipliers, pivot-row scaling, forward substitution and the
// rank-4 matrix-core update are those of wg_gj_blocked_core.
template <typename R, int NT, int N, int LDA>
__device__ __forceinline__ int wg_gj_blocked_core_tall(cx<R>* A, GJBScratch<R, N>* sh) {
  constexpr int BW = 4;
  constexpr int RPL = 2;
  constexpr int NW = NT / WAVE;
  constexpr int NTILE = ((N + 15) / 16) * ((N + 15) / 16);
  constexpr int MAXT = (NTILE + NW - 1) / NW;
  static_assert(N > WAVE && N <= RPL * WAVE && N % BW == 0,
                "tall blocked Gauss-Jordan: 64 < N <= 128, multiple of 4");
  const int t = threadIdx.x;
  const int lane = t & (WAVE - 1);
  const int wave = t / WAVE;
  if (t == 0) sh->singular = 0;
  bool used[RPL] = {false, false};
  bool row_ok[RPL];
#pragma unroll
  for (int j = 0; j < RPL; ++j) row_ok[j] = lane + WAVE * j < N;
  for (int j0 = 0; j0 < N; j0 += BW) {
    if (wave == 0) {
      cx<R> v[RPL][BW], F[RPL][BW], dinv[BW], L[BW][BW];
      int pr[BW];
      int pstep[RPL] = {BW, BW};  // step at which each of this lane's rows pivots (BW: none)
#pragma unroll
      for (int j = 0; j < RPL; ++j)
#pragma unroll
        for (int c = 0; c < BW; ++c)
          v[j][c] = row_ok[j] ? A[(lane + WAVE * j) * LDA + j0 + c] : cx<R>{0, 0};
#pragma unroll
      for (int jj = 0; jj < BW; ++jj) {
        unsigned key = 0;
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          if (row_ok[j] && !used[j]) {
            const float kf = (float)pivot_key(v[j][jj]);
            const unsigned kj = (((__float_as_uint(kf) >> 7) + 1u) << 7) |
                                (unsigned)(2 * WAVE - 1 - (lane + WAVE * j));
            key = kj > key ? kj : key;
          }
        }
        key = wave_max_u32(key);
        int p = 2 * WAVE - 1 - (int)(key & 127u);
        p = p < N ? p : N - 1;
        const int pl = p & (WAVE - 1);
        const bool hi = p >= WAVE;  // uniform: the pivot row is the lane's second row
        // broadcast both candidates and select the value (selecting the
        // register by index would put the arrays in scratch memory)
        auto pick = [&](cx<R> x0, cx<R> x1) {
          const cx<R> a = bcast_lane(x0, pl), b = bcast_lane(x1, pl);
          return hi ? b : a;
        };
        pr[jj] = p;
        bool isp[RPL];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          isp[j] = (lane + WAVE * j == p);
          if (isp[j]) { used[j] = true; pstep[j] = jj; }
          F[j][jj] = v[j][jj];
        }
#pragma unroll
        for (int k2 = 0; k2 < jj; ++k2) L[jj][k2] = pick(F[0][k2], F[1][k2]);
        const cx<R> d = pick(F[0][jj], F[1][jj]);
        dinv[jj] = recip_nr(d);
        if (lane == 0) {
          sh->piv[j0 + jj] = p;
          sh->pr[jj] = p;
          if (d.re == R(0) && d.im == R(0)) sh->singular = 1;
        }
        cx<R> m[RPL];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
          const cx<R> m0 = F[j][jj] * dinv[jj];
          m[j] = cx<R>{isp[j] ? R(0) : m0.re, isp[j] ? R(0) : m0.im};
        }
#pragma unroll
        for (int c = 0; c < BW; ++c) {
          if (c == jj) continue;
          const cx<R> pc = pick(v[0][c], v[1][c]);
#pragma unroll
          for (int j = 0; j < RPL; ++j) {
            v[j][c].re = fma(-m[j].re, pc.re, fma(m[j].im, pc.im, v[j][c].re));
            v[j][c].im = fma(-m[j].re, pc.im, fma(-m[j].im, pc.re, v[j][c].im));
          }
        }
#pragma unroll
        for (int j = 0; j < RPL; ++j)
          v[j][jj] = cx<R>{isp[j] ? R(1) : -m[j].re, isp[j] ? R(0) : -m[j].im};
      }
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        cx<R> sc = {R(1), R(0)};
#pragma unroll
        for (int k = 0; k < BW; ++k)
          if (pstep[j] == k) sc = dinv[k];
        if (row_ok[j]) {
          const int row = lane + WAVE * j;
#pragma unroll
          for (int c = 0; c < BW; ++c) {
            A[row * LDA + j0 + c] = v[j][c] * sc;
            sh->F[row][c] = (pstep[j] < BW && c <= pstep[j]) ? cx<R>{0, 0} : F[j][c] * sc;
          }
        }
      }
      // pivot rows of the other columns: forward substitution in step order
      for (int c = lane; c < N; c += WAVE) {
        if (c >= j0 && c < j0 + BW) continue;
        cx<R> u[BW];
#pragma unroll
        for (int k = 0; k < BW; ++k) {
          cx<R> acc = A[pr[k] * LDA + c];
#pragma unroll
          for (int k2 = 0; k2 < k; ++k2) acc = acc - L[k][k2] * u[k2];
          u[k] = acc * dinv[k];
        }
#pragma unroll
        for (int k = 0; k < BW; ++k) A[pr[k] * LDA + c] = u[k];
      }
    }
    __syncthreads();
    const int p0 = sh->pr[0], p1 = sh->pr[1], p2 = sh->pr[2], p3 = sh->pr[3];
    wg_gemm<R, NW, MAXT, true, 1>(
        N, N, BW, [&](int i, int k) { return sh->F[i][k]; },
        [&](int k, int c) {
          const int pk = k == 0 ? p0 : k == 1 ? p1 : k == 2 ? p2 : p3;
          return A[pk * LDA + c];
        },
        [&](int i, int c, cx<R> val) {
          if (c < j0 || c >= j0 + BW) A[i * LDA + c] = A[i * LDA + c] - val;
        });
    __syncthreads();
  }
  return sh->singular;
}

// The blocked Gauss-Jordan of an N x N matrix (N <= 64) left in pivot order:
// A^-1[i][k] = A[piv[i]][pinv[k]] with piv / pinv in sh on return, so a caller
// that reads through the two maps saves the reordering pass (and its
// registers) of wg_gj_inverse_blocked.  Ends with a workgroup barrier; returns
// nonzero if an exactly-zero pivot was met.
template <typename R, int NT, int N, int LDA>
__device__ __forceinline__ int wg_gj_inverse_pivoted(cx<R>* A, GJBScratch<R, N>* sh) {
  const int t = threadIdx.x;
  const int sing = wg_gj_blocked_core<R, NT, N, LDA>(A, sh);
  for (int j = t; j < N; j += NT) sh->pinv[j] = j;
  __syncthreads();
  for (int j = t; j < N; j += NT) {
    const int pj = sh->piv[j];
    if (pj >= 0 && pj < N) sh->pinv[pj] = j;
  }
  __syncthreads();
  return sing;
}

// In-place inverse of the N x N matrix A (LDS, leading dimension LDA) by the
// blocked Gauss-Jordan above (N <= 64) or its two-rows-per-lane form (N <= 128); store(i, k, v) receives the inverse in natural
// order (A^-1[pinv[i]][piv[k]] = Y[i][k] for the swept Y).  Returns nonzero if
// an exactly-zero pivot was met.
template <typename R, int NT, int N, int LDA, class FS>
__device__ __forceinline__ int wg_gj_inverse_blocked(cx<R>* A, FS store, GJBScratch<R, N>* sh) {
  const int t = threadIdx.x;
  int sing;
  if constexpr (N <= WAVE) sing = wg_gj_blocked_core<R, NT, N, LDA>(A, sh);
  else sing = wg_gj_blocked_core_tall<R, NT, N, LDA>(A, sh);
  for (int j = t; j < N; j += NT) sh->pinv[j] = j;
  __syncthreads();
  for (int j = t; j < N; j += NT) {
    const int pj = sh->piv[j];
    if (pj >= 0 && pj < N) sh->pinv[pj] = j;
  }
  __syncthreads();
  constexpr int PER = (N * N + NT - 1) / NT;
  cx<R> val[PER];
#pragma unroll
  for (int s2 = 0; s2 < PER; ++s2) {
    const int e = t + NT * s2;
    val[s2] = (e < N * N) ? A[(e / N) * LDA + (e % N)] : cx<R>{0, 0};
  }
  __syncthreads();
#pragma unroll
  for (int s2 = 0; s2 < PER; ++s2) {
    const int e = t + NT * s2;
    if (e < N * N) store(sh->pinv[e / N], sh->piv[e % N], val[s2]);
  }
  __syncthreads();
  return sing;
}

}  // namespace negf
