// Full-output RGF (retarded, lesser, greater), batched over the energies of one
// system, driven block by block from the host.
//
// Replaces recursive_gf_cal (dpnegf/negf/recursive_green_cal.py:173-296) when the
// caller needs every block (grd/grl/gru/gr_lc/gr_left and the lesser/greater
// blocks) rather than the transmission only.  The sweep order, the products and
// the reference's quirks are the same:
//  * lesser forward uses L_q gin_q U_q and the backward (B10/B11) uses L_q where
//    the textbook has U_q^dagger (:247-264);
//  * greater uses .conj() where the textbook has .mH (:269-289).
//
// Kernels: bgemm (C = alpha op(A) op(B) + beta C over a batch, operands of the
// form H - z_b S), binv (Gauss-Jordan inverse, one workgroup per matrix),
// bbuild (-(H - z_b S) - Sigma corners).
#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "rgf_kernels.hpp"
#include "rgf_full.hpp"

namespace negf {

// An operand with its op resolved into strides: element (i, k) of op(h - z s)
// sits at offset i * rs + k * cs of batch b, conjugated when csign < 0.
struct OpView {
  const cx<double>* h;
  const cx<double>* s;
  int64_t bstride, sstride, rs, cs;
  double csign;
};
__device__ __forceinline__ OpView op_view(const Opnd& o) {
  const bool tr = (o.op == OP_H || o.op == OP_T);
  return {o.h, o.s, o.bstride, o.sstride, tr ? 1 : (int64_t)o.ld, tr ? (int64_t)o.ld : 1,
          (o.op == OP_H || o.op == OP_C) ? -1.0 : 1.0};
}

// Batched C = alpha op(A) op(B) + beta C: a TM x TN = 32 FR x 32 FC output tile
// per workgroup (FR, FC in {1, 2}), wave w owning the FR x FC block of 16 x 16
// tiles at rows 16 FR (w / 2), columns 16 FC (w % 2).  The host picks FR = 1
// when m < 256 and the last 64-row tile would be at most half full (m mod 64
// in (0, 32]), and FC likewise for n, so below 256 the padding is at most 31
// rows or columns; from 256 on 64-wide tiles are kept, where a half-empty
// last tile is at most 11% of the work.  K is staged
// through LDS in chunks of KC = 16: the 256 threads load the A chunk (TM x 16,
// thread t at column k = t % 16 of rows t / 16 + 16 u) and the B chunk
// (16 x TN, row t / 16, columns t % 16 + 16 u), so sixteen consecutive lanes
// read sixteen consecutive elements.  The next chunk's raw h (and s) values
// are loaded into registers before this chunk's matrix-core steps and turned
// into operand values (h - z s, the conjugation, zeros outside the matrix)
// only when they are stored to LDS after those steps, so no arithmetic waits
// on the loads in between.  One LDS buffer (33 KB at 64 x 64) between two
// barriers per chunk, and registers held to three waves per SIMD: three
// workgroups share a CU (a second buffer, at two workgroups per CU, measured
// 5-8% slower at 96-256).  Indices are clamped, so every load
// is in bounds; a chunk past K contributes zeros.  A's LDS rows are padded to
// 17 elements so that the fragment reads (16 rows at one k) spread over the
// banks.  Matrix-core work is skipped, with wave-uniform control flow, for
// 16 x 16 fragment tiles outside the output and for k-steps past K.
template <bool SA, bool SB, int FR, int FC>
__global__ __launch_bounds__(256, 3) void bgemm_lds_kernel(BGemmArgs a) {
  static_assert((FR == 1 || FR == 2) && (FC == 1 || FC == 2), "fragments per wave");
  typedef mfma<double>::v4 v4;
  constexpr int TM = 32 * FR, TN = 32 * FC, UA = 2 * FR, UB = 2 * FC;
  constexpr int KC = 16, LDAS = KC + 1, LDBS = TN;
  __shared__ cx<double> As[TM * LDAS];
  __shared__ cx<double> Bs[KC * LDBS];
  const int tn = (a.n + TN - 1) / TN;
  const int ti = blockIdx.x / tn, tj = blockIdx.x % tn;
  const int b = blockIdx.y;
  if (a.skip && a.skip[b]) return;  // uniform per workgroup, before any barrier
  const int i0 = ti * TM, j0 = tj * TN;
  const cx<double> zb = a.z ? a.z[b] : cx<double>{0, 0};
  const OpView A = op_view(a.A), Bv = op_view(a.B_);
  const cx<double>* const Ah = A.h + (int64_t)b * A.bstride;
  const cx<double>* const Bh = Bv.h + (int64_t)b * Bv.bstride;
  const cx<double>* const Asp = SA ? A.s + (int64_t)b * A.sstride : nullptr;
  const cx<double>* const Bsp = SB ? Bv.s + (int64_t)b * Bv.sstride : nullptr;
  const int t = threadIdx.x;
  const int lane = t & (WAVE - 1), w = t / WAVE;
  const int li = lane & 15, lk = lane >> 4;
  const int K = a.k;
  // load assignment
  const int ak = t & 15, ar = t >> 4;  // A: column ak of rows ar + 16 u, u < UA
  const int bk = t >> 4, bc = t & 15;  // B: row bk, columns bc + 16 u, u < UB
  int64_t aoff[UA], boff[UB];
  bool arok[UA], bcok[UB];
#pragma unroll
  for (int u = 0; u < UA; ++u) {
    const int r = i0 + ar + 16 * u;
    arok[u] = r < a.m;
    aoff[u] = (int64_t)(arok[u] ? r : a.m - 1) * A.rs;
  }
#pragma unroll
  for (int u = 0; u < UB; ++u) {
    const int c = j0 + bc + 16 * u;
    bcok[u] = c < a.n;
    boff[u] = (int64_t)(bcok[u] ? c : a.n - 1) * Bv.cs;
  }
  cx<double> ra[UA], rb[UB], rsa[UA], rsb[UB];
  auto load_chunk = [&](int k0) __attribute__((always_inline)) {
    const int64_t ka = (int64_t)min(k0 + ak, K - 1) * A.cs;
    const int64_t kb = (int64_t)min(k0 + bk, K - 1) * Bv.rs;
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      ra[u] = Ah[aoff[u] + ka];
      if (SA) rsa[u] = Asp[aoff[u] + ka];
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      rb[u] = Bh[boff[u] + kb];
      if (SB) rsb[u] = Bsp[boff[u] + kb];
    }
  };
  auto store_chunk = [&](int k0) __attribute__((always_inline)) {
    const bool kaok = k0 + ak < K, kbok = k0 + bk < K;
#pragma unroll
    for (int u = 0; u < UA; ++u) {
      cx<double> va = ra[u];
      if (SA) va = va - zb * rsa[u];
      const bool oa = arok[u] && kaok;
      As[(ar + 16 * u) * LDAS + ak] = {oa ? va.re : 0.0, oa ? A.csign * va.im : 0.0};
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
      cx<double> vb = rb[u];
      if (SB) vb = vb - zb * rsb[u];
      const bool ob = bcok[u] && kbok;
      Bs[bk * LDBS + bc + 16 * u] = {ob ? vb.re : 0.0, ob ? Bv.csign * vb.im : 0.0};
    }
  };
  v4 cr[FR][FC], ci[FR][FC];
#pragma unroll
  for (int x = 0; x < FR; ++x)
#pragma unroll
    for (int y = 0; y < FC; ++y) cr[x][y] = ci[x][y] = v4{0, 0, 0, 0};
  const int rbase = 16 * FR * (w / 2), cbase = 16 * FC * (w % 2);
  // NK matrix-core k-steps of the chunk in LDS, for the RX x CY
  // fragment tiles of this wave that overlap the output (the others would only
  // multiply the zero padding).
  auto ksteps = [&](auto RXc, auto CYc, auto NKc) __attribute__((always_inline)) {
    constexpr int RX = decltype(RXc)::value, CY = decltype(CYc)::value;
    constexpr int NK = decltype(NKc)::value;
    static_assert(RX <= FR && CY <= FC, "fragments beyond the wave's block");
#pragma unroll
    for (int kk = 0; kk < NK; ++kk) {
      const int kl = 4 * kk + lk;
      cx<double> pa[FR], pb[FC];
#pragma unroll
      for (int x = 0; x < RX; ++x) pa[x] = As[(rbase + 16 * x + li) * LDAS + kl];
#pragma unroll
      for (int y = 0; y < CY; ++y) pb[y] = Bs[kl * LDBS + cbase + 16 * y + li];
#pragma unroll
      for (int x = 0; x < RX; ++x)
#pragma unroll
        for (int y = 0; y < CY; ++y) {
          cr[x][y] = mfma<double>::op(pa[x].re, pb[y].re, cr[x][y]);
          cr[x][y] = mfma<double>::op(-pa[x].im, pb[y].im, cr[x][y]);
          ci[x][y] = mfma<double>::op(pa[x].re, pb[y].im, ci[x][y]);
          ci[x][y] = mfma<double>::op(pa[x].im, pb[y].re, ci[x][y]);
        }
    }
  };
  // K = nfull KC + rem.  The loop body has no branches, so the accumulators
  // stay in one register class across iterations; its last iteration loads the
  // chunk after nfull - 1 (the remainder, or a chunk past K: clamped in bounds,
  // zeroed on store), which the remainder steps then read.
  const int nfull = K / KC, rem = K % KC;
  auto run = [&](auto RXc, auto CYc) __attribute__((always_inline)) {
    for (int ch = 0; ch < nfull; ++ch) {
      load_chunk((ch + 1) * KC);
      ksteps(RXc, CYc, std::integral_constant<int, KC / 4>());
      __syncthreads();
      store_chunk((ch + 1) * KC);
      __syncthreads();
    }
    switch ((rem + 3) / 4) {
      case 1: ksteps(RXc, CYc, std::integral_constant<int, 1>()); break;
      case 2: ksteps(RXc, CYc, std::integral_constant<int, 2>()); break;
      case 3: ksteps(RXc, CYc, std::integral_constant<int, 3>()); break;
      case 4: ksteps(RXc, CYc, std::integral_constant<int, 4>()); break;
      default: break;
    }
  };
  load_chunk(0);
  store_chunk(0);
  __syncthreads();
  // fragment tiles of this wave inside the output (at most FR x FC), wave-uniform
  const int wr = __builtin_amdgcn_readfirstlane(i0 + rbase);
  const int wc = __builtin_amdgcn_readfirstlane(j0 + cbase);
  const int RX = wr >= a.m ? 0 : (FR == 2 && wr + 16 < a.m ? 2 : 1);
  const int CY = wc >= a.n ? 0 : (FC == 2 && wc + 16 < a.n ? 2 : 1);
  typedef std::integral_constant<int, 0> I0;
  typedef std::integral_constant<int, 1> I1;
  typedef std::integral_constant<int, 2> I2;
  if (RX == 0 || CY == 0) {
    run(I0(), I0());  // loads and barriers only
  } else if (RX == 1 && CY == 1) {
    run(I1(), I1());
  } else if (RX == 1) {  // CY == 2
    if constexpr (FC == 2) run(I1(), I2());
  } else if (CY == 1) {  // RX == 2
    if constexpr (FR == 2) run(I2(), I1());
  } else {
    if constexpr (FR == 2 && FC == 2) run(I2(), I2());
  }
  cx<double>* C = a.C + (int64_t)b * a.cstride;
  const cx<double>* Cin = a.Cin + (int64_t)b * a.cstride;
  const cx<double> alpha = a.alpha, beta = a.beta;
  const bool use_beta = (beta.re != 0.0 || beta.im != 0.0);
#pragma unroll
  for (int x = 0; x < FR; ++x)
#pragma unroll
    for (int y = 0; y < FC; ++y) {
      const int ib = i0 + rbase + 16 * x;
      const int j = j0 + cbase + 16 * y + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = ib + mfma<double>::row(lk, r);
        if (i < a.m && j < a.n) {
          const int64_t o = (int64_t)i * a.ldc + j;
          cx<double> out = alpha * cx<double>{cr[x][y][r], ci[x][y][r]};
          if (use_beta) out = out + beta * Cin[o];
          C[o] = out;
        }
      }
    }
}

// MIN_WAVES: minimum waves per SIMD the register allocation must allow.
template <int NMAX, int NT, int MIN_WAVES = 1>
__global__ __launch_bounds__(NT, MIN_WAVES) void binv_kernel(cx<double>* A, int n, int64_t stride,
                                                  int32_t* status, const int32_t* skip) {
  constexpr int MAXR = gj_maxr(NMAX, NT);
  if (skip && skip[blockIdx.x]) return;  // uniform per workgroup, before any barrier
  __shared__ GJScratch<double, NMAX> gjs;
  cx<double>* M = A + (int64_t)blockIdx.x * stride;
  const int s = wg_gj_inverse<double, NT, NMAX, MAXR>(
      n, [&](int i, int k) { return M[i * n + k]; },
      [&](int i, int k, cx<double> v) { M[i * n + k] = v; }, &gjs);
  if (threadIdx.x == 0 && status && s) atomicOr(&status[blockIdx.x], STATUS_SINGULAR);
}

// Inverse of one n x n block per workgroup, 64 < n <= NP: the block goes to LDS
// padded to NP with an identity block (the padded rows and columns never
// pivot the block's columns, so the n x n corner of the inverse is A^-1), is
// inverted there by the blocked Gauss-Jordan with matrix-core rank-4
// updates, and its n x n corner is written back in natural order.
template <int NP, int NT>
__global__ __launch_bounds__(NT) void binv_lds_kernel(cx<double>* A, int n, int64_t stride,
                                                      int32_t* status, const int32_t* skip) {
  constexpr int LD = NP + 1;
  if (skip && skip[blockIdx.x]) return;  // uniform per workgroup, before any barrier
  __shared__ cx<double> M[NP * LD];
  __shared__ GJBScratch<double, NP> gjb;
  cx<double>* G = A + (int64_t)blockIdx.x * stride;
  for (int e = threadIdx.x; e < NP * NP; e += NT) {
    const int i = e / NP, k = e % NP;
    M[i * LD + k] = (i < n && k < n) ? G[(int64_t)i * n + k]
                                     : cx<double>{i == k ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  const int s = wg_gj_inverse_blocked<double, NT, NP, LD>(
      M, [&](int i, int k, cx<double> v) {
        if (i < n && k < n) G[(int64_t)i * n + k] = v;
      }, &gjb);
  if (threadIdx.x == 0 && status && s) atomicOr(&status[blockIdx.x], STATUS_SINGULAR);
}

// D[b] = -(H - z_b S) - Pb[b] - corner(SigmaL) - corner(SigmaR), Pb ([B, n, n],
// the energy-dependent probe self-energy seP) optional
__global__ void bbuild_kernel(cx<double>* D, const cx<double>* H, const cx<double>* S, int n,
                              const cx<double>* z, int B, const cx<double>* sL, int nL,
                              const cx<double>* sR, int nR, const cx<double>* Pb) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    const int e = (int)(x % ((int64_t)n * n));
    const int i = e / n, k = e % n;
    cx<double> v = H[e] - z[b] * S[e];
    if (Pb) v = v + Pb[x];
    if (sL && i < nL && k < nL) v = v + sL[(int64_t)b * nL * nL + i * nL + k];
    if (sR) {
      const int o = n - nR;
      if (i >= o && k >= o) v = v + sR[(int64_t)b * nR * nR + (i - o) * nR + (k - o)];
    }
    D[x] = -v;
  }
}

// D[b] = S[b]^H for S[b] of shape m x n (row-major), D[b] n x m
__global__ void bconjt_kernel(cx<double>* D, const cx<double>* S, int m, int n, int B) {
  const int64_t tot = (int64_t)B * m * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = x / ((int64_t)m * n);
    const int e = (int)(x % ((int64_t)m * n));
    const int i = e / m, k = e % m;  // D[i][k] = conj(S[k][i])
    D[x] = conj(S[b * m * n + (int64_t)k * n + i]);
  }
}


// D[b][i][k] = S[b][i][k] for a rows x cols sub-block (leading dimensions ldd / lds,
// batch strides sd / ss)
__global__ void bblock_copy_kernel(cx<double>* D, int ldd, int64_t sd, const cx<double>* S,
                                   int lds, int64_t ss, int rows, int cols, int B,
                                   const int32_t* skip) {
  const int64_t tot = (int64_t)B * rows * cols;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = x / ((int64_t)rows * cols);
    if (skip && skip[b]) continue;
    const int e = (int)(x % ((int64_t)rows * cols));
    const int i = e / cols, k = e % cols;
    D[b * sd + (int64_t)i * ldd + k] = S[b * ss + (int64_t)i * lds + k];
  }
}

// D[b] += I for B contiguous n x n matrices
__global__ void badd_identity_kernel(cx<double>* D, int n, int B, const int32_t* skip) {
  const int64_t tot = (int64_t)B * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = x / n;
    if (skip && skip[b]) continue;
    const int i = (int)(x % n);
    D[b * n * n + (int64_t)i * n + i].re += 1.0;
  }
}

// out[i][k] = sign conj(sum_b c_b in[b][k][i]) (c null: c_b = 1) for an m x n
// block: a gradient block summed over the batch, transposed and conjugated
// (with c = z and sign -1, the overlap block's gradient from H's).  A 256-thread workgroup
// takes RED_ELEMS consecutive elements; thread t sums element t % RED_ELEMS
// over the batch entries b = t / RED_ELEMS + RED_PHASES j, so one wave reads
// four 256-byte runs per step and a 64 x 64 block spreads over 256
// workgroups.  The partial sums of an element are added in a fixed order.
constexpr int RED_ELEMS = 16, RED_PHASES = 16;
__global__ __launch_bounds__(256) void breduce_conjt_kernel(cx<double>* out, const cx<double>* in,
                                                            int m, int n, int B, int accumulate,
                                                            const cx<double>* c, double sign) {
  __shared__ cx<double> part[RED_PHASES][RED_ELEMS];
  const int64_t mn = (int64_t)m * n;
  const int el = threadIdx.x % RED_ELEMS, ph = threadIdx.x / RED_ELEMS;
  const int64_t e = (int64_t)blockIdx.x * RED_ELEMS + el;  // in element (k, i) = (e / n, e % n)
  cx<double> acc{0, 0};
  if (e < mn)
    for (int b = ph; b < B; b += RED_PHASES)
      acc = acc + (c ? c[b] * in[(int64_t)b * mn + e] : in[(int64_t)b * mn + e]);
  part[ph][el] = acc;
  __syncthreads();
  if (threadIdx.x < RED_ELEMS && e < mn) {
    cx<double> sum = part[0][el];
    for (int p = 1; p < RED_PHASES; ++p) sum = sum + part[p][el];
    const int k = (int)(e / n), i = (int)(e % n);
    const int64_t o = (int64_t)i * m + k;
    const cx<double> r = {sign * sum.re, -sign * sum.im};
    out[o] = accumulate ? out[o] + r : r;
  }
}

// Gamma = i (Sigma - Sigma^H) padded into an n x n block (top-left or
// bottom-right corner), scaled per batch by scale[b] (null: 1).
__global__ void bgamma_pad_kernel(cx<double>* D, const cx<double>* sig, int ns, int n, int B,
                                  int at_end, const double* scale, double factor) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    const int e = (int)(x % ((int64_t)n * n));
    int i = e / n, k = e % n;
    cx<double> v{0, 0};
    const int o = at_end ? n - ns : 0;
    if (i >= o && k >= o && i < o + ns && k < o + ns) {
      v = gamma_of(sig + (int64_t)b * ns * ns, ns, i - o, k - o);
      const double f = factor * (scale ? scale[b] : 1.0);
      v.re *= f;
      v.im *= f;
    }
    D[x] = v;
  }
}

// Self-energy gradients, per energy: out[b][i][j] (+)= conj(a_b X[b][off + j][off + i])
// for i, j < m, X a batch of ld x ld blocks, a_b = c (2 w[b] when w is given).
__global__ void bcorner_conjt_kernel(cx<double>* out, const cx<double>* X, int ld, int off, int m,
                                     int B, cx<double> c, const double* w, int accumulate) {
  const int64_t tot = (int64_t)B * m * m;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)m * m));
    const int e = (int)(x % ((int64_t)m * m));
    const int i = e / m, j = e % m;
    const cx<double> v = X[(int64_t)b * ld * ld + (int64_t)(off + j) * ld + off + i];
    const double f = w ? 2.0 * w[b] : 1.0;
    const cx<double> a = {c.re * f, c.im * f};
    const cx<double> t = a * v;
    const cx<double> r = {t.re, -t.im};
    out[x] = accumulate ? out[x] + r : r;
  }
}

// out[e] (+)= sum_b Re(c_b * in[b][e]) over a batch of B matrices of `count` elements
__global__ void bweighted_re_sum_kernel(double* out, const cx<double>* in, const cx<double>* c,
                                        int64_t count, int B, int accumulate) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int b = 0; b < B; ++b) {
      const cx<double> v = in[(int64_t)b * count + e];
      acc += c[b].re * v.re - c[b].im * v.im;
    }
    out[e] = accumulate ? out[e] + acc : acc;
  }
}

// C = alpha op(A) op(B) + beta Cin over a batch; Cin (laid out as C) defaults to C.
static hipError_t gemm(hipStream_t st, int B, int m, int n, int k, Opnd A, Opnd Bo,
                       cx<double> alpha, cx<double> beta, cx<double>* C, int64_t cstride,
                       int ldc, const cx<double>* z, const cx<double>* Cin = nullptr,
                       const int32_t* skip = nullptr);
static Opnd mat(const cx<double>* h, int ld, int64_t bstride, int op, const cx<double>* s,
                int64_t sstride);

// out[b][i] (+)= sum_j A[b][i][j] * B[b][j][i]   (A: m x n, B: n x m; diagonal of A B)
__global__ void bdiag_product_kernel(cx<double>* out, const cx<double>* A, int64_t sa,
                                     const cx<double>* Bm, int64_t sb, int m, int n, int B,
                                     int accumulate) {
  const int64_t tot = (int64_t)B * m;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = x / m;
    const int i = (int)(x % m);
    const cx<double>* a = A + b * sa + (int64_t)i * n;
    const cx<double>* bb = Bm + b * sb + i;
    cx<double> acc{0, 0};
    for (int j = 0; j < n; ++j) acc = acc + a[j] * bb[(int64_t)j * m];
    out[x] = accumulate ? out[x] + acc : acc;
  }
}

hipError_t diag_product(cx<double>* out, const cx<double>* A, int64_t sa, const cx<double>* Bm,
                        int64_t sb, int m, int n, int B, int accumulate, hipStream_t st) {
  const int grid = (int)std::min<int64_t>(((int64_t)B * m + 255) / 256, 65535);
  hipLaunchKernelGGL(bdiag_product_kernel, dim3(grid), dim3(256), 0, st, out, A, sa, Bm, sb, m, n,
                     B, accumulate);
  return hipGetLastError();
}

hipError_t zgemm_batched(int B, int m, int n, int k, const cx<double>* A, int lda, int64_t sa,
                         int opA, const cx<double>* Bm, int ldb, int64_t sb, int opB,
                         cx<double> alpha, cx<double> beta, cx<double>* C, int ldc, int64_t sc,
                         hipStream_t st) {
  return gemm(st, B, m, n, k, mat(A, lda, sa, opA, nullptr, 0), mat(Bm, ldb, sb, opB, nullptr, 0), alpha,
              beta, C, sc, ldc, nullptr);
}

// T[b] = Re sum_{i,j} A[b][i][j] B[b][j][i]   (A: m x n, B: n x m), one workgroup per b
__global__ __launch_bounds__(256) void btrace_re_kernel(double* T, const cx<double>* A, int64_t sa,
                                                        const cx<double>* Bm, int64_t sb, int m,
                                                        int n) {
  __shared__ double red[8];
  const int b = blockIdx.x;
  const cx<double>* a = A + (int64_t)b * sa;
  const cx<double>* bb = Bm + (int64_t)b * sb;
  double acc = 0.0;
  for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    const cx<double> x = a[e], y = bb[(int64_t)j * m + i];
    acc += x.re * y.re - x.im * y.im;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) T[b] = acc;
}

hipError_t weighted_re_sum(double* out, const cx<double>* in, const cx<double>* c, int64_t count,
                           int B, int accumulate, hipStream_t st) {
  const int grid = (int)std::min<int64_t>((count + 255) / 256, 65535);
  hipLaunchKernelGGL(bweighted_re_sum_kernel, dim3(grid), dim3(256), 0, st, out, in, c, count, B,
                     accumulate);
  return hipGetLastError();
}


hipError_t batch_sum_conjt(cx<double>* out, const cx<double>* in, int m, int n, int B,
                           int accumulate, hipStream_t st) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(breduce_conjt_kernel, dim3(((int64_t)m * n + RED_ELEMS - 1) / RED_ELEMS),
                     dim3(256), 0, st, out, in, m, n, B, accumulate, (const cx<double>*)nullptr, 1.0);
  return hipGetLastError();
}

// Lopez-Sancho helpers for the batched (large-block) path -----------------------
// alpha0 = h10 - z s10, beta0 = h01 - z s01, eps = eps_s = H, per point
__global__ void ls_init_kernel(cx<double>* al, cx<double>* be, cx<double>* eps, cx<double>* epss,
                               const cx<double>* lead, int64_t lstride, const int32_t* sys,
                               const cx<double>* z, int n, int B, double eta) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    const int e = (int)(x % ((int64_t)n * n));
    const int i = e / n, k = e % n;
    const cx<double>* H = lead + (sys ? (int64_t)sys[b] * lstride : 0);
    const cx<double>* h01 = H + n * n;
    const cx<double>* s01 = H + 3 * n * n;
    const cx<double> zz{z[b].re, z[b].im + eta};
    be[x] = h01[e] - zz * s01[e];
    al[x] = conj(h01[k * n + i]) - zz * conj(s01[k * n + i]);
    eps[x] = H[e];
    epss[x] = H[e];
  }
}

// W = z S - X  (X = eps or eps_s), per point
__global__ void ls_zs_minus_kernel(cx<double>* W, const cx<double>* X, const cx<double>* lead,
                                   int64_t lstride, const int32_t* sys, const cx<double>* z,
                                   int n, int B, double eta, const int32_t* skip) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    if (skip && skip[b]) continue;
    const int e = (int)(x % ((int64_t)n * n));
    const cx<double>* S = lead + (sys ? (int64_t)sys[b] * lstride : 0) + 2 * n * n;
    const cx<double> zz{z[b].re, z[b].im + eta};
    W[x] = zz * S[e] - X[x];
  }
}

// per-point max_ij |alpha_ij| + |beta_ij|; records the first converged iteration
__global__ __launch_bounds__(256) void ls_conv_kernel(const cx<double>* al, const cx<double>* be,
                                                      int n, int32_t* iters, int32_t* done,
                                                      int32_t* remaining, int it, double stop) {
  __shared__ double red[8];
  const int b = blockIdx.x;
  if (done[b]) return;  // converged earlier: frozen (uniform per workgroup)
  const int64_t off = (int64_t)b * n * n;
  double m = 0.0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x)
    m = fmax(m, cabs(al[off + e]) + cabs(be[off + e]));
  m = block_max(m, red);
  if (threadIdx.x == 0) {
    if (m < stop) {
      done[b] = 1;
      iters[b] = it;
    } else {
      atomicAdd(remaining, 1);
    }
  }
}

// out[b] = sys[b] when it names one of the nsys lead records, else 0 with
// STATUS_BAD_INDEX or-ed into status[b]: every later kernel indexes records
// through out, so no read leaves the lead arena.
__global__ void ls_sys_check_kernel(int32_t* out, const int32_t* sys, int nsys, int32_t* status,
                                    int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int s = sys[b];
  const bool ok = s >= 0 && s < nsys;
  out[b] = ok ? s : 0;
  if (!ok && status) status[b] |= STATUS_BAD_INDEX;
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t count) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < count;
       x += (int64_t)gridDim.x * blockDim.x)
    p[x] = v;
}

// points still running after the iteration limit: flag and record the count
__global__ void ls_mark_kernel(const int32_t* done, int32_t* status, int32_t* iters, int B, int it) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < B && !done[b]) {
    if (status) status[b] |= STATUS_LS_NOCONV;
    iters[b] = it;
  }
}

// out = a - b elementwise
// a += t and b += t: the decimation's beta X_a, formed once, enters eps and
// eps_s (surface_green.py:196-200).
__global__ void badd2_kernel(cx<double>* a, cx<double>* b, const cx<double>* t, int64_t count,
                             int64_t per, const int32_t* skip) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < count;
       x += (int64_t)gridDim.x * blockDim.x) {
    if (skip && skip[x / per]) continue;
    const cx<double> v = t[x];
    a[x] = a[x] + v;
    b[x] = b[x] + v;
  }
}

__global__ void bsub_kernel(cx<double>* out, const cx<double>* a, const cx<double>* b, int64_t count) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < count;
       x += (int64_t)gridDim.x * blockDim.x)
    out[x] = a[x] - b[x];
}

// residual test of surface_green.py:211-220 from R = test g_s - I (per point)
__global__ __launch_bounds__(256) void ls_resid_kernel(const cx<double>* Rm, int n, double* resid,
                                                       int32_t* status) {
  __shared__ double red[8];
  const int b = blockIdx.x;
  const int64_t off = (int64_t)b * n * n;
  double m = 0.0;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    cx<double> v = Rm[off + e];
    if (e / n == e % n) v.re -= 1.0;
    m = fmax(m, cabs(v));
  }
  m = block_max(m, red);
  if (threadIdx.x == 0) {
    if (resid) resid[b] = m;
    if (status) {
      if (m >= 3.0e-5) atomicOr(&status[b], STATUS_LS_RESID_FAIL);
      else if (m > 1.0e-8) atomicOr(&status[b], STATUS_LS_RESID_WARN);
    }
  }
}

// test = z S - H - X2 (per point)
__global__ void ls_test_kernel(cx<double>* T, const cx<double>* X2, const cx<double>* lead,
                               int64_t lstride, const int32_t* sys, const cx<double>* z, int n,
                               int B, double eta) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    const int e = (int)(x % ((int64_t)n * n));
    const cx<double>* H = lead + (sys ? (int64_t)sys[b] * lstride : 0);
    const cx<double> zz{z[b].re, z[b].im + eta};
    T[x] = zz * H[2 * n * n + e] - H[e] - X2[x];
  }
}

// coupling operands per point: out = z*s - h (op N) or z*conj(s)^T - conj(h)^T (op H, z not
// conjugated), with h, s of shape r x c from the lead record at element offset `off`
__global__ void ls_coupling_kernel(cx<double>* out, const cx<double>* lead, int64_t lstride,
                                   const int32_t* sys, const cx<double>* z, double eta,
                                   int64_t hoff, int64_t soff, int r, int c, int transpose, int B) {
  const int64_t tot = (int64_t)B * r * c;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)r * c));
    const int e = (int)(x % ((int64_t)r * c));
    const cx<double>* base = lead + (sys ? (int64_t)sys[b] * lstride : 0);
    const cx<double> zz{z[b].re, z[b].im + eta};
    if (!transpose) {
      out[x] = zz * base[soff + e] - base[hoff + e];
    } else {  // element (i, k) of the c x r result = z conj(s[k][i]) - conj(h[k][i])
      const int i = e / r, k = e % r;
      out[x] = zz * conj(base[soff + (int64_t)k * c + i]) - conj(base[hoff + (int64_t)k * c + i]);
    }
  }
}

// ---------------------------------------------------------------------------------
// host side

// Tuning table.  Lookups take the mutex only while the table holds an entry or
// recording is on; otherwise a driver pays two relaxed atomic loads.
namespace {
typedef std::array<int, 4> TuneKey;
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
std::vector<TuneKey> g_tune_seen;
std::atomic<int> g_tune_entries{0};
std::atomic<bool> g_tune_recording{false};

int tune_lookup(int kind, int m, int n, int k) {
  if (g_tune_entries.load(std::memory_order_relaxed) == 0 &&
      !g_tune_recording.load(std::memory_order_relaxed))
    return 0;
  const TuneKey key{kind, m, n, k};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_tune_recording.load(std::memory_order_relaxed) &&
      std::find(g_tune_seen.begin(), g_tune_seen.end(), key) == g_tune_seen.end())
    g_tune_seen.push_back(key);
  const auto it = g_tune.find(key);
  return it == g_tune.end() ? 0 : it->second;
}
}  // namespace

// Built-in choices: 32-wide tiles in a dimension below 256 whose last 64-wide
// tile would be at most half full (there the padding is a large share of the
// work; from 256 on it is at most 11%, and 64-wide tiles run faster than
// that); the LDS blocked inverse for 65..96 (the scalar register-resident
// sweep is latency-bound at n pivot steps with two workgroup barriers each)
static int builtin_tile(int m, int n) {
  const bool fr1 = m < 256 && m % 64 != 0 && m % 64 <= 32;
  const bool fc1 = n < 256 && n % 64 != 0 && n % 64 <= 32;
  return fr1 ? (fc1 ? TILE_32x32 : TILE_32x64) : (fc1 ? TILE_64x32 : TILE_64x64);
}

// The blocked Gauss-Jordan in LDS (binv_lds_kernel, padded to 32, 48, 64 or 96)
// from 17 orbitals on; measured on 4096 blocks: 0.15 / 0.097 ms at n = 20,
// 0.87 / 0.25 ms at 48, 1.37 / 0.57 ms at 64, 4.33 / 2.58 ms at 96 (register /
// LDS form), level at 72.  From 97 on the Schur-complement form: its halves
// take the LDS form and its products run on the matrix cores, where the
// 1024-thread register sweep reaches about 4 TF/s (the large-block sweep with
// 120-orbital blocks: 617 -> 564 ms).
static int builtin_inverse(int n) {
  return n <= 16 ? INV_REGISTER : n <= 96 ? INV_LDS : INV_SCHUR;
}

int tune_builtin(int kind, int m, int n, int k) {
  if (kind == TUNE_GEMM_TILE && m >= 1 && n >= 1 && k >= 1) return builtin_tile(m, n);
  if (kind == TUNE_INVERSE && m >= 1 && m <= GJ_NMAX_LARGE && n == 0 && k == 0)
    return builtin_inverse(m);
  return 0;
}

int tune_set(int kind, int m, int n, int k, int choice) {
  if (kind == TUNE_GEMM_TILE) {
    if (m < 1 || n < 1 || k < 1 || choice < 0 || choice > TILE_64x64) return -1;
  } else if (kind == TUNE_INVERSE) {
    if (n != 0 || k != 0 || m < 1 || m > GJ_NMAX_LARGE || choice < 0 || choice > INV_SCHUR) return -1;
    if (choice == INV_LDS && m > 96) return -1;
    if (choice == INV_SCHUR && m < SCHUR_NMIN) return -1;
  } else {
    return -1;
  }
  const TuneKey key{kind, m, n, k};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (choice == 0) g_tune.erase(key);
  else g_tune[key] = choice;
  g_tune_entries.store((int)g_tune.size(), std::memory_order_relaxed);
  return 0;
}

int tune_get(int kind, int m, int n, int k) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const auto it = g_tune.find(TuneKey{kind, m, n, k});
  return it == g_tune.end() ? 0 : it->second;
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  g_tune_entries.store(0, std::memory_order_relaxed);
}

void tune_record(bool on) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (on) g_tune_seen.clear();
  g_tune_recording.store(on, std::memory_order_relaxed);
}

int tune_recorded(int32_t* out, int cap) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  const int cnt = (int)g_tune_seen.size();
  for (int i = 0; i < cnt && i < cap; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = g_tune_seen[i][j];
  return cnt;
}

static hipError_t gemm(hipStream_t st, int B, int m, int n, int k, Opnd A, Opnd Bo,
                       cx<double> alpha, cx<double> beta, cx<double>* C, int64_t cstride,
                       int ldc, const cx<double>* z, const cx<double>* Cin,
                       const int32_t* skip) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (k <= 0) return hipErrorInvalidValue;  // the kernel clamps indices into [0, k)
  BGemmArgs a;
  a.m = m; a.n = n; a.k = k; a.B = B;
  a.A = A; a.B_ = Bo; a.alpha = alpha; a.beta = beta;
  a.C = C; a.Cin = Cin ? Cin : C; a.cstride = cstride; a.ldc = ldc; a.z = z; a.skip = skip;
  // the tuning table's shape for (m, n, k), else the built-in one
  int c = tune_lookup(TUNE_GEMM_TILE, m, n, k);
  if (c == 0) c = builtin_tile(m, n);
  const bool fr1 = c == TILE_32x32 || c == TILE_32x64;
  const bool fc1 = c == TILE_32x32 || c == TILE_64x32;
  const int tm = fr1 ? 32 : 64, tn = fc1 ? 32 : 64;
  const dim3 grid(((m + tm - 1) / tm) * ((n + tn - 1) / tn), B);
  const bool sa = A.s != nullptr, sb = Bo.s != nullptr;
  auto launch = [&](auto FRc, auto FCc) {
    constexpr int FR = decltype(FRc)::value, FC = decltype(FCc)::value;
    if (sa && sb) hipLaunchKernelGGL((bgemm_lds_kernel<true, true, FR, FC>), grid, dim3(256), 0, st, a);
    else if (sa) hipLaunchKernelGGL((bgemm_lds_kernel<true, false, FR, FC>), grid, dim3(256), 0, st, a);
    else if (sb) hipLaunchKernelGGL((bgemm_lds_kernel<false, true, FR, FC>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((bgemm_lds_kernel<false, false, FR, FC>), grid, dim3(256), 0, st, a);
  };
  typedef std::integral_constant<int, 1> I1;
  typedef std::integral_constant<int, 2> I2;
  if (fr1 && fc1) launch(I1(), I1());
  else if (fr1) launch(I1(), I2());
  else if (fc1) launch(I2(), I1());
  else launch(I2(), I2());
  return hipGetLastError();
}

#define CHK(x)                                  \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return e_;            \
  } while (0)

int64_t inverse_workspace(int n, int B) {
  return n >= SCHUR_NMIN ? 3 * (int64_t)B * n * n : 0;
}

// The form inverse() takes for n: the tuning table's, else the built-in one;
// the Schur form needs the caller's workspace (without one, n <= 128 falls
// back to the register sweep).  Above GJ_NMAX_LARGE always the Schur form.
static int inverse_form(int n, bool has_ws) {
  if (n > GJ_NMAX_LARGE) return INV_SCHUR;
  int form = tune_lookup(TUNE_INVERSE, n, 0, 0);
  if (form == 0 || (form == INV_LDS && n > 96) || (form == INV_SCHUR && n < SCHUR_NMIN))
    form = builtin_inverse(n);
  if (form == INV_SCHUR && !has_ws) form = INV_REGISTER;
  return form;
}

static hipError_t block_copy(hipStream_t st, cx<double>* D, int ldd, int64_t sd,
                             const cx<double>* S, int lds, int64_t ss, int rows, int cols, int B,
                             const int32_t* skip = nullptr) {
  const int64_t tot = (int64_t)B * rows * cols;
  hipLaunchKernelGGL(bblock_copy_kernel, dim3((int)std::min<int64_t>((tot + 255) / 256, 65535)),
                     dim3(256), 0, st, D, ldd, sd, S, lds, ss, rows, cols, B, skip);
  return hipGetLastError();
}

// In-place inverse of B contiguous n x n matrices (form: inverse_form).  The
// register and LDS forms: one workgroup per matrix, Gauss-Jordan with partial
// pivoting.  The Schur form (up to BLOCK_NMAX): two diagonal halves h1 = n/2,
// h2 = n - h1, each inverted by this same function (recursively, until the
// halves are at most GJ_NMAX_LARGE),
// joined through the Schur complement S = A22 - A21 A11^-1 A12:
//   A^-1 = [[X + X A12 Y A21 X, -X A12 Y], [-Y A21 X, Y]],  X = A11^-1, Y = S^-1.
// No pivoting across the halves: every matrix inverted on this path is
// z S - H - Sigma (or a decimation W = z S - eps) with eta > 0, whose
// anti-Hermitian part is positive definite, so A11 and S are nonsingular --
// but they can be far worse conditioned than A, so the assembled inverse takes
// one Newton step in FP64, X <- X + X (I - A X), which squares its relative
// error.  A_copy: an existing copy of A (same layout) for that step's residual,
// which saves copying A into the workspace.  ws: inverse_workspace(n, B) elements.
// Halves up to GJ_NMAX_LARGE take a one-workgroup Gauss-Jordan form (a Schur
// form nested inside this one measured 1.7e-10 against the 1e-10 parity bar
// on 256/192/240 blocks).  A half above GJ_NMAX_LARGE takes the region Xc
// below as its own workspace:
// Xc is not used until the Newton step, and it holds B n^2 elements, at least
// the 3 B ceil(n/2)^2 a half needs for n >= 7.
static hipError_t inverse(hipStream_t st, cx<double>* A, int n, int B, int32_t* status,
                          cx<double>* ws, const cx<double>* A_copy = nullptr,
                          const int32_t* skip = nullptr) {
  const int64_t s = (int64_t)n * n;
  const int form = inverse_form(n, ws != nullptr);
  if (form == INV_LDS) {
    if (n <= 32)
      hipLaunchKernelGGL((binv_lds_kernel<32, 128>), dim3(B), dim3(128), 0, st, A, n, s, status, skip);
    else if (n <= 48)
      hipLaunchKernelGGL((binv_lds_kernel<48, 192>), dim3(B), dim3(192), 0, st, A, n, s, status, skip);
    else if (n <= 64)
      hipLaunchKernelGGL((binv_lds_kernel<64, 256>), dim3(B), dim3(256), 0, st, A, n, s, status, skip);
    else
      hipLaunchKernelGGL((binv_lds_kernel<96, 512>), dim3(B), dim3(512), 0, st, A, n, s, status, skip);
  }
  else if (n <= 16) hipLaunchKernelGGL((binv_kernel<16, 256>), dim3(B), dim3(256), 0, st, A, n, s, status, skip);
  else if (n <= 32) hipLaunchKernelGGL((binv_kernel<32, 256>), dim3(B), dim3(256), 0, st, A, n, s, status, skip);
  else if (n <= 64)  // 8 rows per thread and 71 registers: three workgroups per CU
    hipLaunchKernelGGL((binv_kernel<64, 512, 4>), dim3(B), dim3(512), 0, st, A, n, s, status, skip);
  else if (n <= 96)  // registers capped at 128: two workgroups per CU
    hipLaunchKernelGGL((binv_kernel<96, 512, 4>), dim3(B), dim3(512), 0, st, A, n, s, status, skip);
  else if (form == INV_REGISTER)
    hipLaunchKernelGGL((binv_kernel<GJ_NMAX_LARGE, 1024>), dim3(B), dim3(1024), 0, st, A, n, s, status, skip);
  else if (n <= BLOCK_NMAX) {
    if (!ws) return hipErrorInvalidValue;
    const int h1 = n / 2, h2 = n - h1;
    const int64_t s1 = (int64_t)h1 * h1, s2 = (int64_t)h2 * h2, s12 = (int64_t)h1 * h2;
    cx<double>* A0w = ws;            // A as given             [n x n]
    const cx<double>* A0 = A0w;
    cx<double>* Xc = A0w + B * s;    // copy of the inverse    [n x n]
    cx<double>* X = Xc + B * s;      // A11^-1   [h1 x h1]  (then I - A X, [n x n])
    cx<double>* Y = X + B * s1;      // S^-1     [h2 x h2]
    cx<double>* T1 = Y + B * s2;     // X A12    [h1 x h2]
    cx<double>* T2 = T1 + B * s12;   // A21 X    [h2 x h1]
    const cx<double> one{1, 0}, mone{-1, 0}, zero{0, 0};
    auto M = [](const cx<double>* p, int ld, int64_t bs) {
      return mat(p, ld, bs, OP_N, nullptr, 0);
    };
    cx<double>* A12 = A + h1;
    cx<double>* A21 = A + (int64_t)h1 * n;
    cx<double>* A22 = A21 + h1;
    if (A_copy) A0 = A_copy;
    else CHK(block_copy(st, A0w, n, s, A, n, s, n, n, B, skip));
    CHK(block_copy(st, X, h1, s1, A, n, s, h1, h1, B, skip));
    CHK(inverse(st, X, h1, B, status, h1 > GJ_NMAX_LARGE ? Xc : nullptr, nullptr, skip));
    CHK(gemm(st, B, h1, h2, h1, M(X, h1, s1), M(A12, n, s), one, zero, T1, s12, h2, nullptr, nullptr, skip));
    CHK(block_copy(st, Y, h2, s2, A22, n, s, h2, h2, B, skip));
    CHK(gemm(st, B, h2, h2, h1, M(A21, n, s), M(T1, h2, s12), mone, one, Y, s2, h2, nullptr, nullptr, skip));
    CHK(inverse(st, Y, h2, B, status, h2 > GJ_NMAX_LARGE ? Xc : nullptr, nullptr, skip));
    CHK(gemm(st, B, h2, h1, h1, M(A21, n, s), M(X, h1, s1), one, zero, T2, s12, h1, nullptr, nullptr, skip));
    CHK(gemm(st, B, h1, h2, h2, M(T1, h2, s12), M(Y, h2, s2), mone, zero, A12, s, n, nullptr, nullptr, skip));
    CHK(gemm(st, B, h2, h1, h2, M(Y, h2, s2), M(T2, h1, s12), mone, zero, A21, s, n, nullptr, nullptr, skip));
    CHK(block_copy(st, A, n, s, X, h1, s1, h1, h1, B, skip));
    CHK(gemm(st, B, h1, h1, h2, M(A12, n, s), M(T2, h1, s12), mone, one, A, s, n, nullptr, nullptr, skip));
    CHK(block_copy(st, A22, n, s, Y, h2, s2, h2, h2, B, skip));
    // Newton step: R = I - A0 A^-1 (in X's space), A^-1 += copy(A^-1) R
    cx<double>* Rm = X;
    CHK(gemm(st, B, n, n, n, M(A0, n, s), M(A, n, s), mone, zero, Rm, s, n, nullptr, nullptr, skip));
    const int gid = (int)std::min<int64_t>((B * (int64_t)n + 255) / 256, 65535);
    hipLaunchKernelGGL(badd_identity_kernel, dim3(gid), dim3(256), 0, st, Rm, n, B, skip);
    CHK(hipGetLastError());
    CHK(block_copy(st, Xc, n, s, A, n, s, n, n, B, skip));
    CHK(gemm(st, B, n, n, n, M(Xc, n, s), M(Rm, n, s), one, one, A, s, n, nullptr, nullptr, skip));
    return hipSuccess;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t zinv_batched(hipStream_t st, cx<double>* A, int n, int B, int32_t* status) {
  if (n < 1 || n > GJ_NMAX_LARGE || B < 0) return hipErrorInvalidValue;
  if (B == 0) return hipSuccess;
  return inverse(st, A, n, B, status, nullptr);
}

static Opnd mat(const cx<double>* h, int ld, int64_t bstride, int op = OP_N,
                const cx<double>* s = nullptr, int64_t sstride = 0);
static Opnd mat(const cx<double>* h, int ld, int64_t bstride, int op, const cx<double>* s,
                int64_t sstride) {
  Opnd o;
  o.h = h; o.s = s; o.bstride = bstride; o.sstride = sstride; o.ld = ld; o.op = op;
  return o;
}

hipError_t rgf_full(const RgfFullArgs& a, hipStream_t st) {
  const int K = a.K, B = a.B;
  const std::vector<int> n(a.bsz, a.bsz + K);
  const cx<double> one{1, 0}, mone{-1, 0}, zero{0, 0};
  // offsets into the per-block arenas
  std::vector<int64_t> od(K), ou(K > 1 ? K - 1 : 1), olc(K), hoff(K);
  int64_t sd = 0, su = 0, slc = 0, sh = 0;
  for (int q = 0; q < K; ++q) {
    od[q] = sd; sd += (int64_t)B * n[q] * n[q];
    olc[q] = slc; slc += (int64_t)B * n[q] * n[K - 1];
    hoff[q] = sh; sh += 2 * (int64_t)n[q] * n[q];
    if (q < K - 1) { ou[q] = su; su += (int64_t)B * n[q] * n[q + 1]; sh += 4 * (int64_t)n[q] * n[q + 1]; }
  }
  const cx<double>* H = a.arena;
  auto hd = [&](int q) { return H + hoff[q]; };
  auto sdp = [&](int q) { return H + hoff[q] + (int64_t)n[q] * n[q]; };
  auto hu = [&](int q) { return H + hoff[q] + 2 * (int64_t)n[q] * n[q]; };
  auto suu = [&](int q) { return hu(q) + (int64_t)n[q] * n[q + 1]; };
  auto hl = [&](int q) { return suu(q) + (int64_t)n[q] * n[q + 1]; };
  auto sl = [&](int q) { return hl(q) + (int64_t)n[q + 1] * n[q]; };
  // U_q - z SU_q and L_q - z SL_q as (shared) operands
  auto Uop = [&](int q, int op = OP_N) { return mat(hu(q), n[q + 1], 0, op, a.ortho ? nullptr : suu(q), 0); };
  auto Lop = [&](int q, int op = OP_N) { return mat(hl(q), n[q], 0, op, a.ortho ? nullptr : sl(q), 0); };
  const cx<double>* z = a.z;
  const int64_t W = (int64_t)B * a.nmax * a.nmax;
  cx<double>* w0 = a.work;
  cx<double>* w1 = a.work + W;
  cx<double>* w2 = a.work + 2 * W;
  cx<double>* w3 = a.work + 3 * W;
  cx<double>* w4 = a.work + 4 * W;
  auto gl = [&](int q) { return a.gl + od[q]; };
  auto grd = [&](int q) { return a.grd + od[q]; };
  auto grl = [&](int q) { return a.grl + ou[q]; };
  auto gru = [&](int q) { return a.gru + ou[q]; };
  const int bthreads = 256;
  auto build = [&](int q, bool left, bool right) {
    const int64_t tot = (int64_t)B * n[q] * n[q];
    const int grid = (int)std::min<int64_t>((tot + bthreads - 1) / bthreads, 65535);
    hipLaunchKernelGGL(bbuild_kernel, dim3(grid), dim3(bthreads), 0, st, gl(q), hd(q), sdp(q), n[q],
                       z, B, left ? a.sigL : nullptr, a.nL, right ? a.sigR : nullptr, a.nR,
                       a.seP ? a.seP + od[q] : nullptr);
    return hipGetLastError();
  };

  // ---- forward: left-connected g_q  (:198-207)
  CHK(build(0, true, K == 1));
  CHK(inverse(st, gl(0), n[0], B, a.status, a.inv_ws));
  for (int q = 0; q + 1 < K; ++q) {
    const int nq = n[q], n1 = n[q + 1];
    CHK(gemm(st, B, nq, n1, nq, mat(gl(q), nq, (int64_t)nq * nq), Uop(q), one, zero, w0,
             (int64_t)nq * n1, n1, z));
    CHK(build(q + 1, false, q + 1 == K - 1));
    CHK(gemm(st, B, n1, n1, nq, Lop(q), mat(w0, n1, (int64_t)nq * n1), mone, one, gl(q + 1),
             (int64_t)n1 * n1, n1, z));
    CHK(inverse(st, gl(q + 1), n1, B, a.status, a.inv_ws));
  }

  // ---- backward: retarded blocks and the last column  (:209-238)
  const int nK = n[K - 1];
  CHK(hipMemcpyAsync(grd(K - 1), gl(K - 1), sizeof(cx<double>) * B * nK * nK,
                     hipMemcpyDeviceToDevice, st));
  cx<double>* col = w3;  // G_{q,K-1}, ping-pong between w3 and w4 unless lc is kept
  cx<double>* colnext = w4;
  if (a.lc) {
    col = a.lc + olc[K - 1];
    CHK(hipMemcpyAsync(col, gl(K - 1), sizeof(cx<double>) * B * nK * nK, hipMemcpyDeviceToDevice, st));
  } else {
    CHK(hipMemcpyAsync(col, gl(K - 1), sizeof(cx<double>) * B * nK * nK, hipMemcpyDeviceToDevice, st));
  }
  for (int q = K - 2; q >= 0; --q) {
    const int nq = n[q], n1 = n[q + 1];
    // gU = g_q U_q
    CHK(gemm(st, B, nq, n1, nq, mat(gl(q), nq, (int64_t)nq * nq), Uop(q), one, zero, w0,
             (int64_t)nq * n1, n1, z));
    // T = G_{q+1,q+1} L_q ; grl = T g_q
    CHK(gemm(st, B, n1, nq, n1, mat(grd(q + 1), n1, (int64_t)n1 * n1), Lop(q), one, zero, w1,
             (int64_t)n1 * nq, nq, z));
    CHK(gemm(st, B, n1, nq, nq, mat(w1, nq, (int64_t)n1 * nq), mat(gl(q), nq, (int64_t)nq * nq),
             one, zero, grl(q), (int64_t)n1 * nq, nq, z));
    // gru = gU G_{q+1,q+1}
    CHK(gemm(st, B, nq, n1, n1, mat(w0, n1, (int64_t)nq * n1), mat(grd(q + 1), n1, (int64_t)n1 * n1),
             one, zero, gru(q), (int64_t)nq * n1, n1, z));
    // grd = g_q + gU grl
    CHK(hipMemcpyAsync(grd(q), gl(q), sizeof(cx<double>) * B * nq * nq, hipMemcpyDeviceToDevice, st));
    CHK(gemm(st, B, nq, nq, n1, mat(w0, n1, (int64_t)nq * n1), mat(grl(q), nq, (int64_t)n1 * nq),
             one, one, grd(q), (int64_t)nq * nq, nq, z));
    // G_{q,K-1} = gU G_{q+1,K-1}
    cx<double>* dst = a.lc ? a.lc + olc[q] : colnext;
    CHK(gemm(st, B, nq, nK, n1, mat(w0, n1, (int64_t)nq * n1), mat(col, nK, (int64_t)n1 * nK), one,
             zero, dst, (int64_t)nq * nK, nK, z));
    if (!a.lc) { colnext = col; }
    col = dst;
  }
  CHK(hipMemcpyAsync(a.g_trans, col, sizeof(cx<double>) * B * n[0] * nK, hipMemcpyDeviceToDevice, st));

  // ---- lesser (:241-264), greater (:267-289) and the bilinear G S G (pass 2,
  // no counterpart in the reference: the density adjoint's G W G, with
  // x_{q+1} = g (S + L x_q U) g forward and, backward,
  //   X_{q+1,q} = G_{q+1,q+1} L x_q + X_{q+1,q+1} L g_q
  //   X_{q,q+1} = x_q U G_{q+1,q+1} + g_q U X_{q+1,q+1}
  //   X_qq = x_q + G_{q,q+1} L x_q + x_q U G_{q+1,q} + g_q U X_{q+1,q+1} L g_q)
  for (int pass = 0; pass < 3; ++pass) {
    const cx<double>* sin = pass == 0 ? a.s_in : pass == 1 ? a.s_out : a.s_bil;
    cx<double>* gin = pass == 0 ? a.gin : pass == 1 ? a.gip : a.gbi;
    cx<double>* gnd = pass == 0 ? a.gnd : pass == 1 ? a.gpd : a.gbd;
    cx<double>* gnl = pass == 0 ? a.gnl : pass == 1 ? a.gpl : a.gbl;
    cx<double>* gnu = pass == 0 ? a.gnu : pass == 1 ? a.gpu : a.gbu;
    if (!sin) continue;
    // lesser conjugates with ^H, the greater quirk with plain conj, bilinear not at all
    const int cj = pass == 0 ? OP_H : pass == 1 ? OP_C : OP_N;
    auto gi = [&](int q) { return gin + od[q]; };
    auto si = [&](int q) { return sin + od[q]; };
    // gin_0 = g_0 s_0 g_0^(cj)
    CHK(gemm(st, B, n[0], n[0], n[0], mat(gl(0), n[0], (int64_t)n[0] * n[0]),
             mat(si(0), n[0], (int64_t)n[0] * n[0]), one, zero, w0, (int64_t)n[0] * n[0], n[0], z));
    CHK(gemm(st, B, n[0], n[0], n[0], mat(w0, n[0], (int64_t)n[0] * n[0]),
             mat(gl(0), n[0], (int64_t)n[0] * n[0], cj), one, zero, gi(0), (int64_t)n[0] * n[0], n[0], z));
    for (int q = 0; q + 1 < K; ++q) {
      const int nq = n[q], n1 = n[q + 1];
      // src = s_{q+1} + L gin_q U^(cj? conj : N)
      CHK(gemm(st, B, n1, nq, nq, Lop(q), mat(gi(q), nq, (int64_t)nq * nq), one, zero, w0,
               (int64_t)n1 * nq, nq, z));
      CHK(hipMemcpyAsync(w1, si(q + 1), sizeof(cx<double>) * B * n1 * n1, hipMemcpyDeviceToDevice, st));
      CHK(gemm(st, B, n1, n1, nq, mat(w0, nq, (int64_t)n1 * nq), Uop(q, pass == 1 ? OP_C : OP_N), one,
               one, w1, (int64_t)n1 * n1, n1, z));
      // gin_{q+1} = g src g^(cj)
      CHK(gemm(st, B, n1, n1, n1, mat(gl(q + 1), n1, (int64_t)n1 * n1), mat(w1, n1, (int64_t)n1 * n1),
               one, zero, w0, (int64_t)n1 * n1, n1, z));
      CHK(gemm(st, B, n1, n1, n1, mat(w0, n1, (int64_t)n1 * n1), mat(gl(q + 1), n1, (int64_t)n1 * n1, cj),
               one, zero, gi(q + 1), (int64_t)n1 * n1, n1, z));
    }
    CHK(hipMemcpyAsync(gnd + od[K - 1], gi(K - 1), sizeof(cx<double>) * B * nK * nK,
                       hipMemcpyDeviceToDevice, st));
    for (int q = K - 2; q >= 0; --q) {
      const int nq = n[q], n1 = n[q + 1];
      const int64_t sq = (int64_t)nq * nq, s1 = (int64_t)n1 * n1, sx = (int64_t)nq * n1;
      cx<double>* nl = gnl + ou[q];
      cx<double>* nd = gnd + od[q];
      const cx<double>* nd1 = gnd + od[q + 1];
      // lg = L^(c) g_q^(cj)  [n1 x nq]  (lesser: L g^H ; greater: conj(L) conj(g))
      CHK(gemm(st, B, n1, nq, nq, Lop(q, pass == 1 ? OP_C : OP_N), mat(gl(q), nq, sq, cj), one, zero,
               w0, (int64_t)n1 * nq, nq, z));
      // gnl = (G_{q+1,q+1} L) gin_q + gnd_{q+1} lg
      CHK(gemm(st, B, n1, nq, n1, mat(grd(q + 1), n1, s1), Lop(q), one, zero, w1, (int64_t)n1 * nq,
               nq, z));
      CHK(gemm(st, B, n1, nq, nq, mat(w1, nq, (int64_t)n1 * nq), mat(gi(q), nq, sq), one, zero, nl,
               (int64_t)n1 * nq, nq, z));
      CHK(gemm(st, B, n1, nq, n1, mat(nd1, n1, s1), mat(w0, nq, (int64_t)n1 * nq), one, one, nl,
               (int64_t)n1 * nq, nq, z));
      // gnd = gin_q + (g_q U) gnd_{q+1} lg + gin_q U^(c) (gru or grl)^(cj) + gru L gin_q
      CHK(hipMemcpyAsync(nd, gi(q), sizeof(cx<double>) * B * nq * nq, hipMemcpyDeviceToDevice, st));
      CHK(gemm(st, B, nq, n1, nq, mat(gl(q), nq, sq), Uop(q), one, zero, w1, sx, n1, z));
      CHK(gemm(st, B, nq, n1, n1, mat(w1, n1, sx), mat(nd1, n1, s1), one, zero, w2, sx, n1, z));
      CHK(gemm(st, B, nq, nq, n1, mat(w2, n1, sx), mat(w0, nq, (int64_t)n1 * nq), one, one, nd, sq, nq, z));
      CHK(gemm(st, B, nq, n1, nq, mat(gi(q), nq, sq), Uop(q, pass == 1 ? OP_C : OP_N), one, zero, w1,
               sx, n1, z));
      if (pass == 0)  // ... gru[q]^H  (n1 x nq operand from gru stored nq x n1)
        CHK(gemm(st, B, nq, nq, n1, mat(w1, n1, sx), mat(gru(q), n1, sx, OP_H), one, one, nd, sq, nq, z));
      else            // ... conj(grl[q]) (greater) or grl[q] (bilinear)  (grl stored n1 x nq)
        CHK(gemm(st, B, nq, nq, n1, mat(w1, n1, sx),
                 mat(grl(q), nq, (int64_t)n1 * nq, pass == 1 ? OP_C : OP_N), one, one, nd, sq, nq, z));
      if (pass == 2)  // X_{q,q+1} = (x_q U) G_{q+1,q+1} + (g_q U) X_{q+1,q+1}  (the latter in w2)
        CHK(gemm(st, B, nq, n1, n1, mat(w1, n1, sx), mat(grd(q + 1), n1, s1), one, one, gnu + ou[q], sx,
                 n1, z, w2));
      CHK(gemm(st, B, nq, nq, n1, mat(gru(q), n1, sx), Lop(q), one, zero, w1, sq, nq, z));
      CHK(gemm(st, B, nq, nq, nq, mat(w1, nq, sq), mat(gi(q), nq, sq), one, one, nd, sq, nq, z));
      // gnu = gnl^H (lesser and greater)
      if (pass < 2) {
        const int64_t tot = (int64_t)B * nq * n1;
        const int grid = (int)std::min<int64_t>((tot + 255) / 256, 65535);
        hipLaunchKernelGGL(bconjt_kernel, dim3(grid), dim3(256), 0, st, gnu + ou[q], nl, n1, nq, B);
        CHK(hipGetLastError());
      }
    }
  }
  return hipSuccess;
}


int64_t rgf_transmission_ops_workspace(int K, const int32_t* bsz, int B) {
  int nmax = 0;
  for (int q = 0; q < K; ++q) nmax = std::max(nmax, (int)bsz[q]);
  return 8 * (int64_t)B * nmax * nmax + inverse_workspace(nmax, B);
}

// Transmission for any block size (<= BLOCK_NMAX) by batched block operations:
// the same sweep as the fused kernel (S_q = G^L_{0,q} U_q carried forward), then
// T = Re Tr[Gamma_L Q Gamma_R Q^H] with Q the corner of G_{0,K-1}.
hipError_t rgf_transmission_ops(const RgfOpsArgs& a, hipStream_t st) {
  const int K = a.K, B = a.B;
  const std::vector<int> n(a.bsz, a.bsz + K);
  int nmax = 0;
  std::vector<int64_t> hoff(K);
  int64_t sh = 0;
  for (int q = 0; q < K; ++q) {
    nmax = std::max(nmax, n[q]);
    hoff[q] = sh;
    sh += 2 * (int64_t)n[q] * n[q];
    if (q + 1 < K) sh += 4 * (int64_t)n[q] * n[q + 1];
  }
  const int64_t m2 = (int64_t)B * nmax * nmax;
  cx<double>* g = a.work;
  cx<double>* Y = g + m2;
  cx<double>* P = Y + m2;
  cx<double>* P2 = P + m2;
  cx<double>* gamL = P2 + m2;
  cx<double>* gamR = gamL + m2;
  cx<double>* V = gamR + m2;
  cx<double>* W = V + m2;
  cx<double>* iws = W + m2;  // inverse workspace
  const cx<double> one{1, 0}, mone{-1, 0}, zero{0, 0};
  const cx<double>* H = a.arena;
  auto hd = [&](int q) { return H + hoff[q]; };
  auto sdp = [&](int q) { return H + hoff[q] + (int64_t)n[q] * n[q]; };
  auto hu = [&](int q) { return H + hoff[q] + 2 * (int64_t)n[q] * n[q]; };
  auto suu = [&](int q) { return hu(q) + (int64_t)n[q] * n[q + 1]; };
  auto hl = [&](int q) { return suu(q) + (int64_t)n[q] * n[q + 1]; };
  auto sl = [&](int q) { return hl(q) + (int64_t)n[q + 1] * n[q]; };
  auto Uop = [&](int q) { return mat(hu(q), n[q + 1], 0, OP_N, a.ortho ? nullptr : suu(q), 0); };
  auto Lop = [&](int q) { return mat(hl(q), n[q], 0, OP_N, a.ortho ? nullptr : sl(q), 0); };
  auto grid_of = [](int64_t tot) { return (int)std::min<int64_t>((tot + 255) / 256, 65535); };
  auto build = [&](int q, bool left, bool right) {
    hipLaunchKernelGGL(bbuild_kernel, dim3(grid_of((int64_t)B * n[q] * n[q])), dim3(256), 0, st, g,
                       hd(q), sdp(q), n[q], a.z, B, left ? a.sigL : nullptr, a.nL,
                       right ? a.sigR : nullptr, a.nR, nullptr);
    return hipGetLastError();
  };
  const int n0 = n[0], nK = n[K - 1], nL = a.nL, nR = a.nR, c0 = nK - nR;
  CHK(build(0, true, K == 1));
  CHK(inverse(st, g, n0, B, a.status, iws));
  // S_q = G^L_{0,q} U_q = S_{q-1} Y_q with Y_q = g_q U_q (S_0 = Y_0), batch
  // stride n0 n_{q+1} and ld n_{q+1}; only the first nL rows, the rows Gamma_L
  // reaches, are carried
  cx<double>* S = P;
  cx<double>* S2 = P2;
  for (int q = 0; q + 1 < K; ++q) {
    const int nq = n[q], n1 = n[q + 1];
    CHK(gemm(st, B, nq, n1, nq, mat(g, nq, (int64_t)nq * nq), Uop(q), one, zero, Y, (int64_t)nq * n1,
             n1, a.z));
    if (q == 0) {
      std::swap(S, Y);  // S_0 = Y_0 (first nL rows read through ld n1)
    } else {
      CHK(gemm(st, B, nL, n1, nq, mat(S, nq, (int64_t)n0 * nq), mat(Y, n1, (int64_t)nq * n1), one,
               zero, S2, (int64_t)n0 * n1, n1, a.z));
      std::swap(S, S2);
    }
    const cx<double>* Yq = q == 0 ? S : Y;
    CHK(build(q + 1, false, q + 1 == K - 1));
    CHK(gemm(st, B, n1, n1, nq, Lop(q), mat(Yq, n1, (int64_t)nq * n1), mone, one, g,
             (int64_t)n1 * n1, n1, a.z));
    CHK(inverse(st, g, n1, B, a.status, iws));
  }
  // Q = G_{0,K-1}[0:nL, c0:c0+nR] = S_{K-2} g_{K-1}[:, c0:] (g_0's corner for K = 1)
  const cx<double>* Q = g + c0;
  int ldq = nK;
  int64_t sq = (int64_t)nK * nK;
  if (K > 1) {  // S_{K-2}: batch stride n0 nK, ld nK
    CHK(gemm(st, B, nL, nR, nK, mat(S, nK, (int64_t)n0 * nK), mat(g + c0, nK, (int64_t)nK * nK),
             one, zero, S2, (int64_t)nL * nR, nR, a.z));
    Q = S2; ldq = nR; sq = (int64_t)nL * nR;
  }
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(grid_of((int64_t)B * nL * nL)), dim3(256), 0, st, gamL,
                     a.sigL, nL, nL, B, 0, (const double*)nullptr, 1.0);
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(grid_of((int64_t)B * nR * nR)), dim3(256), 0, st, gamR,
                     a.sigR, nR, nR, B, 0, (const double*)nullptr, 1.0);
  CHK(hipGetLastError());
  // V = Q Gamma_R   [nL x nR];  W = V Q^H  [nL x nL];  T = Re Tr[Gamma_L W]
  CHK(gemm(st, B, nL, nR, nR, mat(Q, ldq, sq), mat(gamR, nR, (int64_t)nR * nR), one,
           zero, V, (int64_t)nL * nR, nR, a.z));
  CHK(gemm(st, B, nL, nL, nR, mat(V, nR, (int64_t)nL * nR), mat(Q, ldq, sq, OP_H),
           one, zero, W, (int64_t)nL * nL, nL, a.z));
  hipLaunchKernelGGL(btrace_re_kernel, dim3(B), dim3(256), 0, st, a.T, gamL, (int64_t)nL * nL, W,
                     (int64_t)nL * nL, nL, nL);
  return hipGetLastError();
}


int64_t lead_ls_batched_workspace(int n, int nred, int B) {
  const int m = std::max(n, nred);
  return 9 * (int64_t)B * m * m + inverse_workspace(m, B) + 4 * (int64_t)B;  // + int32 flags
}

hipError_t lead_ls_batched(const LeadArgs& a, cx<double>* work, hipStream_t st) {
  const int n = a.n, nr = a.nred, B = a.P;
  const int m = std::max(n, nr);
  const int64_t m2 = (int64_t)B * m * m;
  cx<double>* al = work;
  cx<double>* be = al + m2;
  cx<double>* eps = be + m2;
  cx<double>* W = eps + m2;
  cx<double>* Ai = W + m2;
  cx<double>* Bi = Ai + m2;
  cx<double>* Tm = Bi + m2;
  cx<double>* X1 = Tm + m2;
  cx<double>* X2 = X1 + m2;
  cx<double>* iws = X2 + m2;  // inverse workspace
  // flags: done [B], iteration counts [B], remaining [1], checked sys_idx [B]
  // (3 B + 1 int32 of the 16 B the workspace reserves)
  int32_t* flags = (int32_t*)(iws + inverse_workspace(m, B));
  int32_t* done = flags;
  int32_t* iters = a.iters ? a.iters : flags + B;
  int32_t* remaining = flags + 2 * B;
  int32_t* sysc = flags + 2 * B + 1;
  cx<double>* epss = a.gs;  // eps_s accumulates in the g_s output
  const cx<double> one{1, 0}, zero{0, 0}, mone{-1, 0};
  const int64_t nn = (int64_t)n * n;
  auto grid_of = [](int64_t tot) { return (int)std::min<int64_t>((tot + 255) / 256, 65535); };
  const int gridn = grid_of((int64_t)B * nn);
  int32_t* status = a.status;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(2 * (int64_t)B + 1)), dim3(256), 0, st, flags, 0,
                     2 * (int64_t)B + 1);
  if (status) hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(B)), dim3(256), 0, st, status, 0, (int64_t)B);
  CHK(hipGetLastError());
  const int32_t* sys_idx = nullptr;
  if (a.sys_idx) {
    hipLaunchKernelGGL(ls_sys_check_kernel, dim3((B + 255) / 256), dim3(256), 0, st, sysc, a.sys_idx,
                       a.nsys, status, B);
    CHK(hipGetLastError());
    sys_idx = sysc;
  }
  hipLaunchKernelGGL(ls_init_kernel, dim3(gridn), dim3(256), 0, st, al, be, eps, epss, a.lead,
                     a.lead_stride, sys_idx, a.e, n, B, a.eta);
  CHK(hipGetLastError());
  auto M = [&](cx<double>* p) { return mat(p, n, nn); };
  // Points that have converged are frozen (done[b]: every kernel of the loop
  // leaves them untouched), as the reference stops iterating a point once it
  // converges.  The host reads the count of running points only every few
  // iterations, and never while the stream is being captured into a graph:
  // then all 101 iterations are launched and converged points cost a read of
  // their flag.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  CHK(hipStreamIsCapturing(st, &cap));
  const bool capturing = cap != hipStreamCaptureStatusNone;
  const int32_t* skip = done;
  int it = 0;
  bool converged = false;
  while (true) {
    ++it;
    hipLaunchKernelGGL(ls_zs_minus_kernel, dim3(gridn), dim3(256), 0, st, W, eps, a.lead,
                       a.lead_stride, sys_idx, a.e, n, B, a.eta, skip);
    CHK(hipGetLastError());
    // X_a = W^-1 alpha, X_b = W^-1 beta solved as the reference solves them (surface_green.py:189-194):
    // through the inverse, refined once in FP64, X += W^-1 (rhs - W X).  Without the
    // refinement, W^-1 times rhs loses the Dyson residual near band edges at small eta.
    CHK(hipMemcpyAsync(X1, W, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
    CHK(inverse(st, W, n, B, status, iws, X1, skip));
    cx<double>* rhs[2] = {al, be};
    cx<double>* sol[2] = {Ai, Bi};
    for (int q = 0; q < 2; ++q) {
      CHK(gemm(st, B, n, n, n, M(W), M(rhs[q]), one, zero, sol[q], nn, n, nullptr, nullptr, skip));
      // X2 = rhs - W X
      CHK(gemm(st, B, n, n, n, M(X1), M(sol[q]), mone, one, X2, nn, n, nullptr, rhs[q], skip));
      CHK(gemm(st, B, n, n, n, M(W), M(X2), one, one, sol[q], nn, n, nullptr, nullptr, skip));
    }
    // eps += alpha X_b + beta X_a ; eps_s += beta X_a (beta X_a formed once, as the
    // reference does) ; alpha <- alpha X_a ; beta <- beta X_b.  A frozen point's
    // alpha / beta slots go stale through the swaps; they are never read again.
    CHK(gemm(st, B, n, n, n, M(al), M(Bi), one, one, eps, nn, n, nullptr, nullptr, skip));
    CHK(gemm(st, B, n, n, n, M(be), M(Ai), one, zero, Tm, nn, n, nullptr, nullptr, skip));
    hipLaunchKernelGGL(badd2_kernel, dim3(gridn), dim3(256), 0, st, eps, epss, Tm, (int64_t)B * nn,
                       nn, skip);
    CHK(hipGetLastError());
    CHK(gemm(st, B, n, n, n, M(al), M(Ai), one, zero, Tm, nn, n, nullptr, nullptr, skip));
    std::swap(al, Tm);
    CHK(gemm(st, B, n, n, n, M(be), M(Bi), one, zero, Tm, nn, n, nullptr, nullptr, skip));
    std::swap(be, Tm);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, st, remaining, 0, (int64_t)1);
    hipLaunchKernelGGL(ls_conv_kernel, dim3(B), dim3(256), 0, st, al, be, n, iters, done, remaining,
                       it, a.stop);
    CHK(hipGetLastError());
    if (!capturing && (it >= 101 || (it >= 8 && it % 4 == 0))) {
      int32_t left = 0;
      CHK(hipMemcpyAsync(&left, remaining, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      CHK(hipStreamSynchronize(st));
      if (left == 0) { converged = true; break; }
    }
    if (it >= 101) break;
  }
  if (!converged) {
    // points still running at the iteration limit (surface_green.py:222-224)
    hipLaunchKernelGGL(ls_mark_kernel, dim3((B + 255) / 256), dim3(256), 0, st, done, status, iters,
                       B, it);
    CHK(hipGetLastError());
  }
  // g_s = (z S - eps_s)^-1
  hipLaunchKernelGGL(ls_zs_minus_kernel, dim3(gridn), dim3(256), 0, st, W, epss, a.lead,
                     a.lead_stride, sys_idx, a.e, n, B, a.eta, nullptr);
  CHK(hipGetLastError());
  if (inverse_form(n, true) != INV_SCHUR) {
    // one FP64 Newton step on the one-workgroup inverse, X <- X + X (I - A X)
    // (the Schur form takes its own)
    CHK(hipMemcpyAsync(X1, W, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
    CHK(inverse(st, W, n, B, status, iws));
    CHK(gemm(st, B, n, n, n, M(X1), M(W), mone, zero, X2, nn, n, nullptr));
    hipLaunchKernelGGL(badd_identity_kernel, dim3(grid_of((int64_t)B * n)), dim3(256), 0, st, X2, n,
                       B, nullptr);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(Tm, W, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
    CHK(gemm(st, B, n, n, n, M(Tm), M(X2), one, one, W, nn, n, nullptr));
  } else {
    CHK(inverse(st, W, n, B, status, iws));
  }
  CHK(hipMemcpyAsync(a.gs, W, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
  const cx<double>* gs = a.gs;
  // residual: X1 = (z s01 - h01) g_s ; X2 = X1 (z s10 - h10) ; R = (zS - H - X2) g_s
  hipLaunchKernelGGL(ls_coupling_kernel, dim3(gridn), dim3(256), 0, st, Ai, a.lead, a.lead_stride,
                     sys_idx, a.e, a.eta, nn, 3 * nn, n, n, 0, B);
  hipLaunchKernelGGL(ls_coupling_kernel, dim3(gridn), dim3(256), 0, st, Bi, a.lead, a.lead_stride,
                     sys_idx, a.e, a.eta, nn, 3 * nn, n, n, 1, B);
  CHK(hipGetLastError());
  CHK(gemm(st, B, n, n, n, M(Ai), mat(gs, n, nn), one, zero, X1, nn, n, nullptr));
  CHK(gemm(st, B, n, n, n, M(X1), M(Bi), one, zero, X2, nn, n, nullptr));
  hipLaunchKernelGGL(ls_test_kernel, dim3(gridn), dim3(256), 0, st, Tm, X2, a.lead, a.lead_stride,
                     sys_idx, a.e, n, B, a.eta);
  CHK(hipGetLastError());
  CHK(gemm(st, B, n, n, n, M(Tm), mat(gs, n, nn), one, zero, X1, nn, n, nullptr));
  hipLaunchKernelGGL(ls_resid_kernel, dim3(B), dim3(256), 0, st, X1, n, a.resid, status);
  CHK(hipGetLastError());
  // Sigma
  const int64_t dl = 4 * nn, sdl = 4 * nn + (int64_t)nr * n;
  if (a.mode == 0) {
    // tau1 = e S_DL - H_DL [nr x n], tau2 = e S_DL^H - H_DL^H [n x nr]  (real-axis e: eta 0)
    hipLaunchKernelGGL(ls_coupling_kernel, dim3(grid_of((int64_t)B * nr * n)), dim3(256), 0, st, Ai,
                       a.lead, a.lead_stride, sys_idx, a.e, 0.0, dl, sdl, nr, n, 0, B);
    hipLaunchKernelGGL(ls_coupling_kernel, dim3(grid_of((int64_t)B * nr * n)), dim3(256), 0, st, Bi,
                       a.lead, a.lead_stride, sys_idx, a.e, 0.0, dl, sdl, nr, n, 1, B);
    CHK(hipGetLastError());
    CHK(gemm(st, B, n, nr, n, mat(gs, n, nn), mat(Bi, nr, (int64_t)n * nr), one, zero, X1,
             (int64_t)n * nr, nr, nullptr));
    CHK(gemm(st, B, nr, nr, n, mat(Ai, n, (int64_t)nr * n), mat(X1, nr, (int64_t)n * nr), one, zero,
             a.sigma, (int64_t)nr * nr, nr, nullptr));
  } else {
    CHK(hipMemcpyAsync(W, gs, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
    CHK(inverse(st, W, n, B, status, iws));
    if (a.mode == 1) {
      // Sigma = (e S - H) - g_s^-1 : the coupling kernel with h = H (offset 0), s = S (2nn)
      hipLaunchKernelGGL(ls_coupling_kernel, dim3(gridn), dim3(256), 0, st, X1, a.lead,
                         a.lead_stride, sys_idx, a.e, 0.0, (int64_t)0, 2 * nn, n, n, 0, B);
      hipLaunchKernelGGL(bsub_kernel, dim3(gridn), dim3(256), 0, st, a.sigma, X1, W, (int64_t)B * nn);
      CHK(hipGetLastError());
    } else {
      CHK(hipMemcpyAsync(a.sigma, W, sizeof(cx<double>) * B * nn, hipMemcpyDeviceToDevice, st));
    }
  }
  return hipSuccess;
}

int64_t rgf_grad_workspace(int K, const int32_t* bsz, int B) {
  int64_t nsq = 0, nx = 0, nlc = 0;
  int nmax = 0;
  for (int q = 0; q < K; ++q) {
    nsq += (int64_t)bsz[q] * bsz[q];
    nlc += (int64_t)bsz[q] * bsz[K - 1];
    if (q + 1 < K) nx += (int64_t)bsz[q] * bsz[q + 1];
    nmax = std::max(nmax, (int)bsz[q]);
  }
  const int64_t m2 = (int64_t)nmax * nmax;
  // gl, grd | grl, gru | lc | g_trans | work(5) | gammaL, gammaR, C | R x2, G0 x2, D x2, W
  return B * (2 * nsq + 2 * std::max<int64_t>(nx, 1) + nlc + m2 + 5 * m2 + 3 * m2 + 7 * m2) +
         inverse_workspace(nmax, B);
}

// bcorner_conjt_kernel over the B energies of an adjoint
static hipError_t sig_corner(cx<double>* out, const cx<double>* X, int ld, int off, int m, cx<double> c,
                             const double* w, int accumulate, int B, hipStream_t st) {
  const int64_t tot = (int64_t)B * m * m;
  hipLaunchKernelGGL(bcorner_conjt_kernel, dim3((int)std::min<int64_t>((tot + 255) / 256, 65535)),
                     dim3(256), 0, st, out, X, ld, off, m, B, c, w, accumulate);
  return hipGetLastError();
}

hipError_t rgf_grad(const RgfGradArgs& g, hipStream_t st) {
  const int K = g.K, B = g.B;
  const std::vector<int> n(g.bsz, g.bsz + K);
  int64_t nsq = 0, nx = 0, nlc = 0;
  int nmax = 0;
  std::vector<int64_t> od(K), ou(K > 1 ? K - 1 : 1), olc(K), hoff(K), gsq(K), gx(K > 1 ? K - 1 : 1);
  int64_t sh = 0, s_sq = 0, s_x = 0;
  for (int q = 0; q < K; ++q) {
    od[q] = (int64_t)B * nsq; gsq[q] = s_sq;
    nsq += (int64_t)n[q] * n[q]; s_sq += (int64_t)n[q] * n[q];
    olc[q] = (int64_t)B * nlc; nlc += (int64_t)n[q] * n[K - 1];
    hoff[q] = sh; sh += 2 * (int64_t)n[q] * n[q];
    if (q + 1 < K) {
      ou[q] = (int64_t)B * nx; gx[q] = s_x;
      nx += (int64_t)n[q] * n[q + 1]; s_x += (int64_t)n[q] * n[q + 1];
      sh += 4 * (int64_t)n[q] * n[q + 1];
    }
    nmax = std::max(nmax, n[q]);
  }
  const int64_t m2 = (int64_t)B * nmax * nmax;
  cx<double>* w = g.work;
  RgfFullArgs a;
  std::memset(&a, 0, sizeof(a));
  a.K = K; a.B = B; a.nmax = nmax; a.ortho = g.ortho; a.bsz = g.bsz; a.arena = g.arena; a.z = g.z;
  a.sigL = g.sigL; a.nL = g.nL; a.sigR = g.sigR; a.nR = g.nR; a.status = g.status;
  a.gl = w; w += (int64_t)B * nsq;
  a.grd = w; w += (int64_t)B * nsq;
  a.grl = w; w += (int64_t)B * std::max<int64_t>(nx, 1);
  a.gru = w; w += (int64_t)B * std::max<int64_t>(nx, 1);
  a.lc = w; w += (int64_t)B * nlc;
  a.g_trans = w; w += m2;
  a.work = w; w += 5 * m2;
  a.inv_ws = g.work + (rgf_grad_workspace(K, g.bsz, B) - inverse_workspace(nmax, B));
  CHK(rgf_full(a, st));
  cx<double>* gamL = w; w += m2;
  cx<double>* gamR = w; w += m2;
  cx<double>* C = w; w += m2;
  cx<double>* Rb[2] = {w, w + m2}; w += 2 * m2;
  cx<double>* G0[2] = {w, w + m2}; w += 2 * m2;
  cx<double>* D[2] = {w, w + m2}; w += 2 * m2;
  cx<double>* Wm = w; w += m2;
  const cx<double> one{1, 0}, zero{0, 0};
  const cx<double>* H = g.arena;
  auto hu = [&](int q) { return H + hoff[q] + 2 * (int64_t)n[q] * n[q]; };
  auto suu = [&](int q) { return hu(q) + (int64_t)n[q] * n[q + 1]; };
  auto Uop = [&](int q) { return mat(hu(q), n[q + 1], 0, OP_N, g.ortho ? nullptr : suu(q), 0); };
  const int n0 = n[0], nK = n[K - 1];
  auto lc = [&](int q) { return a.lc + olc[q]; };
  auto gl = [&](int q) { return a.gl + od[q]; };
  auto grd = [&](int q) { return a.grd + od[q]; };
  auto launch_grid = [](int64_t tot) { return (int)std::min<int64_t>((tot + 255) / 256, 65535); };

  // C = Gamma_R G_{0,K-1}^H (2 w Gamma_L)      [nK x n0]
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(launch_grid((int64_t)B * n0 * n0)), dim3(256), 0, st,
                     gamL, g.sigL, g.nL, n0, B, 0, g.w, 2.0);
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(launch_grid((int64_t)B * nK * nK)), dim3(256), 0, st,
                     gamR, g.sigR, g.nR, nK, B, 1, (const double*)nullptr, 1.0);
  CHK(hipGetLastError());
  CHK(gemm(st, B, nK, n0, nK, mat(gamR, nK, (int64_t)nK * nK), mat(lc(0), nK, (int64_t)n0 * nK, OP_H),
           one, zero, Wm, (int64_t)nK * n0, n0, g.z));
  CHK(gemm(st, B, nK, n0, n0, mat(Wm, n0, (int64_t)nK * n0), mat(gamL, n0, (int64_t)n0 * n0), one,
           zero, C, (int64_t)nK * n0, n0, g.z));
  // first block row: G_{0,0} = grd_0; G_{0,b} = R_b G_bb, R_1 = g_0 U_0, R_{b+1} = R_b g_b U_b
  auto reduce = [&](cx<double>* out, const cx<double>* in, int m_, int n_) {
    hipLaunchKernelGGL(breduce_conjt_kernel, dim3((m_ * n_ + RED_ELEMS - 1) / RED_ELEMS), dim3(256),
                       0, st, out, in, m_, n_, B, 0, (const cx<double>*)nullptr, 1.0);
    return hipGetLastError();
  };
  // M = z S - H: the overlap block's gradient is -sum_b conj(z_b) (W_b)^H
  // against H's sum_b (W_b)^H
  auto reduce_s = [&](cx<double>* out, const cx<double>* in, int m_, int n_) -> hipError_t {
    if (!out) return hipSuccess;
    hipLaunchKernelGGL(breduce_conjt_kernel, dim3((m_ * n_ + RED_ELEMS - 1) / RED_ELEMS), dim3(256),
                       0, st, out, in, m_, n_, B, 0, g.z, -1.0);
    return hipGetLastError();
  };
  const cx<double>* g0b = grd(0);  // G_{0,0}
  int cur = 0;
  // D_0 = C G_{0,0}
  CHK(gemm(st, B, nK, n0, n0, mat(C, n0, (int64_t)nK * n0), mat(g0b, n0, (int64_t)n0 * n0), one,
           zero, D[0], (int64_t)nK * n0, n0, g.z));
  for (int q = 0; q < K; ++q) {
    const int nq = n[q];
    // grad_hd[q] = sum_b (G_{q,K-1} D_q)^H
    CHK(gemm(st, B, nq, nq, nK, mat(lc(q), nK, (int64_t)nq * nK), mat(D[cur], nq, (int64_t)nK * nq),
             one, zero, Wm, (int64_t)nq * nq, nq, g.z));
    CHK(reduce(g.grad_hd + gsq[q], Wm, nq, nq));
    CHK(reduce_s(g.grad_sd ? g.grad_sd + gsq[q] : nullptr, Wm, nq, nq));
    // the self-energies enter M_0 / M_{K-1} as H does: their corners of
    // (G_{q,K-1} D_q)^H, per energy
    if (q == 0 && g.grad_sigL) CHK(sig_corner(g.grad_sigL, Wm, n0, 0, g.nL, one, nullptr, 0, B, st));
    if (q == K - 1 && g.grad_sigR)
      CHK(sig_corner(g.grad_sigR, Wm, nK, nK - g.nR, g.nR, one, nullptr, 0, B, st));
    if (q + 1 == K) break;
    const int n1 = n[q + 1];
    // R_{q+1} = (q == 0 ? g_0 : R_q g_q) U_q
    const cx<double>* left = gl(0);
    int64_t lstride = (int64_t)n0 * n0;
    if (q > 0) {
      CHK(gemm(st, B, n0, nq, nq, mat(Rb[cur], nq, (int64_t)n0 * nq), mat(gl(q), nq, (int64_t)nq * nq),
               one, zero, G0[1 - cur], (int64_t)n0 * nq, nq, g.z));
      left = G0[1 - cur];
      lstride = (int64_t)n0 * nq;
    }
    CHK(gemm(st, B, n0, n1, nq, mat(left, nq, lstride), Uop(q), one, zero, Rb[1 - cur],
             (int64_t)n0 * n1, n1, g.z));
    // G_{0,q+1} = R_{q+1} G_{q+1,q+1};  D_{q+1} = C G_{0,q+1}
    CHK(gemm(st, B, n0, n1, n1, mat(Rb[1 - cur], n1, (int64_t)n0 * n1),
             mat(grd(q + 1), n1, (int64_t)n1 * n1), one, zero, G0[cur], (int64_t)n0 * n1, n1, g.z));
    CHK(gemm(st, B, nK, n1, n0, mat(C, n0, (int64_t)nK * n0), mat(G0[cur], n1, (int64_t)n0 * n1), one,
             zero, D[1 - cur], (int64_t)nK * n1, n1, g.z));
    // grad_hu[q] = sum_b (G_{q+1,K-1} D_q)^H ;  grad_hl[q] = sum_b (G_{q,K-1} D_{q+1})^H
    CHK(gemm(st, B, n1, nq, nK, mat(lc(q + 1), nK, (int64_t)n1 * nK), mat(D[cur], nq, (int64_t)nK * nq),
             one, zero, Wm, (int64_t)n1 * nq, nq, g.z));
    CHK(reduce(g.grad_hu + gx[q], Wm, n1, nq));
    CHK(reduce_s(g.grad_su ? g.grad_su + gx[q] : nullptr, Wm, n1, nq));
    CHK(gemm(st, B, nq, n1, nK, mat(lc(q), nK, (int64_t)nq * nK), mat(D[1 - cur], n1, (int64_t)nK * n1),
             one, zero, Wm, (int64_t)nq * n1, n1, g.z));
    CHK(reduce(g.grad_hl + gx[q], Wm, nq, n1));
    CHK(reduce_s(g.grad_sl ? g.grad_sl + gx[q] : nullptr, Wm, nq, n1));
    cur = 1 - cur;
  }
  // and through Gamma: with Q = G_{0,K-1}, T = Tr[Gamma_L Q Gamma_R Q^H] and
  // Gamma = i (Sigma - Sigma^H) add i 2 w (Q Gamma_R Q^H) to Sigma_L's and
  // i (Q^H (2 w Gamma_L) Q) to Sigma_R's term of dJ = Re Tr[V dSigma]
  if (g.grad_sigL) {
    CHK(gemm(st, B, n0, nK, nK, mat(lc(0), nK, (int64_t)n0 * nK), mat(gamR, nK, (int64_t)nK * nK),
             one, zero, Rb[0], (int64_t)n0 * nK, nK, g.z));
    CHK(gemm(st, B, n0, n0, nK, mat(Rb[0], nK, (int64_t)n0 * nK),
             mat(lc(0), nK, (int64_t)n0 * nK, OP_H), one, zero, G0[0], (int64_t)n0 * n0, n0, g.z));
    CHK(sig_corner(g.grad_sigL, G0[0], n0, 0, g.nL, cx<double>{0, 1}, g.w, 1, B, st));
  }
  if (g.grad_sigR) {
    CHK(gemm(st, B, nK, n0, n0, mat(lc(0), nK, (int64_t)n0 * nK, OP_H), mat(gamL, n0, (int64_t)n0 * n0),
             one, zero, Rb[1], (int64_t)nK * n0, n0, g.z));
    CHK(gemm(st, B, nK, nK, n0, mat(Rb[1], n0, (int64_t)nK * n0), mat(lc(0), nK, (int64_t)n0 * nK),
             one, zero, G0[1], (int64_t)nK * nK, nK, g.z));
    CHK(sig_corner(g.grad_sigR, G0[1], nK, nK - g.nR, g.nR, cx<double>{0, 1}, nullptr, 1, B, st));
  }
  return hipSuccess;
}

// ---- adjoint with sqrt(K) checkpointing ------------------------------------
// The same gradients as rgf_grad, in O(K / s + s) blocks per energy instead of
// O(K): the forward sweep keeps the left-connected g_c and the first-row
// product R_c = g_0 U_0 ... g_{c-1} U_{c-1} only at c = 0, s, 2s, ...; the
// gradient sweep then runs down the blocks one segment [c, c + s) at a time,
// from the right, recomputing the segment's g_q and R_q from its checkpoint
// and carrying G_{q+1,q+1}, G_{q+1,K-1} and D_{q+1} = C G_{0,q+1} across
// segment boundaries.  Cost: one more forward sweep than rgf_grad.

// Per-energy element counts, in workspace order.
struct CkptLayout {
  int nseg, nmax;
  int64_t ck, m2, per;
};
static CkptLayout ckpt_layout(int K, const int32_t* bsz, int stride) {
  CkptLayout L;
  L.nmax = 0;
  for (int q = 0; q < K; ++q) L.nmax = std::max(L.nmax, (int)bsz[q]);
  L.nseg = (K + stride - 1) / stride;
  L.ck = 0;
  for (int j = 0; j < L.nseg; ++j) {
    const int64_t nc = bsz[j * stride];
    L.ck += nc * nc + (j > 0 ? (int64_t)bsz[0] * nc : 0);
  }
  L.m2 = (int64_t)L.nmax * L.nmax;
  // checkpoints | segment g, R (stride each) | carries G_qq, G_{q,K-1}, D (x2) |
  // C, Gamma_L, Gamma_R | forward g, R (x2) | five temporaries
  L.per = L.ck + (2 * (int64_t)stride + 6 + 3 + 4 + 5) * L.m2;
  return L;
}

int64_t rgf_grad_ckpt_workspace(int K, const int32_t* bsz, int B, int stride) {
  if (stride < 1) return -1;
  const CkptLayout L = ckpt_layout(K, bsz, stride);
  return (int64_t)B * L.per + inverse_workspace(L.nmax, B);
}

hipError_t rgf_grad_ckpt(const RgfGradArgs& g, int stride, hipStream_t st) {
  const int K = g.K, B = g.B;
  if (stride < 1) return hipErrorInvalidValue;
  const std::vector<int> n(g.bsz, g.bsz + K);
  const CkptLayout Ly = ckpt_layout(K, g.bsz, stride);
  const int n0 = n[0], nK = n[K - 1], nseg = Ly.nseg;
  const int64_t M2 = (int64_t)B * Ly.m2;
  std::vector<int64_t> hoff(K), gsq(K), gx(K > 1 ? K - 1 : 1);
  int64_t sh = 0, s_sq = 0, s_x = 0;
  for (int q = 0; q < K; ++q) {
    hoff[q] = sh; sh += 2 * (int64_t)n[q] * n[q];
    gsq[q] = s_sq; s_sq += (int64_t)n[q] * n[q];
    if (q + 1 < K) {
      gx[q] = s_x; s_x += (int64_t)n[q] * n[q + 1];
      sh += 4 * (int64_t)n[q] * n[q + 1];
    }
  }
  const cx<double>* H = g.arena;
  auto hd = [&](int q) { return H + hoff[q]; };
  auto sdp = [&](int q) { return H + hoff[q] + (int64_t)n[q] * n[q]; };
  auto hu = [&](int q) { return H + hoff[q] + 2 * (int64_t)n[q] * n[q]; };
  auto suu = [&](int q) { return hu(q) + (int64_t)n[q] * n[q + 1]; };
  auto hl = [&](int q) { return suu(q) + (int64_t)n[q] * n[q + 1]; };
  auto sl = [&](int q) { return hl(q) + (int64_t)n[q + 1] * n[q]; };
  auto Uop = [&](int q) { return mat(hu(q), n[q + 1], 0, OP_N, g.ortho ? nullptr : suu(q), 0); };
  auto Lop = [&](int q) { return mat(hl(q), n[q], 0, OP_N, g.ortho ? nullptr : sl(q), 0); };
  auto M = [&](const cx<double>* p, int ld, int rows, int op = OP_N) {
    return mat(p, ld, (int64_t)rows * ld, op);
  };
  // workspace, in the order of ckpt_layout
  cx<double>* w = g.work;
  std::vector<cx<double>*> ckg(nseg), ckR(nseg, nullptr);
  for (int j = 0; j < nseg; ++j) {
    const int nc = n[j * stride];
    ckg[j] = w; w += (int64_t)B * nc * nc;
    if (j > 0) { ckR[j] = w; w += (int64_t)B * n0 * nc; }
  }
  cx<double>* segg = w; w += stride * M2;  // slot i: g_{c+i}
  cx<double>* segR = w; w += stride * M2;  // slot i: R_{c+i} (n0 x n_{c+i}); unused for q = 0
  cx<double>* Gd[2] = {w, w + M2}; w += 2 * M2;  // G_qq
  cx<double>* Gc[2] = {w, w + M2}; w += 2 * M2;  // G_{q,K-1}
  cx<double>* D[2] = {w, w + M2}; w += 2 * M2;   // C G_{0,q}
  cx<double>* C = w; w += M2;
  cx<double>* gamL = w; w += M2;
  cx<double>* gamR = w; w += M2;
  cx<double>* fg[2] = {w, w + M2}; w += 2 * M2;
  cx<double>* fR[2] = {w, w + M2}; w += 2 * M2;
  cx<double>* t0 = w; w += M2;
  cx<double>* t1 = w; w += M2;
  cx<double>* t2 = w; w += M2;
  cx<double>* t3 = w; w += M2;
  cx<double>* Wm = w; w += M2;
  cx<double>* inv_ws = w;
  const cx<double> one{1, 0}, mone{-1, 0}, zero{0, 0};
  auto slot = [&](cx<double>* base, int i) { return base + i * M2; };
  auto copy = [&](cx<double>* dst, const cx<double>* src, int64_t elems) {
    return hipMemcpyAsync(dst, src, sizeof(cx<double>) * elems, hipMemcpyDeviceToDevice, st);
  };
  auto build = [&](cx<double>* dst, int q) {
    const int64_t tot = (int64_t)B * n[q] * n[q];
    const int grid = (int)std::min<int64_t>((tot + 255) / 256, 65535);
    hipLaunchKernelGGL(bbuild_kernel, dim3(grid), dim3(256), 0, st, dst, hd(q), sdp(q), n[q], g.z, B,
                       q == 0 ? g.sigL : nullptr, g.nL, q == K - 1 ? g.sigR : nullptr, g.nR, nullptr);
    return hipGetLastError();
  };
  // g_{q+1} from g_q  (:198-207)
  auto forward = [&](const cx<double>* gq, int q, cx<double>* gnext) -> hipError_t {
    const int nq = n[q], n1 = n[q + 1];
    CHK(gemm(st, B, nq, n1, nq, M(gq, nq, nq), Uop(q), one, zero, t0, (int64_t)nq * n1, n1, g.z));
    CHK(build(gnext, q + 1));
    CHK(gemm(st, B, n1, n1, nq, Lop(q), M(t0, n1, nq), mone, one, gnext, (int64_t)n1 * n1, n1, g.z));
    return inverse(st, gnext, n1, B, g.status, inv_ws);
  };
  // R_{q+1} = R_q g_q U_q, R_1 = g_0 U_0  (Rq unused for q = 0)
  auto rnext = [&](const cx<double>* Rq, const cx<double>* gq, int q, cx<double>* Rout) -> hipError_t {
    const int nq = n[q], n1 = n[q + 1];
    const cx<double>* left = gq;
    if (q > 0) {
      CHK(gemm(st, B, n0, nq, nq, M(Rq, nq, n0), M(gq, nq, nq), one, zero, t1, (int64_t)n0 * nq, nq,
               g.z));
      left = t1;
    }
    return gemm(st, B, n0, n1, nq, M(left, nq, q > 0 ? n0 : nq), Uop(q), one, zero, Rout,
                (int64_t)n0 * n1, n1, g.z);
  };
  auto reduce = [&](cx<double>* out, const cx<double>* in, int m_, int n_) {
    hipLaunchKernelGGL(breduce_conjt_kernel, dim3((m_ * n_ + RED_ELEMS - 1) / RED_ELEMS), dim3(256),
                       0, st, out, in, m_, n_, B, 0, (const cx<double>*)nullptr, 1.0);
    return hipGetLastError();
  };
  // M = z S - H: the overlap block's gradient is -sum_b conj(z_b) (W_b)^H
  // against H's sum_b (W_b)^H
  auto reduce_s = [&](cx<double>* out, const cx<double>* in, int m_, int n_) -> hipError_t {
    if (!out) return hipSuccess;
    hipLaunchKernelGGL(breduce_conjt_kernel, dim3((m_ * n_ + RED_ELEMS - 1) / RED_ELEMS), dim3(256),
                       0, st, out, in, m_, n_, B, 0, g.z, -1.0);
    return hipGetLastError();
  };

  // ---- forward: checkpoints and G_{0,K-1} = R_{K-1} g_{K-1}
  CHK(build(fg[0], 0));
  CHK(inverse(st, fg[0], n0, B, g.status, inv_ws));
  CHK(copy(ckg[0], fg[0], (int64_t)B * n0 * n0));
  int cg = 0, cr = 0;
  for (int q = 0; q + 1 < K; ++q) {
    CHK(rnext(fR[cr], fg[cg], q, fR[1 - cr]));
    CHK(forward(fg[cg], q, fg[1 - cg]));
    cg = 1 - cg;
    cr = 1 - cr;
    if ((q + 1) % stride == 0) {
      const int j = (q + 1) / stride, nc = n[q + 1];
      CHK(copy(ckg[j], fg[cg], (int64_t)B * nc * nc));
      CHK(copy(ckR[j], fR[cr], (int64_t)B * n0 * nc));
    }
  }
  const cx<double>* g0K = fg[cg];  // K = 1: G_{0,0} = g_0
  if (K > 1) {
    CHK(gemm(st, B, n0, nK, nK, M(fR[cr], nK, n0), M(fg[cg], nK, nK), one, zero, t2,
             (int64_t)n0 * nK, nK, g.z));
    g0K = t2;
  }
  // C = Gamma_R G_{0,K-1}^H (2 w Gamma_L)      [nK x n0]
  auto grid_of = [](int64_t tot) { return (int)std::min<int64_t>((tot + 255) / 256, 65535); };
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(grid_of((int64_t)B * n0 * n0)), dim3(256), 0, st, gamL,
                     g.sigL, g.nL, n0, B, 0, g.w, 2.0);
  hipLaunchKernelGGL(bgamma_pad_kernel, dim3(grid_of((int64_t)B * nK * nK)), dim3(256), 0, st, gamR,
                     g.sigR, g.nR, nK, B, 1, (const double*)nullptr, 1.0);
  CHK(hipGetLastError());
  CHK(gemm(st, B, nK, n0, nK, M(gamR, nK, nK), M(g0K, nK, n0, OP_H), one, zero, Wm,
           (int64_t)nK * n0, n0, g.z));
  CHK(gemm(st, B, nK, n0, n0, M(Wm, n0, nK), M(gamL, n0, n0), one, zero, C, (int64_t)nK * n0, n0,
           g.z));
  // the self-energies' terms through Gamma (as in rgf_grad), before t0..t3
  // serve the sweep; their terms through G follow at q = 0 and q = K - 1
  if (g.grad_sigL) {
    CHK(gemm(st, B, n0, nK, nK, M(g0K, nK, n0), M(gamR, nK, nK), one, zero, t0, (int64_t)n0 * nK,
             nK, g.z));
    CHK(gemm(st, B, n0, n0, nK, M(t0, nK, n0), M(g0K, nK, n0, OP_H), one, zero, t1, (int64_t)n0 * n0,
             n0, g.z));
    CHK(sig_corner(g.grad_sigL, t1, n0, 0, g.nL, cx<double>{0, 1}, g.w, 0, B, st));
  }
  if (g.grad_sigR) {
    CHK(gemm(st, B, nK, n0, n0, M(g0K, nK, n0, OP_H), M(gamL, n0, n0), one, zero, t0,
             (int64_t)nK * n0, n0, g.z));
    CHK(gemm(st, B, nK, nK, n0, M(t0, n0, nK), M(g0K, nK, n0), one, zero, t3, (int64_t)nK * nK, nK,
             g.z));
    CHK(sig_corner(g.grad_sigR, t3, nK, nK - g.nR, g.nR, cx<double>{0, 1}, nullptr, 0, B, st));
  }

  // ---- gradient sweep, segment by segment from the right
  int cur = 0;  // carries of block q in [cur], of block q + 1 in [1 - cur]
  for (int j = nseg - 1; j >= 0; --j) {
    const int c = j * stride, e = std::min(c + stride, K);
    CHK(copy(slot(segg, 0), ckg[j], (int64_t)B * n[c] * n[c]));
    if (c > 0) CHK(copy(slot(segR, 0), ckR[j], (int64_t)B * n0 * n[c]));
    for (int i = 0; c + i + 1 < e; ++i) {
      CHK(forward(slot(segg, i), c + i, slot(segg, i + 1)));
      CHK(rnext(slot(segR, i), slot(segg, i), c + i, slot(segR, i + 1)));
    }
    for (int q = e - 1; q >= c; --q) {
      const int i = q - c, nq = n[q];
      const cx<double>* gq = slot(segg, i);
      if (q == K - 1) {
        CHK(copy(Gd[cur], gq, (int64_t)B * nq * nq));
        CHK(copy(Gc[cur], gq, (int64_t)B * nq * nK));
      } else {
        const int n1 = n[q + 1];
        // gU = g_q U_q;  G_{q+1,q} = G_{q+1,q+1} L_q g_q;  G_qq = g_q + gU G_{q+1,q}
        CHK(gemm(st, B, nq, n1, nq, M(gq, nq, nq), Uop(q), one, zero, t0, (int64_t)nq * n1, n1, g.z));
        CHK(gemm(st, B, n1, nq, n1, M(Gd[1 - cur], n1, n1), Lop(q), one, zero, t1, (int64_t)n1 * nq,
                 nq, g.z));
        CHK(gemm(st, B, n1, nq, nq, M(t1, nq, n1), M(gq, nq, nq), one, zero, t2, (int64_t)n1 * nq, nq,
                 g.z));
        CHK(copy(Gd[cur], gq, (int64_t)B * nq * nq));
        CHK(gemm(st, B, nq, nq, n1, M(t0, n1, nq), M(t2, nq, n1), one, one, Gd[cur], (int64_t)nq * nq,
                 nq, g.z));
        // G_{q,K-1} = gU G_{q+1,K-1}
        CHK(gemm(st, B, nq, nK, n1, M(t0, n1, nq), M(Gc[1 - cur], nK, n1), one, zero, Gc[cur],
                 (int64_t)nq * nK, nK, g.z));
      }
      // G_{0,q} = R_q G_qq (G_{0,0} = G_00);  D_q = C G_{0,q}
      const cx<double>* g0q = Gd[cur];
      if (q > 0) {
        CHK(gemm(st, B, n0, nq, nq, M(slot(segR, i), nq, n0), M(Gd[cur], nq, nq), one, zero, t3,
                 (int64_t)n0 * nq, nq, g.z));
        g0q = t3;
      }
      CHK(gemm(st, B, nK, nq, n0, M(C, n0, nK), M(g0q, nq, n0), one, zero, D[cur], (int64_t)nK * nq,
               nq, g.z));
      // grad_hd[q] = sum_b (G_{q,K-1} D_q)^H
      CHK(gemm(st, B, nq, nq, nK, M(Gc[cur], nK, nq), M(D[cur], nq, nK), one, zero, Wm,
               (int64_t)nq * nq, nq, g.z));
      CHK(reduce(g.grad_hd + gsq[q], Wm, nq, nq));
      CHK(reduce_s(g.grad_sd ? g.grad_sd + gsq[q] : nullptr, Wm, nq, nq));
      if (q == 0 && g.grad_sigL) CHK(sig_corner(g.grad_sigL, Wm, n0, 0, g.nL, one, nullptr, 1, B, st));
      if (q == K - 1 && g.grad_sigR)
        CHK(sig_corner(g.grad_sigR, Wm, nK, nK - g.nR, g.nR, one, nullptr, 1, B, st));
      if (q + 1 < K) {
        const int n1 = n[q + 1];
        // grad_hu[q] = sum_b (G_{q+1,K-1} D_q)^H ;  grad_hl[q] = sum_b (G_{q,K-1} D_{q+1})^H
        CHK(gemm(st, B, n1, nq, nK, M(Gc[1 - cur], nK, n1), M(D[cur], nq, nK), one, zero, Wm,
                 (int64_t)n1 * nq, nq, g.z));
        CHK(reduce(g.grad_hu + gx[q], Wm, n1, nq));
        CHK(reduce_s(g.grad_su ? g.grad_su + gx[q] : nullptr, Wm, n1, nq));
        CHK(gemm(st, B, nq, n1, nK, M(Gc[cur], nK, nq), M(D[1 - cur], n1, nK), one, zero, Wm,
                 (int64_t)nq * n1, n1, g.z));
        CHK(reduce(g.grad_hl + gx[q], Wm, nq, n1));
        CHK(reduce_s(g.grad_sl ? g.grad_sl + gx[q] : nullptr, Wm, nq, n1));
      }
      cur = 1 - cur;
    }
  }
  return hipSuccess;
}

}  // namespace negf
