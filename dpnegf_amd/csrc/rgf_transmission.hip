// Fused transmission sweep: one workgroup per (E,k) point, blocks in LDS.
//
// Replaces the forward loop of recursive_gf_cal (dpnegf/negf/recursive_green_cal.py
// :198-207), the g_trans product of its backward loop (:223-228) and the trace of
// DeviceProperty._cal_tc_ (dpnegf/negf/device_property.py:383-404).  Instead of
// storing every g_q and walking back, the sweep carries S_q = G^L_{0,q} U_q,
// which needs no product of its own: with Y_q = g_q U_q, formed anyway for the
// Dyson term,
//
//   g_0 = (-M_0)^{-1},  S_{-1} = I
//   Y_q = g_q U_q
//   S_q = S_{q-1} Y_q,  g_{q+1} = (-M_{q+1} - L_q Y_q)^{-1}   (one stacked product)
//
// so S_{K-2} g_{K-1} = g_0 U_0 g_1 ... U_{K-2} g_{K-1} = G_{0,K-1}, the product
// the reference builds right-to-left: three block products per block instead
// of the four of carrying G^L_{0,q} itself (P U_q and P g_{q+1}).  Only the
// corner Q = G_{0,K-1}[0:nL, K-1 corner] is formed at the end, and
// T = Re Tr[Gamma_L Q Gamma_R Q^dagger], Gamma = i(Sigma - Sigma^dagger) formed
// on the fly from the self-energies.
//
// Arena layout of one system (complex128, row-major), for q = 0..K-1:
//   hd_q [n_q x n_q], sd_q [n_q x n_q], and for q < K-1:
//   hu_q [n_q x n_q+1], su_q, hl_q [n_q+1 x n_q], sl_q.
#include "rgf_kernels.hpp"

namespace negf {

// Residual max |I - A X0| above which the FP32 block inverse X0 is not
// refined but replaced by the FP64 inverse (one Newton step squares the
// relative error: 1e-4 -> about 1e-8 before the FP64 rounding floor).
constexpr double REFINE_TOL = 1e-4;

// Where the carried S lives: in LDS next to g, or (SG) outside LDS, so that g
// alone takes LDS.  FP64 FIXED at NMAX = 48: 79 KB of LDS holds two workgroups
// per CU, 41 KB four (three-wave workgroups, below), and the serial pivot
// chains of more points overlap.  SG keeps S transposed in registers (the
// kernel body says how); held in the caller's work buffer instead it moved
// about 80 GB per launch through the memory fabric at four workgroups per CU.
// W = 48, K = 128 ribbon sweep: 71.1 ms (S in LDS, two workgroups per CU) ->
// 64.1 (work buffer, three) -> 58.5 (work buffer, four) -> 55.3 ms (registers,
// four, g^-1 left in pivot order).  At NMAX = 64, S in the work buffer (133 ->
// 71 KB, two workgroups per CU instead of one) measured 31.21 against 31.24 ms
// on the W = 64, K = 64 sweep: S stays in LDS there.
template <typename R, int NMAX, bool FIXED, bool REFINE>
constexpr bool rgf_s_out_of_lds() {
  return FIXED && !REFINE && sizeof(R) == 8 && NMAX == 48;
}
// Threads per workgroup.  With S out of LDS, three waves and 40.9 KB of LDS
// (g, the panel multipliers and 16-bit pivot indices) let four workgroups
// share a CU: four pivot chains overlap instead of three, and the 9-tile
// products split 3/3/3 over the waves (with four waves the slowest also had
// three).
template <typename R, int NMAX, bool FIXED, bool REFINE>
constexpr int rgf_threads() {
  return rgf_s_out_of_lds<R, NMAX, FIXED, REFINE>() ? 192 : 256;
}
// waves per SIMD the register budget is held to (workgroups per CU x waves / 4)
template <typename R, int NMAX, bool FIXED, bool REFINE, int NTH>
constexpr int rgf_waves_per_eu() {
  if (rgf_s_out_of_lds<R, NMAX, FIXED, REFINE>()) return 4 * (NTH / 64) / 4;
  return (NMAX <= 48 && !REFINE ? 2 : 1) * NTH / 256;
}

// FIXED: every block has exactly NMAX rows (compile-time shapes for the inverse
// and the products).  Up to NMAX = 48 two workgroups share a CU (LDS), so the
// register budget is held to two waves per SIMD (four three-wave workgroups
// and three waves per SIMD with S in global memory).
// REFINE (with R = double): every block inverse is computed in FP32 and
// refined by one FP64 Newton step, X = X0 + X0 (I - A X0); a block whose FP32
// residual exceeds REFINE_TOL is inverted in FP64 instead (status bit
// STATUS_REFINE_FALLBACK).  Products and the carried P stay FP64.
// NTH: threads per workgroup (NTH / 64 waves share every product and update).
template <typename R, int NMAX, bool ORTHO, bool FIXED, bool REFINE = false, int NTH = 256>
__global__ __launch_bounds__(NTH)
__attribute__((amdgpu_waves_per_eu(rgf_waves_per_eu<R, NMAX, FIXED, REFINE, NTH>())))
void rgf_transmission_kernel(RgfArgs a) {
  constexpr bool SG = rgf_s_out_of_lds<R, NMAX, FIXED, REFINE>();
  static_assert(!REFINE || (sizeof(R) == 8 && NMAX <= 48), "refinement: FP64 kernel, NMAX <= 48");
  constexpr int NW = NTH / WAVE;
  constexpr int LD = NMAX + 1;
  constexpr int NT = (NMAX / 16) * (NMAX / 16);
  constexpr int MAXT_STACK = (2 * NT + NW - 1) / NW;
  constexpr int MAXT_SQ = (NT + NW - 1) / NW;
  constexpr int MAXR = gj_maxr(NMAX, NTH);
  __shared__ cx<R> PG[(SG ? 1 : 2) * NMAX * LD];  // [S; g], or g alone
  constexpr int SLD = SG ? NMAX : LD;             // leading dimension of S
  __shared__ GJScratch<R, NMAX> gjs;
  __shared__ GJBScratch<R, NMAX> gjb;
  __shared__ double red[NW];
  // FP32 copy / inverse of the block and the FP64 residual (refined variant only)
  constexpr int NR = REFINE ? NMAX * LD : 1;
  __shared__ cx<float> XF[NR];
  __shared__ cx<double> RS[NR];
  __shared__ GJScratch<float, REFINE ? NMAX : 16> gjsf;
  __shared__ GJBScratch<float, REFINE ? NMAX : 16> gjbf;
  int refine_fallback = 0;

  const int pt = blockIdx.x;
  if (pt >= a.P) return;  // uniform per workgroup
  const int sys = a.sys_idx ? a.sys_idx[pt] : 0;
  if (sys < 0 || sys >= a.nsys) {  // uniform per workgroup, before any barrier
    if (threadIdx.x == 0) {
      a.T[pt] = __builtin_nan("");
      if (a.status) a.status[pt] = STATUS_BAD_INDEX;
    }
    return;
  }
  const cx<double>* base = a.arena + (int64_t)sys * a.sys_stride;
  const cx<double> zd = a.z[pt];
  const int K = a.K;
  const int n0 = a.bsz[0];
  const int nlast = a.bsz[K - 1];
  const cx<double>* sL = a.sigL + (int64_t)pt * a.nL * a.nL;
  const cx<double>* sR = a.sigR + (int64_t)pt * a.nR * a.nR;
  const int nL = a.nL, nR = a.nR;

  // S_{q-1}, n0 rows (FIXED: n0 = NMAX, so point pt's slot is work + pt n0 nmax)
  cx<R>* S = SG ? (cx<R>*)a.work + (int64_t)pt * NMAX * NMAX : PG;
  cx<R>* g = SG ? PG : PG + n0 * LD;
  // SG: the carried S is held transposed in registers.  Wave w keeps column
  // block w of S^T (rows 16 I + lg + 4 r of tile I in element r, the
  // accumulator layout of the f64 matrix-core instruction), which is exactly
  // the B operand of S^T <- Y^T S^T for k-step 4 I + r: the product needs no
  // data movement, and S never reaches LDS or memory (the work buffer holds
  // only the final corner Q).  Held in global memory instead, S moved about
  // 80 GB per launch through the memory fabric at four workgroups per CU.
  typedef typename mfma<R>::v4 v4;
  constexpr int NB = NMAX / 16;
  static_assert(!SG || NW == NB, "register S^T: one wave per 16-column block");
  const int li = threadIdx.x & 15, lg = (threadIdx.x & (WAVE - 1)) >> 4;
  const int wv = threadIdx.x / WAVE;
  v4 str[SG ? NB : 1], sti[SG ? NB : 1];

  // M(i,k) = H - z S (+ Sigma corners), in double before rounding to R
  auto Mval = [&](const cx<double>* h, const cx<double>* s, int nb, int i, int k, bool left,
                  bool right) -> cx<double> {
    cx<double> v = h[i * nb + k];
    if (ORTHO) {
      if (i == k) v = v - zd;
    } else {
      v = v - zd * s[i * nb + k];
    }
    if (left && i < nL && k < nL) v = v + sL[i * nL + k];
    if (right) {
      const int o = nb - nR;
      if (i >= o && k >= o) v = v + sR[(i - o) * nR + (k - o)];
    }
    return v;
  };

  const cx<double>* hd0 = base;
  const cx<double>* sd0 = base + n0 * n0;
  // FIXED: blocked inverse of g in place; otherwise the register-resident one
  auto inverse_full = [&](int n) __attribute__((always_inline)) -> int {
    if constexpr (FIXED)
      return wg_gj_inverse_blocked<R, NTH, NMAX, LD>(
          g, [&](int i, int k, cx<R> v) { g[i * LD + k] = v; }, &gjb);
    else
      return wg_gj_inverse<R, NTH, NMAX, MAXR>(
          n, [&](int i, int k) { return g[i * LD + k]; },
          [&](int i, int k, cx<R> v) { g[i * LD + k] = v; }, &gjs);
  };
  // g <- g^-1 in place; every thread calls it, and it ends with a barrier.
  // SG: g^-1 stays in pivot order, g^-1[i][k] = g[piv[i]][pinv[k]], and the
  // consumers (the product Y = g U and the final corner) read through the maps.
  auto inverse_g = [&](int n) __attribute__((always_inline)) -> int {
    if constexpr (SG) {
      return wg_gj_inverse_pivoted<R, NTH, NMAX, LD>(g, &gjb);
    } else if constexpr (!REFINE) {
      return inverse_full(n);
    } else {
      for (int e = threadIdx.x; e < n * n; e += NTH) {
        const int i = e / n, k = e % n;
        XF[i * LD + k] = cvt<float>(g[i * LD + k]);
      }
      __syncthreads();
      int sg;
      if constexpr (FIXED)
        sg = wg_gj_inverse_blocked<float, NTH, NMAX, LD>(
            XF, [&](int i, int k, cx<float> v) { XF[i * LD + k] = v; }, &gjbf);
      else
        sg = wg_gj_inverse<float, NTH, NMAX, gj_maxr(NMAX, NTH)>(
            n, [&](int i, int k) { return XF[i * LD + k]; },
            [&](int i, int k, cx<float> v) { XF[i * LD + k] = v; }, &gjsf);
      __syncthreads();
      // RS = I - A X0 in FP64 (A still in g), and its largest entry
      double rmax = 0.0;
      auto ra = [&](int i, int k) { return g[i * LD + k]; };
      auto rb = [&](int k, int j) { return cvt<double>(XF[k * LD + j]); };
      auto re = [&](int i, int j, cx<double> v) {
        const cx<double> r = cx<double>{(i == j ? 1.0 : 0.0) - v.re, -v.im};
        RS[i * LD + j] = r;
        rmax = fmax(rmax, abs1(r));
      };
      if constexpr (FIXED) wg_gemm_fixed<double, NW, NMAX, NMAX, NMAX, false, false>(ra, rb, re);
      else wg_gemm<double, NW, MAXT_SQ>(n, n, n, ra, rb, re);
      rmax = block_max(rmax, red);  // ends with the residual visible to all
      if (sg || !(rmax < REFINE_TOL)) {  // uniform: singular, NaN or too far for one step
        refine_fallback = 1;
        return inverse_full(n);
      }
      // g <- X0 + X0 RS
      auto xa = [&](int i, int k) { return cvt<double>(XF[i * LD + k]); };
      auto xb = [&](int k, int j) { return RS[k * LD + j]; };
      auto xe = [&](int i, int j, cx<double> v) { g[i * LD + j] = cvt<double>(XF[i * LD + j]) + v; };
      if constexpr (FIXED) wg_gemm_fixed<double, NW, NMAX, NMAX, NMAX, false, false>(xa, xb, xe);
      else wg_gemm<double, NW, MAXT_SQ>(n, n, n, xa, xb, xe);
      __syncthreads();
      return 0;
    }
  };
  for (int e = threadIdx.x; e < n0 * n0; e += NTH) {
    const int i = e / n0, k = e % n0;
    g[i * LD + k] = cvt<R>(-Mval(hd0, sd0, n0, i, k, true, K == 1));
  }
  __syncthreads();
  int sing = inverse_g(n0);
  if constexpr (SG) {
#pragma unroll
    for (int I = 0; I < NB; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        str[I][r] = (16 * I + lg + 4 * r == 16 * wv + li) ? R(1) : R(0);
        sti[I][r] = R(0);
      }
  } else {
    for (int e = threadIdx.x; e < n0 * n0; e += NTH) {
      const int i = e / n0, k = e % n0;
      S[i * SLD + k] = cx<R>{i == k ? R(1) : R(0), R(0)};
    }
    __syncthreads();
  }

  int64_t off = 0;
  for (int q = 0; q < K - 1; ++q) {
    const int nq = a.bsz[q], n1 = a.bsz[q + 1];
    const cx<double>* hu = base + off + 2 * (int64_t)nq * nq;
    const cx<double>* su = hu + (int64_t)nq * n1;
    const cx<double>* hl = su + (int64_t)nq * n1;
    const cx<double>* sl = hl + (int64_t)n1 * nq;
    const int64_t off1 = off + 2 * (int64_t)nq * nq + 4 * (int64_t)nq * n1;
    const cx<double>* hd1 = base + off1;
    const cx<double>* sd1 = hd1 + (int64_t)n1 * n1;
    const bool last = (q + 1 == K - 1);

    // Y = g_q (U_q - z SU_q), in place in g.  SG, with g_q in pivot order:
    // Y[i] = Z[piv[i]] for Z = g (U with row m taken from row piv[m]), so U is
    // read through piv and Z's row r is stored as Y's row pinv[r].
    auto a1 = [&](int i, int k) { return g[i * LD + k]; };
    auto b1 = [&](int k, int j) {
      const int kr = SG ? (int)gjb.piv[k] : k;
      cx<double> u = hu[kr * n1 + j];
      if (!ORTHO) u = u - zd * su[kr * n1 + j];
      return cvt<R>(u);
    };
    auto e1 = [&](int i, int j, cx<R> v) {
      const int ir = SG ? (int)gjb.pinv[i] : i;
      g[ir * LD + j] = v;
    };
    if constexpr (FIXED) wg_gemm_fixed<R, NW, NMAX, NMAX, NMAX, false, true>(a1, b1, e1);
    else wg_gemm<R, NW, MAXT_SQ>(nq, n1, nq, a1, b1, e1);
    __syncthreads();
    // [S; g] <- [S Y; -M_{q+1} - (L_q - z SL_q) Y]: rows below n0 read S from
    // LDS, the others L_q through L2 (a 16-row tile never mixes the two when
    // the blocks are multiples of 16)
    auto a2 = [&](int i, int k) {
      if (i < n0) return S[i * SLD + k];
      cx<double> l = hl[(i - n0) * nq + k];
      if (!ORTHO) l = l - zd * sl[(i - n0) * nq + k];
      return cvt<R>(l);
    };
    auto b2 = [&](int k, int j) { return g[k * LD + j]; };
    auto e2 = [&](int i, int j, cx<R> v) {
      if (i < n0) S[i * SLD + j] = v;
      else
        g[(i - n0) * LD + j] =
            cvt<R>(-(cvt<double>(v) + Mval(hd1, sd1, n1, i - n0, j, false, last)));
    };
    // every tile's A operand is loaded one k-step ahead (prefetching only the
    // L_q tiles measured 78.8 against 71.1 ms at NMAX 48, with 244 B of
    // register spills per lane against 64 B)
    if constexpr (SG) {
      // S^T <- Y^T S^T in registers (A operand Y^T read from LDS), then
      // g <- -M_{q+1} - L_q Y, whose epilogue barrier follows every read of Y
      v4 nr[NB], ni[NB];
#pragma unroll
      for (int I = 0; I < NB; ++I) {
        nr[I] = v4{0, 0, 0, 0};
        ni[I] = v4{0, 0, 0, 0};
      }
      // one base address; k-step and tile offsets are immediates
      const cx<R>* yb = g + lg * LD + li;
#pragma unroll
      for (int kk = 0; kk < NMAX / 4; ++kk) {
        const R br = str[kk >> 2][kk & 3], bi = sti[kk >> 2][kk & 3];
#pragma unroll
        for (int I = 0; I < NB; ++I) {
          const cx<R> y = yb[4 * kk * LD + 16 * I];
          nr[I] = mfma<R>::op(y.re, br, nr[I]);
          nr[I] = mfma<R>::op(-y.im, bi, nr[I]);
          ni[I] = mfma<R>::op(y.re, bi, ni[I]);
          ni[I] = mfma<R>::op(y.im, br, ni[I]);
        }
      }
#pragma unroll
      for (int I = 0; I < NB; ++I) {
        str[I] = nr[I];
        sti[I] = ni[I];
      }
      wg_gemm_fixed<R, NW, NMAX, NMAX, NMAX, true, false, true, 1>(
          [&](int i, int k) { return a2(i + n0, k); }, b2,
          [&](int i, int j, cx<R> v) { e2(i + n0, j, v); });
    } else if constexpr (FIXED) wg_gemm_fixed<R, NW, 2 * NMAX, NMAX, NMAX, true, false>(a2, b2, e2);
    else wg_gemm<R, NW, MAXT_STACK>(n0 + n1, n1, nq, a2, b2, e2);
    __syncthreads();
    sing |= inverse_g(n1);  // ends with a workgroup barrier
    off = off1;
  }

  // Q = (S_{K-2} g_{K-1})[0:nL, c0:c0+nR] = G_{0,K-1} corner, in place in S
  // (S = I for K = 1); then
  // T = Re sum_{a,b} GammaL[a][b] (Q GammaR Q^dagger)[b][a]
  const int c0 = nlast - nR;
  if constexpr (SG) {
    // Q^T = (g[:, c0:c0+nR])^T S^T: tile (J, wv) from the registers, written
    // to the work buffer as Q
#pragma unroll
    for (int J = 0; J < NB; ++J) {
      v4 qr = v4{0, 0, 0, 0}, qi = v4{0, 0, 0, 0};
      const int jc = 16 * J + li;
#pragma unroll
      for (int kk = 0; kk < NMAX / 4; ++kk) {
        const R br = str[kk >> 2][kk & 3], bi = sti[kk >> 2][kk & 3];
        // g_{K-1}[k][c0 + jc] = g[piv[k]][pinv[c0 + jc]] (pivot order)
        const cx<R> y = jc < nR ? g[(int)gjb.piv[4 * kk + lg] * LD + (int)gjb.pinv[c0 + jc]]
                                : cx<R>{0, 0};
        qr = mfma<R>::op(y.re, br, qr);
        qr = mfma<R>::op(-y.im, bi, qr);
        qi = mfma<R>::op(y.re, bi, qi);
        qi = mfma<R>::op(y.im, br, qi);
      }
      const int i = 16 * wv + li;  // row of Q
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 16 * J + lg + 4 * r;  // column of Q
        if (i < nL && j < nR) S[i * SLD + j] = cx<R>{qr[r], qi[r]};
      }
    }
  } else {
    wg_gemm<R, NW, MAXT_SQ>(
        nL, nR, nlast, [&](int i, int k) { return S[i * SLD + k]; },
        [&](int k, int j) { return g[k * LD + c0 + j]; },
        [&](int i, int j, cx<R> v) { S[i * SLD + j] = v; });
  }
  __syncthreads();
  cx<R>* V = g;
  wg_gemm<R, NW, MAXT_SQ>(
      nL, nR, nR, [&](int i, int k) { return S[i * SLD + k]; },
      [&](int k, int j) { return cvt<R>(gamma_of(sR, nR, k, j)); },
      [&](int i, int j, cx<R> v) { V[i * LD + j] = v; });
  __syncthreads();
  double acc = 0.0;
  wg_gemm<R, NW, MAXT_SQ>(
      nL, nL, nR, [&](int i, int k) { return V[i * LD + k]; },
      [&](int k, int j) { return conj(S[j * SLD + k]); },
      [&](int i, int j, cx<R> x) {
        const cx<double> gl = gamma_of(sL, nL, j, i);
        acc += gl.re * (double)x.re - gl.im * (double)x.im;
      });
  const double T = block_sum(acc, red);
  if (threadIdx.x == 0) {
    a.T[pt] = T;
    if (a.status)
      a.status[pt] = (sing ? STATUS_SINGULAR : 0) | (refine_fallback ? STATUS_REFINE_FALLBACK : 0);
  }
}

template <typename R, int NMAX, bool FIXED, bool REFINE = false, int NTH = 256>
static hipError_t launch_nmax(const RgfArgs& a, bool ortho, hipStream_t st) {
  dim3 grid(a.P), block(NTH);
  if (ortho)
    hipLaunchKernelGGL((rgf_transmission_kernel<R, NMAX, true, FIXED, REFINE, NTH>), grid, block, 0,
                       st, a);
  else
    hipLaunchKernelGGL((rgf_transmission_kernel<R, NMAX, false, FIXED, REFINE, NTH>), grid, block, 0,
                       st, a);
  return hipGetLastError();
}

// 256 threads: at NMAX = 48, 512 threads (4 waves per SIMD, 128 registers with
// spills) measured 87.5 ms against 76.5 ms on the ribbon W = 48, K = 128 sweep.
template <typename R, int NMAX>
static hipError_t launch_sel(const RgfArgs& a, bool uniform, int nmax, bool ortho, hipStream_t st) {
  if (uniform && nmax == NMAX)
    return launch_nmax<R, NMAX, true, false, rgf_threads<R, NMAX, true, false>()>(a, ortho, st);
  return launch_nmax<R, NMAX, false>(a, ortho, st);
}

// uniform: every block has nmax rows
template <typename R>
hipError_t launch_rgf_transmission(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                   hipStream_t st) {
  if (nmax <= 16) return launch_sel<R, 16>(a, uniform, nmax, ortho, st);
  if (nmax <= 32) return launch_sel<R, 32>(a, uniform, nmax, ortho, st);
  if (nmax <= 48) return launch_sel<R, 48>(a, uniform, nmax, ortho, st);
  if (nmax <= 64) return launch_sel<R, 64>(a, uniform, nmax, ortho, st);
  return hipErrorInvalidValue;
}

template hipError_t launch_rgf_transmission<double>(const RgfArgs&, int, bool, bool, hipStream_t);

hipError_t launch_rgf_transmission_mixed(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                         hipStream_t st) {
  return launch_rgf_transmission<float>(a, nmax, uniform, ortho, st);
}

template <int NMAX>
static hipError_t launch_refined_sel(const RgfArgs& a, bool uniform, int nmax, bool ortho,
                                     hipStream_t st) {
  if (uniform && nmax == NMAX) return launch_nmax<double, NMAX, true, true>(a, ortho, st);
  return launch_nmax<double, NMAX, false, true>(a, ortho, st);
}

hipError_t launch_rgf_transmission_refined(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                           hipStream_t st) {
  if (nmax <= 16) return launch_refined_sel<16>(a, uniform, nmax, ortho, st);
  if (nmax <= 32) return launch_refined_sel<32>(a, uniform, nmax, ortho, st);
  if (nmax <= 48) return launch_refined_sel<48>(a, uniform, nmax, ortho, st);
  return hipErrorInvalidValue;
}

}  // namespace negf
