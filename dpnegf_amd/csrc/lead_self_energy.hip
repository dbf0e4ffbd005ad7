// Batched Lopez-Sancho decimation + lead self-energy, FP64, one workgroup per
// (E,k) point.
//
// Replaces _surface_green_scipy_core / _surface_green_numba_core
// (dpnegf/negf/surface_green.py:102-159, 169-226) and the Sigma assembly of
// selfEnergy (surface_green.py:72-86), which the reference runs per (E,k) in a
// joblib process pool (dpnegf/negf/lead_property.py:599-683).
//
// Per iteration, as the reference (:189-204):
//   [X_a | X_b] = W \ [alpha | beta],  W = zS - eps   (solve, not W^-1 times)
//   eps   += alpha X_b + beta X_a
//   eps_s += beta X_a
//   alpha <- alpha X_a,  beta <- beta X_b
// until max_ij |alpha_ij| + |beta_ij| < stop (1e-40 in the reference, :208).
// Then g_s = (zS - eps_s)^-1, the Dyson residual of :211-220, and
//   Sigma = (e S_DL - H_DL) g_s (e S_DL^H - H_DL^H)      (:86, real-axis e)
// or the no-coupling forms of :72-80.
//
// The solve is Gauss-Jordan elimination on [W | alpha | beta] (forward stable
// like the reference's LU solve; forming W^-1 and multiplying instead loses
// the Dyson residual near band edges at small eta).  Matrices are padded to
// NMAX with an identity block in W and zero rows/columns in alpha and beta,
// which leaves the n x n results exactly as without padding.
//
// LDS: AUG = [W | alpha | beta] (NMAX x 3 NMAX); after the solve W's columns
// receive alpha again (from a copy in the workspace) and SP holds beta, so the
// products read [alpha; beta] from [AUG[:, :N]; SP] and X from AUG through the
// pivot order.  eps and the alpha copy live in the workspace, eps_s in the g_s
// output; each is read and written once per iteration.
#include "rgf_kernels.hpp"

namespace negf {

template <int NMAX, bool ORTHO, int NTH>
__global__ __launch_bounds__(NTH) void lead_self_energy_kernel(LeadArgs a) {
  typedef double R;
  constexpr int N = NMAX;
  constexpr int NW = NTH / WAVE;
  constexpr int LDG = 3 * N + 1;
  constexpr int LDP = N + 1;
  constexpr int NTILE = (N / 16) * (N / 16);
  constexpr int MAXT2 = (2 * NTILE + NW - 1) / NW;
  constexpr int MAXT1 = (NTILE + NW - 1) / NW;
  __shared__ cx<R> AUG[N * LDG];
  __shared__ cx<R> SP[N * LDP];
  __shared__ GJBScratch<R, N> gjb;
  __shared__ double red[NW];

  const int pt = blockIdx.x;
  if (pt >= a.P) return;
  const int n = a.n, nr = a.nred;
  const int sys = a.sys_idx ? a.sys_idx[pt] : 0;
  const cx<R>* H = a.lead + (int64_t)sys * a.lead_stride;
  const cx<R>* h01 = H + n * n;
  const cx<R>* S = h01 + n * n;
  const cx<R>* s01 = S + n * n;
  const cx<R>* hDL = s01 + n * n;
  const cx<R>* sDL = hDL + nr * n;
  const cx<R> e = a.e[pt];
  const cx<R> z = {e.re, e.im + a.eta};
  cx<R>* eps = a.eps + (int64_t)pt * 2 * n * n;
  cx<R>* acopy = eps + n * n;
  cx<R>* epss = a.gs + (int64_t)pt * n * n;

  auto zS = [&](int i, int k) -> cx<R> {
    if (ORTHO) return (i == k) ? z : cx<R>{0, 0};
    return z * S[i * n + k];
  };
  auto zs01 = [&](int i, int k) -> cx<R> {
    if (ORTHO) return cx<R>{0, 0};
    return z * s01[i * n + k];
  };
  // Padded copy of the n x n matrix f(i, k) into AUG[:, :N] (identity outside).
  auto stage = [&](auto f) __attribute__((always_inline)) {
    for (int x = threadIdx.x; x < N * N; x += NTH) {
      const int i = x / N, k = x % N;
      AUG[i * LDG + k] = (i < n && k < n) ? f(i, k) : cx<R>{i == k ? R(1) : R(0), 0};
    }
  };
  // alpha0 = h10 - z s10, beta0 = h01 - z s01, eps = eps_s = H   (:170-177)
  for (int x = threadIdx.x; x < N * N; x += NTH) {
    const int i = x / N, k = x % N;
    cx<R> al = {0, 0}, be = {0, 0};
    if (i < n && k < n) {
      be = h01[i * n + k] - zs01(i, k);
      // s10 = s01^H enters as z * conj(s01[k][i]); z itself is not conjugated
      al = conj(h01[k * n + i]);
      if (!ORTHO) al = al - z * conj(s01[k * n + i]);
      acopy[i * n + k] = al;
      eps[i * n + k] = H[i * n + k];
      epss[i * n + k] = H[i * n + k];
    }
    AUG[i * LDG + N + k] = al;
    AUG[i * LDG + 2 * N + k] = be;
    SP[i * LDP + k] = be;
  }

  int status = 0;
  int it = 0;
  while (true) {
    ++it;
    stage([&](int i, int k) { return zS(i, k) - eps[i * n + k]; });
    __syncthreads();
    status |= wg_gj_solve_blocked<R, NTH, N, 2 * N, LDG>(AUG, &gjb);
    for (int x = threadIdx.x; x < N * N; x += NTH) {
      const int i = x / N, k = x % N;
      AUG[i * LDG + k] = (i < n && k < n) ? acopy[i * n + k] : cx<R>{0, 0};
    }
    __syncthreads();
    // [alpha; beta] X_a:  alpha X_a is the new alpha, beta X_a enters eps and eps_s
    wg_gemm<R, NW, MAXT2>(
        2 * N, N, N,
        [&](int i, int k) { return i < N ? AUG[i * LDG + k] : SP[(i - N) * LDP + k]; },
        [&](int k, int j) { return AUG[gjb.piv[k] * LDG + N + j]; },
        [&](int i, int j, cx<R> v) {
          if (i < N) {
            AUG[i * LDG + N + j] = v;
            if (i < n && j < n) acopy[i * n + j] = v;
          } else {
            const int r = i - N;
            if (r < n && j < n) {
              eps[r * n + j] = eps[r * n + j] + v;
              epss[r * n + j] = epss[r * n + j] + v;
            }
          }
        });
    // [alpha; beta] X_b:  alpha X_b enters eps, beta X_b is the new beta.  The
    // barrier inside orders these eps updates after the ones above.
    wg_gemm<R, NW, MAXT2>(
        2 * N, N, N,
        [&](int i, int k) { return i < N ? AUG[i * LDG + k] : SP[(i - N) * LDP + k]; },
        [&](int k, int j) { return AUG[gjb.piv[k] * LDG + 2 * N + j]; },
        [&](int i, int j, cx<R> v) {
          if (i < N) {
            if (i < n && j < n) eps[i * n + j] = eps[i * n + j] + v;
          } else {
            const int r = i - N;
            AUG[r * LDG + 2 * N + j] = v;
            SP[r * LDP + j] = v;
          }
        });
    __syncthreads();
    double m = 0.0;
    for (int x = threadIdx.x; x < n * n; x += NTH) {
      const int i = x / n, k = x % n;
      m = fmax(m, cabs(AUG[i * LDG + N + k]) + cabs(AUG[i * LDG + 2 * N + k]));
    }
    m = block_max(m, red);
    if (m < a.stop) break;
    if (it >= 101) {
      status |= STATUS_LS_NOCONV;
      break;
    }
  }

  // g_s = (zS - eps_s)^-1, written over eps_s (the g_s output)   (:209)
  stage([&](int i, int k) { return zS(i, k) - epss[i * n + k]; });
  __syncthreads();
  status |= wg_gj_inverse_blocked<R, NTH, N, LDG>(
      AUG, [&](int i, int k, cx<R> v) { if (i < n && k < n) epss[i * n + k] = v; }, &gjb);
  const cx<R>* gs = epss;
  // residual: max |(zS - H - (z s01 - h01) g_s (z s10 - h10)) g_s - I|   (:211-212)
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return zs01(i, k) - h01[i * n + k]; },
      [&](int k, int j) { return gs[k * n + j]; },
      [&](int i, int j, cx<R> v) { SP[i * LDP + j] = v; });
  __syncthreads();
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return SP[i * LDP + k]; },
      [&](int k, int j) {
        cx<R> s10 = ORTHO ? cx<R>{0, 0} : z * conj(s01[j * n + k]);
        return s10 - conj(h01[j * n + k]);
      },
      [&](int i, int j, cx<R> v) { AUG[i * LDG + N + j] = zS(i, j) - H[i * n + j] - v; });
  __syncthreads();
  double rm = 0.0;
  wg_gemm<R, NW, MAXT1>(
      n, n, n, [&](int i, int k) { return AUG[i * LDG + N + k]; },
      [&](int k, int j) { return gs[k * n + j]; },
      [&](int i, int j, cx<R> v) {
        if (i == j) v.re -= 1.0;
        rm = fmax(rm, cabs(v));
      });
  rm = block_max(rm, red);
  if (rm >= 3.0e-5) status |= STATUS_LS_RESID_FAIL;
  else if (rm > 1.0e-8) status |= STATUS_LS_RESID_WARN;

  cx<R>* sig = a.sigma + (int64_t)pt * nr * nr;
  if (a.mode == 0) {
    // Y = g_s (e S_DL^H - H_DL^H) [n x nr];  Sigma = (e S_DL - H_DL) Y
    wg_gemm<R, NW, MAXT1>(
        n, nr, n, [&](int i, int k) { return gs[i * n + k]; },
        [&](int k, int j) { return e * conj(sDL[j * n + k]) - conj(hDL[j * n + k]); },
        [&](int i, int j, cx<R> v) { AUG[i * LDG + N + j] = v; });
    __syncthreads();
    wg_gemm<R, NW, MAXT1>(
        nr, nr, n, [&](int i, int k) { return e * sDL[i * n + k] - hDL[i * n + k]; },
        [&](int k, int j) { return AUG[k * LDG + N + j]; },
        [&](int i, int j, cx<R> v) { sig[i * nr + j] = v; });
  } else {
    // g_s^-1 by a second inverse; mode 1: (e S - H) - g_s^-1, mode 2: g_s^-1
    const int mode = a.mode;
    stage([&](int i, int k) { return gs[i * n + k]; });
    __syncthreads();
    status |= wg_gj_inverse_blocked<R, NTH, N, LDG>(
        AUG,
        [&](int i, int k, cx<R> v) {
          if (i < n && k < n) {
            if (mode == 1) sig[i * n + k] = e * S[i * n + k] - H[i * n + k] - v;
            else sig[i * n + k] = v;
          }
        },
        &gjb);
  }
  if (threadIdx.x == 0) {
    if (a.iters) a.iters[pt] = it;
    if (a.status) a.status[pt] = status;
    if (a.resid) a.resid[pt] = rm;
  }
}

#ifndef LS_THREADS_48
#define LS_THREADS_48 512
#endif
// Threads per workgroup: n = 48 leaves one workgroup per CU (LDS), so it runs
// 8 waves rather than 4 (measured faster with identical results).
template <int NMAX> constexpr int ls_threads() { return NMAX == 48 ? LS_THREADS_48 : 256; }

template <int NMAX>
static hipError_t launch_ls(const LeadArgs& a, bool ortho, hipStream_t st) {
  constexpr int NTH = ls_threads<NMAX>();
  dim3 grid(a.P), block(NTH);
  if (ortho)
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, true, NTH>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, false, NTH>), grid, block, 0, st, a);
  return hipGetLastError();
}

int64_t lead_self_energy_workspace(int n, int P) { return 2 * (int64_t)P * n * n; }

hipError_t launch_lead_self_energy(const LeadArgs& a, bool ortho, hipStream_t st) {
  const int m = a.n > a.nred ? a.n : a.nred;
  if (m <= 16) return launch_ls<16>(a, ortho, st);
  if (m <= 32) return launch_ls<32>(a, ortho, st);
  if (m <= 48) return launch_ls<48>(a, ortho, st);
  return hipErrorInvalidValue;
}

}  // namespace negf
