// Batched Lopez-Sancho decimation + lead self-energy, FP64, one workgroup per
// (E,k) point.
//
// Replaces _surface_green_scipy_core / _surface_green_numba_core
// (dpnegf/negf/surface_green.py:102-159, 169-226) and the Sigma assembly of
// selfEnergy (surface_green.py:72-86), which the reference runs per (E,k) in a
// joblib process pool (dpnegf/negf/lead_property.py:599-683).
//
// Per iteration, with W = zS - eps inverted by Gauss-Jordan (the reference
// solves W X = [alpha | beta] instead; same products, other association):
//   Ai = alpha W^-1, Bi = beta W^-1
//   eps   += Ai beta + Bi alpha        (one product with k = 2n)
//   eps_s += Bi alpha
//   alpha <- Ai alpha,  beta <- Bi beta
// until max_ij |alpha_ij| + |beta_ij| < stop (1e-40 in the reference, :208).
// Then g_s = (zS - eps_s)^-1, the Dyson residual of :211-220, and
//   Sigma = (e S_DL - H_DL) g_s (e S_DL^H - H_DL^H)      (:86, real-axis e)
// or the no-coupling forms of :72-80.
//
// LDS: AB holds [beta; alpha] (2n x n), AIBI holds W^-1 and then [Ai | Bi]
// (n x 2n).  eps and eps_s live in global memory (eps in the workspace, eps_s
// in the g_s output), read/written once per iteration.
#include "rgf_kernels.hpp"

namespace negf {

template <int NMAX, bool ORTHO>
__global__ __launch_bounds__(256) void lead_self_energy_kernel(LeadArgs a) {
  typedef double R;
  constexpr int LD = NMAX + 1;
  constexpr int LD2 = 2 * NMAX + 1;
  constexpr int NT = (NMAX / 16) * (NMAX / 16);
  constexpr int MAXT2 = (2 * NT + 3) / 4;
  constexpr int MAXT1 = (NT + 3) / 4;
  constexpr int MAXR = gj_maxr(NMAX, 256);
  __shared__ cx<R> AB[2 * NMAX * LD];
  __shared__ cx<R> AIBI[NMAX * LD2];
  __shared__ GJScratch<R, NMAX> gjs;
  __shared__ double red[8];

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
  cx<R>* eps = a.eps + (int64_t)pt * n * n;
  cx<R>* epss = a.gs + (int64_t)pt * n * n;
  cx<R>* beta = AB;
  cx<R>* alpha = AB + n * LD;

  auto zS = [&](int i, int k) -> cx<R> {
    if (ORTHO) return (i == k) ? z : cx<R>{0, 0};
    return z * S[i * n + k];
  };
  auto zs01 = [&](int i, int k) -> cx<R> {
    if (ORTHO) return cx<R>{0, 0};
    return z * s01[i * n + k];
  };
  // alpha0 = h10 - z s10, beta0 = h01 - z s01, eps = eps_s = H   (:170-177)
  for (int x = threadIdx.x; x < n * n; x += 256) {
    const int i = x / n, k = x % n;
    beta[i * LD + k] = h01[i * n + k] - zs01(i, k);
    // s10 = s01^H enters as z * conj(s01[k][i]); z itself is not conjugated
    cx<R> al = conj(h01[k * n + i]);
    if (!ORTHO) al = al - z * conj(s01[k * n + i]);
    alpha[i * LD + k] = al;
    eps[x] = H[x];
    epss[x] = H[x];
  }
  __syncthreads();

  int status = 0;
  int it = 0;
  while (true) {
    ++it;
    status |= wg_gj_inverse<R, 256, NMAX, MAXR>(
        n, [&](int i, int k) { return zS(i, k) - eps[i * n + k]; },
        [&](int i, int k, cx<R> v) { AIBI[i * LD2 + k] = v; }, &gjs);
    __syncthreads();
    // [Bi; Ai] = [beta; alpha] W^-1, stored as [Ai | Bi]
    wg_gemm<R, 4, MAXT2>(
        2 * n, n, n, [&](int i, int k) { return AB[i * LD + k]; },
        [&](int k, int j) { return AIBI[k * LD2 + j]; },
        [&](int r, int j, cx<R> v) {
          if (r < n) AIBI[r * LD2 + n + j] = v;
          else AIBI[(r - n) * LD2 + j] = v;
        });
    __syncthreads();
    // eps += [Ai | Bi] [beta; alpha]
    wg_gemm<R, 4, MAXT1>(
        n, n, 2 * n, [&](int i, int k) { return AIBI[i * LD2 + k]; },
        [&](int k, int j) { return AB[k * LD + j]; },
        [&](int i, int j, cx<R> v) { eps[i * n + j] = eps[i * n + j] + v; });
    // eps_s += Bi alpha
    wg_gemm<R, 4, MAXT1>(
        n, n, n, [&](int i, int k) { return AIBI[i * LD2 + n + k]; },
        [&](int k, int j) { return alpha[k * LD + j]; },
        [&](int i, int j, cx<R> v) { epss[i * n + j] = epss[i * n + j] + v; });
    // alpha <- Ai alpha
    wg_gemm<R, 4, MAXT1>(
        n, n, n, [&](int i, int k) { return AIBI[i * LD2 + k]; },
        [&](int k, int j) { return alpha[k * LD + j]; },
        [&](int i, int j, cx<R> v) { alpha[i * LD + j] = v; });
    // beta <- Bi beta
    wg_gemm<R, 4, MAXT1>(
        n, n, n, [&](int i, int k) { return AIBI[i * LD2 + n + k]; },
        [&](int k, int j) { return beta[k * LD + j]; },
        [&](int i, int j, cx<R> v) { beta[i * LD + j] = v; });
    __syncthreads();
    double m = 0.0;
    for (int x = threadIdx.x; x < n * n; x += 256) {
      const int i = x / n, k = x % n;
      m = fmax(m, cabs(alpha[i * LD + k]) + cabs(beta[i * LD + k]));
    }
    m = block_max(m, red);
    if (m < a.stop) break;
    if (it >= 101) {
      status |= STATUS_LS_NOCONV;
      break;
    }
  }

  // g_s = (zS - eps_s)^-1, written over eps_s (the g_s output)
  status |= wg_gj_inverse<R, 256, NMAX, MAXR>(
      n, [&](int i, int k) { return zS(i, k) - epss[i * n + k]; },
      [&](int i, int k, cx<R> v) { epss[i * n + k] = v; }, &gjs);
  __syncthreads();
  const cx<R>* gs = epss;
  // residual: max |(zS - H - (z s01 - h01) g_s (z s10 - h10)) g_s - I|   (:211-212)
  wg_gemm<R, 4, MAXT1>(
      n, n, n, [&](int i, int k) { return zs01(i, k) - h01[i * n + k]; },
      [&](int k, int j) { return gs[k * n + j]; },
      [&](int i, int j, cx<R> v) { AIBI[i * LD2 + j] = v; });
  __syncthreads();
  wg_gemm<R, 4, MAXT1>(
      n, n, n, [&](int i, int k) { return AIBI[i * LD2 + k]; },
      [&](int k, int j) {
        cx<R> s10 = ORTHO ? cx<R>{0, 0} : z * conj(s01[j * n + k]);
        return s10 - conj(h01[j * n + k]);
      },
      [&](int i, int j, cx<R> v) { AB[i * LD + j] = zS(i, j) - H[i * n + j] - v; });
  __syncthreads();
  double rm = 0.0;
  wg_gemm<R, 4, MAXT1>(
      n, n, n, [&](int i, int k) { return AB[i * LD + k]; },
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
    wg_gemm<R, 4, MAXT1>(
        n, nr, n, [&](int i, int k) { return gs[i * n + k]; },
        [&](int k, int j) { return e * conj(sDL[j * n + k]) - conj(hDL[j * n + k]); },
        [&](int i, int j, cx<R> v) { AIBI[i * LD2 + j] = v; });
    __syncthreads();
    wg_gemm<R, 4, MAXT1>(
        nr, nr, n, [&](int i, int k) { return e * sDL[i * n + k] - hDL[i * n + k]; },
        [&](int k, int j) { return AIBI[k * LD2 + j]; },
        [&](int i, int j, cx<R> v) { sig[i * nr + j] = v; });
  } else {
    // g_s^-1 by a second inverse; mode 1: (e S - H) - g_s^-1, mode 2: g_s^-1
    const int mode = a.mode;
    status |= wg_gj_inverse<R, 256, NMAX, MAXR>(
        n, [&](int i, int k) { return gs[i * n + k]; },
        [&](int i, int k, cx<R> v) {
          if (mode == 1) sig[i * n + k] = e * S[i * n + k] - H[i * n + k] - v;
          else sig[i * n + k] = v;
        },
        &gjs);
  }
  if (threadIdx.x == 0) {
    if (a.iters) a.iters[pt] = it;
    if (a.status) a.status[pt] = status;
    if (a.resid) a.resid[pt] = rm;
  }
}

template <int NMAX>
static hipError_t launch_ls(const LeadArgs& a, bool ortho, hipStream_t st) {
  dim3 grid(a.P), block(256);
  if (ortho)
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, true>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((lead_self_energy_kernel<NMAX, false>), grid, block, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_lead_self_energy(const LeadArgs& a, bool ortho, hipStream_t st) {
  const int m = a.n > a.nred ? a.n : a.nred;
  if (m <= 16) return launch_ls<16>(a, ortho, st);
  if (m <= 32) return launch_ls<32>(a, ortho, st);
  if (m <= 48) return launch_ls<48>(a, ortho, st);
  return hipErrorInvalidValue;
}

}  // namespace negf
