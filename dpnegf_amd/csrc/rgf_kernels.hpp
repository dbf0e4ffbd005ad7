// Kernel argument blocks and launchers shared by the C ABI and the kernels.
#pragma once
#include "wg_linalg.hpp"

namespace negf {

enum : int32_t {
  STATUS_SINGULAR = 1,       // exactly-zero pivot in a block inverse
  STATUS_LS_NOCONV = 2,      // Lopez-Sancho: > 100 iterations
  STATUS_LS_RESID_FAIL = 4,  // Lopez-Sancho: Dyson residual >= 3e-5
  STATUS_LS_RESID_WARN = 8,  // Lopez-Sancho: Dyson residual > 1e-8
  STATUS_REFINE_FALLBACK = 16,  // mixed precision: refinement did not converge
  STATUS_BAD_INDEX = 32,     // sys_idx[p] outside [0, nsys): point skipped
  STATUS_SYNC_TIMEOUT = 64,  // lead kernel: a wave waited past its limit for a panel; point given up
};

// Largest per-thread row count of wg_gj_inverse over all n <= nmax.
constexpr int gj_maxr(int nmax, int nt) {
  int m = 1;
  for (int n = 1; n <= nmax; ++n) {
    const int G = nt / n;
    const int r = (n + G - 1) / G;
    if (r > m) m = r;
  }
  return m;
}

// Gamma = i (Sigma - Sigma^dagger), element (i, j) of an n x n Sigma.
__device__ __forceinline__ cx<double> gamma_of(const cx<double>* s, int n, int i, int j) {
  const cx<double> a = s[i * n + j], b = s[j * n + i];
  const double wr = a.re - b.re, wi = a.im + b.im;  // a - conj(b)
  return cx<double>{-wi, wr};
}

struct RgfArgs {
  int K;
  const int32_t* bsz;        // device [K]
  const cx<double>* arena;   // device, nsys systems
  int64_t sys_stride;        // elements between systems
  const int32_t* sys_idx;    // device [P] or null (all system 0)
  int nsys;                  // systems in the arena (sys_idx range)
  int P;
  const cx<double>* z;       // device [P]: E + E_ref + i eta
  const cx<double>* sigL;    // device [P, nL, nL]
  int nL;
  const cx<double>* sigR;    // device [P, nR, nR]
  int nR;
  double* T;                 // device [P]
  int32_t* status;           // device [P] or null
  cx<double>* work;          // device [P, n_0, nmax]: per-point scratch of the kernels that keep
                             // the carried S out of LDS (rgf_s_out_of_lds): the final corner Q
};

template <typename R>
hipError_t launch_rgf_transmission(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                   hipStream_t st);
hipError_t launch_rgf_transmission_mixed(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                         hipStream_t st);
// FP32 block inverses refined by one FP64 Newton step, FP64 products; nmax <= 48.
hipError_t launch_rgf_transmission_refined(const RgfArgs& a, int nmax, bool uniform, bool ortho,
                                           hipStream_t st);
constexpr int REFINED_NMAX = 48;

struct LeadArgs {
  int n;                     // principal-layer size
  int nred;                  // rows of the device-lead coupling (Sigma is nred x nred)
  int mode;                  // 0: Sigma = tau g tau', 1: (eS - H) - g^{-1}, 2: g^{-1}
  const cx<double>* lead;    // per system: H, h01, S, s01 [n x n], hDL, sDL [nred x n]
  int64_t lead_stride;
  const int32_t* sys_idx;    // [P] or null
  int nsys;                  // lead records (sys_idx range)
  int P;
  const cx<double>* e;       // [P] E + E_ref (no eta)
  double eta;                // eta_lead
  cx<double>* eps;           // workspace [P, n, n] (fused kernel: D = eps - eps_s)
  cx<double>* gs;            // out [P, n, n]
  cx<double>* sigma;         // out [P, nred, nred]
  int32_t* iters;            // [P] or null
  int32_t* status;           // [P] or null
  double* resid;             // [P] or null
  double stop;               // decimation stop threshold on max(|alpha| + |beta|)
};

hipError_t launch_lead_self_energy(const LeadArgs& a, bool ortho, hipStream_t st);
// Workspace (complex elements) of the fused kernel: n x n per point (the
// decimation's running sum D = eps - eps_s; eps_s accumulates in g_s).
int64_t lead_self_energy_workspace(int n, int P);

}  // namespace negf
