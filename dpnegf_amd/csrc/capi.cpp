// extern "C" boundary of libdpnegf_amd.  Plain pointers and sizes only; every
// data pointer is device memory unless its name starts with h_.  Entry points
// return 0 on success and a negative code on failure; dpnegf_last_error()
// gives the message.  Nothing here allocates device memory: workspaces are
// passed in by the caller.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "dpnegf_amd.h"
#include "rgf_full.hpp"
#include "rgf_kernels.hpp"

using negf::cx;
static_assert(DPNEGF_FULL_NMAX == negf::BLOCK_NMAX, "block limit of the batched path");

namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(DPNEGF_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

int64_t arena_elems(int K, const int32_t* n) {
  int64_t s = 0;
  for (int q = 0; q < K; ++q) {
    s += 2 * (int64_t)n[q] * n[q];
    if (q + 1 < K) s += 4 * (int64_t)n[q] * n[q + 1];
  }
  return s;
}

int check_blocks(int K, const int32_t* h_bsz, int nmax_allowed, int* nmax_out) {
  if (K < 1) return fail(DPNEGF_ERR_ARG, "K must be >= 1 (got %d)", K);
  if (!h_bsz) return fail(DPNEGF_ERR_ARG, "h_bsz is null");
  int m = 0;
  for (int q = 0; q < K; ++q) {
    if (h_bsz[q] < 1) return fail(DPNEGF_ERR_ARG, "block %d has size %d", q, h_bsz[q]);
    if (h_bsz[q] > m) m = h_bsz[q];
  }
  if (m > nmax_allowed)
    return fail(DPNEGF_ERR_UNSUPPORTED, "block size %d exceeds this entry point's limit %d", m,
                nmax_allowed);
  *nmax_out = m;
  return 0;
}
}  // namespace

extern "C" {

int dpnegf_abi_version(void) { return DPNEGF_ABI_VERSION; }

const char* dpnegf_last_error(void) { return g_err.c_str(); }

int64_t dpnegf_arena_elems(int K, const int32_t* h_bsz) {
  if (K < 1 || !h_bsz) return -1;
  return arena_elems(K, h_bsz);
}

int64_t dpnegf_rgf_transmission_workspace_elems(int K, const int32_t* h_bsz, int P) {
  int nmax = 0;
  if (P < 0 || check_blocks(K, h_bsz, DPNEGF_TRANSMISSION_NMAX, &nmax)) return -1;
  return (int64_t)P * h_bsz[0] * nmax;
}

int dpnegf_rgf_transmission(void* stream, int precision, int K, const int32_t* h_bsz,
                            const int32_t* bsz, const void* arena, int64_t sys_stride, int nsys,
                            const int32_t* sys_idx, int P, const void* z, const void* sigL, int nL,
                            const void* sigR, int nR, int ortho, double* T, void* work,
                            int32_t* status) {
  int nmax = 0;
  if (int r = check_blocks(K, h_bsz, DPNEGF_TRANSMISSION_NMAX, &nmax)) return r;
  if (P < 0) return fail(DPNEGF_ERR_ARG, "P must be >= 0");
  if (P == 0) return 0;
  if (!bsz || !arena || !z || !sigL || !sigR || !T || !work)
    return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (nL < 1 || nL > h_bsz[0])
    return fail(DPNEGF_ERR_ARG, "left self-energy size %d not in [1, %d]", nL, h_bsz[0]);
  if (nR < 1 || nR > h_bsz[K - 1])
    return fail(DPNEGF_ERR_ARG, "right self-energy size %d not in [1, %d]", nR, h_bsz[K - 1]);
  if (nsys < 1) return fail(DPNEGF_ERR_ARG, "nsys must be >= 1");
  if (nsys > 1 && !sys_idx) return fail(DPNEGF_ERR_ARG, "sys_idx required when nsys > 1");
  if (nsys > 1 && sys_stride < arena_elems(K, h_bsz))
    return fail(DPNEGF_ERR_ARG, "sys_stride smaller than one system's arena");
  if (precision != DPNEGF_FP64 && precision != DPNEGF_MIXED && precision != DPNEGF_MIXED_REFINED)
    return fail(DPNEGF_ERR_ARG, "unknown precision %d", precision);
  if (precision == DPNEGF_MIXED_REFINED && nmax > negf::REFINED_NMAX)
    return fail(DPNEGF_ERR_UNSUPPORTED, "refined mixed precision needs blocks <= %d (got %d)",
                negf::REFINED_NMAX, nmax);
  negf::RgfArgs a;
  a.K = K; a.bsz = bsz; a.arena = (const cx<double>*)arena; a.sys_stride = sys_stride;
  a.sys_idx = sys_idx; a.nsys = nsys; a.P = P; a.z = (const cx<double>*)z;
  a.sigL = (const cx<double>*)sigL; a.nL = nL; a.sigR = (const cx<double>*)sigR; a.nR = nR;
  a.T = T; a.status = status; a.work = (cx<double>*)work;
  bool uniform = true;
  for (int q = 0; q < K; ++q) uniform = uniform && h_bsz[q] == nmax;
  hipError_t e;
  if (precision == DPNEGF_FP64)
    e = negf::launch_rgf_transmission<double>(a, nmax, uniform, ortho != 0, (hipStream_t)stream);
  else if (precision == DPNEGF_MIXED)
    e = negf::launch_rgf_transmission_mixed(a, nmax, uniform, ortho != 0, (hipStream_t)stream);
  else
    e = negf::launch_rgf_transmission_refined(a, nmax, uniform, ortho != 0, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rgf_transmission launch");
  return 0;
}

// Arguments of both lead entries; 1 when there is nothing to do (P == 0),
// < 0 on an argument error, 0 otherwise.
static int lead_args(negf::LeadArgs& a, int n, int nred, int mode, const void* lead,
                     int64_t lead_stride, int nsys, const int32_t* sys_idx, int P, const void* e,
                     double eta, double stop, void* gs, void* sigma, int32_t* iters,
                     int32_t* status, double* resid) {
  if (n < 1 || nred < 1) return fail(DPNEGF_ERR_ARG, "bad lead sizes n=%d nred=%d", n, nred);
  if (n > DPNEGF_FULL_NMAX || nred > DPNEGF_FULL_NMAX)
    return fail(DPNEGF_ERR_UNSUPPORTED, "lead size %d / coupling rows %d exceed %d", n, nred,
                DPNEGF_FULL_NMAX);
  if (mode < 0 || mode > 2) return fail(DPNEGF_ERR_ARG, "mode must be 0, 1 or 2");
  if (mode != 0 && nred != n) return fail(DPNEGF_ERR_ARG, "modes 1/2 need nred == n");
  if (P < 0) return fail(DPNEGF_ERR_ARG, "P must be >= 0");
  if (P == 0) return 1;
  if (!lead || !e || !gs || !sigma) return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (nsys < 1) return fail(DPNEGF_ERR_ARG, "nsys must be >= 1");
  if (nsys > 1 && !sys_idx) return fail(DPNEGF_ERR_ARG, "sys_idx required when nsys > 1");
  // records: H, H01, S, S01 (n x n), then H_DL, S_DL (nred x n) in mode 0 only
  const int64_t need = 4 * (int64_t)n * n + (mode == 0 ? 2 * (int64_t)nred * n : 0);
  if (nsys > 1 && lead_stride < need) return fail(DPNEGF_ERR_ARG, "lead_stride too small");
  a.n = n; a.nred = nred; a.mode = mode; a.lead = (const cx<double>*)lead;
  a.lead_stride = lead_stride; a.sys_idx = sys_idx; a.nsys = nsys; a.P = P;
  a.e = (const cx<double>*)e;
  a.eta = eta; a.eps = nullptr; a.gs = (cx<double>*)gs; a.sigma = (cx<double>*)sigma;
  a.iters = iters; a.status = status; a.resid = resid; a.stop = stop;
  return 0;
}

static bool lead_fused(int n, int nred) {
  return n <= DPNEGF_LEAD_NMAX && nred <= DPNEGF_LEAD_NMAX;
}

static int64_t lead_workspace(int n, int nred, int P) {
  return lead_fused(n, nred) ? negf::lead_self_energy_workspace(n, P)
                             : negf::lead_ls_batched_workspace(n, nred, P);
}

// One launch (fused kernel) or one batched decimation over the points of a.
static hipError_t lead_run(const negf::LeadArgs& a, cx<double>* ws, int ortho, hipStream_t st) {
  if (lead_fused(a.n, a.nred)) {
    negf::LeadArgs b = a;
    b.eps = ws;
    return negf::launch_lead_self_energy(b, ortho != 0, st);
  }
  return negf::lead_ls_batched(a, ws, st);
}

int dpnegf_lead_self_energy(void* stream, int n, int nred, int mode, const void* lead,
                            int64_t lead_stride, int nsys, const int32_t* sys_idx, int P,
                            const void* e, double eta, double stop, void* eps_ws, void* gs,
                            void* sigma, int ortho, int32_t* iters, int32_t* status,
                            double* resid) {
  negf::LeadArgs a;
  const int r = lead_args(a, n, nred, mode, lead, lead_stride, nsys, sys_idx, P, e, eta, stop, gs,
                          sigma, iters, status, resid);
  if (r) return r < 0 ? r : 0;
  if (!eps_ws) return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (!lead_fused(n, nred) && P > 65535)
    return fail(DPNEGF_ERR_ARG, "P=%d exceeds 65535 per call for blocks > %d", P,
                DPNEGF_LEAD_NMAX);
  hipError_t err = lead_run(a, (cx<double>*)eps_ws, ortho, (hipStream_t)stream);
  if (err != hipSuccess) return hip_fail(err, "lead_self_energy launch");
  return 0;
}

// ---- no-allocation lead entry ---------------------------------------------------
// The decimation workspace of dpnegf_lead_self_energy_noalloc lives in this
// file-scope device array; every GPU the library runs on holds its own copy
// (DPNEGF_NOALLOC_SCRATCH_ELEMS complex128, 128 MiB) until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it is a
// stand-in, not a device address): it uses the per-device address that
// hipGetSymbolAddress returns, looked up once per device on first use.
static __device__ __align__(256) cx<double> scratch[DPNEGF_NOALLOC_SCRATCH_ELEMS];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
cx<double>* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the stream's device (the current device for
// the null stream), looked up with that device current.
int scratch_address(hipStream_t st, cx<double>** out) {
  int cur = 0;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  int dev = cur;
  if (st) {
    e = hipStreamGetDevice(st, &dev);
    if (e != hipSuccess) return hip_fail(e, "hipStreamGetDevice");
  }
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
    return fail(DPNEGF_ERR_UNSUPPORTED, "device %d: the scratch table covers devices 0..%d", dev,
                SCRATCH_MAX_DEVICES - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    if (dev != cur && (e = hipSetDevice(dev)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (dev != cur) {
      const hipError_t e2 = hipSetDevice(cur);
      if (e == hipSuccess && e2 != hipSuccess) return hip_fail(e2, "hipSetDevice");
    }
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (cx<double>*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

int dpnegf_lead_self_energy_noalloc(void* stream, int n, int nred, int mode, const void* lead,
                                    int64_t lead_stride, int nsys, const int32_t* sys_idx, int P,
                                    const void* e, double eta, double stop, void* gs, void* sigma,
                                    int ortho, int32_t* iters, int32_t* status, double* resid) {
  negf::LeadArgs a;
  const int r = lead_args(a, n, nred, mode, lead, lead_stride, nsys, sys_idx, P, e, eta, stop, gs,
                          sigma, iters, status, resid);
  if (r) return r < 0 ? r : 0;
  // points per launch: the workspace is linear in the point count
  const int64_t per = lead_workspace(n, nred, 1);
  int64_t cap = DPNEGF_NOALLOC_SCRATCH_ELEMS / per;
  if (!lead_fused(n, nred) && cap > 65535) cap = 65535;
  if (cap < 1)
    return fail(DPNEGF_ERR_UNSUPPORTED,
                "lead of %d orbitals (%d coupling rows) needs %lld workspace elements per point; "
                "the no-allocation scratch holds %lld",
                n, nred, (long long)per, (long long)DPNEGF_NOALLOC_SCRATCH_ELEMS);
  cx<double>* ws = nullptr;
  if (int rc = scratch_address((hipStream_t)stream, &ws)) return rc;
  // equal chunks, rounded up to whole waves of 256 workgroups where that fits
  const int64_t nchunk = (P + cap - 1) / cap;
  int64_t chunk = (P + nchunk - 1) / nchunk;
  if (nchunk > 1 && (chunk + 255) / 256 * 256 <= cap) chunk = (chunk + 255) / 256 * 256;
  for (int64_t c0 = 0; c0 < P; c0 += chunk) {
    const int64_t c1 = c0 + chunk < P ? c0 + chunk : P;
    negf::LeadArgs b = a;
    b.P = (int)(c1 - c0);
    b.sys_idx = a.sys_idx ? a.sys_idx + c0 : nullptr;
    b.e = a.e + c0;
    b.gs = a.gs + c0 * n * n;
    b.sigma = a.sigma + c0 * nred * nred;
    b.iters = a.iters ? a.iters + c0 : nullptr;
    b.status = a.status ? a.status + c0 : nullptr;
    b.resid = a.resid ? a.resid + c0 : nullptr;
    // stream order keeps each chunk's use of the scratch after the previous one's
    hipError_t err = lead_run(b, ws, ortho, (hipStream_t)stream);
    if (err != hipSuccess) return hip_fail(err, "lead_self_energy_noalloc launch");
  }
  return 0;
}

int64_t dpnegf_lead_self_energy_workspace_elems(int n, int nred, int P) {
  if (n < 1 || nred < 1 || P < 0) return -1;
  return lead_workspace(n, nred, P);
}

int64_t dpnegf_rgf_full_workspace_elems(int K, const int32_t* h_bsz, int B) {
  int nmax = 0;
  if (!h_bsz || K < 1 || B < 0) return -1;
  for (int q = 0; q < K; ++q) nmax = h_bsz[q] > nmax ? h_bsz[q] : nmax;
  return 5 * (int64_t)B * nmax * nmax + negf::inverse_workspace(nmax, B);
}

int dpnegf_rgf_full(void* stream, int K, const int32_t* h_bsz, const void* arena, int B,
                    const void* z, const void* sigL, int nL, const void* sigR, int nR, int ortho,
                    const void* seP, void* gl, void* grd, void* grl, void* gru, void* lc, void* g_trans,
                    const void* s_in, void* gin, void* gnd, void* gnl, void* gnu,
                    const void* s_out, void* gip, void* gpd, void* gpl, void* gpu,
                    const void* s_bil, void* gbi, void* gbd, void* gbl, void* gbu, void* work,
                    int32_t* status) {
  int nmax = 0;
  if (int r = check_blocks(K, h_bsz, DPNEGF_FULL_NMAX, &nmax)) return r;
  if (B < 0) return fail(DPNEGF_ERR_ARG, "B must be >= 0");
  if (B == 0) return 0;
  if (B > 65535) return fail(DPNEGF_ERR_ARG, "B=%d exceeds 65535 per call", B);
  if (!arena || !z || !gl || !grd || !g_trans || !work)
    return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (K > 1 && (!grl || !gru)) return fail(DPNEGF_ERR_ARG, "grl/gru required for K > 1");
  if ((sigL && (nL < 1 || nL > h_bsz[0])) || (sigR && (nR < 1 || nR > h_bsz[K - 1])))
    return fail(DPNEGF_ERR_ARG, "self-energy larger than its corner block");
  if (s_in && (!gin || !gnd || (K > 1 && (!gnl || !gnu))))
    return fail(DPNEGF_ERR_ARG, "lesser outputs missing");
  if (s_out && (!gip || !gpd || (K > 1 && (!gpl || !gpu))))
    return fail(DPNEGF_ERR_ARG, "greater outputs missing");
  if (s_bil && (!gbi || !gbd || (K > 1 && (!gbl || !gbu))))
    return fail(DPNEGF_ERR_ARG, "bilinear outputs missing");
  negf::RgfFullArgs a;
  std::memset(&a, 0, sizeof(a));
  a.K = K; a.B = B; a.nmax = nmax; a.ortho = ortho; a.bsz = h_bsz;
  a.arena = (const cx<double>*)arena; a.z = (const cx<double>*)z;
  a.sigL = (const cx<double>*)sigL; a.nL = nL; a.sigR = (const cx<double>*)sigR; a.nR = nR;
  a.seP = (const cx<double>*)seP;
  a.gl = (cx<double>*)gl; a.grd = (cx<double>*)grd; a.grl = (cx<double>*)grl;
  a.gru = (cx<double>*)gru; a.lc = (cx<double>*)lc; a.g_trans = (cx<double>*)g_trans;
  a.s_in = (const cx<double>*)s_in; a.gin = (cx<double>*)gin; a.gnd = (cx<double>*)gnd;
  a.gnl = (cx<double>*)gnl; a.gnu = (cx<double>*)gnu;
  a.s_out = (const cx<double>*)s_out; a.gip = (cx<double>*)gip; a.gpd = (cx<double>*)gpd;
  a.gpl = (cx<double>*)gpl; a.gpu = (cx<double>*)gpu;
  a.s_bil = (const cx<double>*)s_bil; a.gbi = (cx<double>*)gbi; a.gbd = (cx<double>*)gbd;
  a.gbl = (cx<double>*)gbl; a.gbu = (cx<double>*)gbu;
  a.work = (cx<double>*)work; a.status = status;
  a.inv_ws = negf::inverse_workspace(nmax, B) > 0 ? a.work + 5 * (int64_t)B * nmax * nmax : nullptr;
  hipError_t e = negf::rgf_full(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rgf_full");
  return 0;
}

int64_t dpnegf_rgf_grad_workspace_elems(int K, const int32_t* h_bsz, int B) {
  if (!h_bsz || K < 1 || B < 0) return -1;
  return negf::rgf_grad_workspace(K, h_bsz, B);
}

// Arguments of both adjoint entries; 1 when there is nothing to do (B == 0),
// < 0 on an argument error, 0 otherwise.
static int grad_args(negf::RgfGradArgs& a, int K, const int32_t* h_bsz, const void* arena, int B,
                     const void* z, const void* sigL, int nL, const void* sigR, int nR, int ortho,
                     const double* w, void* grad_hd, void* grad_hu, void* grad_hl,
                     void* grad_sigL, void* grad_sigR, void* grad_sd, void* grad_su,
                     void* grad_sl, void* work, int32_t* status) {
  int nmax = 0;
  if (int r = check_blocks(K, h_bsz, DPNEGF_FULL_NMAX, &nmax)) return r;
  if (B < 0) return fail(DPNEGF_ERR_ARG, "B must be >= 0");
  if (B == 0) return 1;
  if (B > 65535) return fail(DPNEGF_ERR_ARG, "B=%d exceeds 65535 per call", B);
  if (!arena || !z || !sigL || !sigR || !w || !grad_hd || !work)
    return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (K > 1 && (!grad_hu || !grad_hl)) return fail(DPNEGF_ERR_ARG, "grad_hu/grad_hl required");
  if (nL < 1 || nL > h_bsz[0] || nR < 1 || nR > h_bsz[K - 1])
    return fail(DPNEGF_ERR_ARG, "self-energy larger than its corner block");
  a.K = K; a.B = B; a.ortho = ortho; a.bsz = h_bsz; a.arena = (const cx<double>*)arena;
  a.z = (const cx<double>*)z; a.sigL = (const cx<double>*)sigL; a.nL = nL;
  a.sigR = (const cx<double>*)sigR; a.nR = nR; a.w = w; a.grad_hd = (cx<double>*)grad_hd;
  a.grad_hu = (cx<double>*)grad_hu; a.grad_hl = (cx<double>*)grad_hl;
  a.grad_sigL = (cx<double>*)grad_sigL; a.grad_sigR = (cx<double>*)grad_sigR;
  a.grad_sd = (cx<double>*)grad_sd; a.grad_su = (cx<double>*)grad_su;
  a.grad_sl = (cx<double>*)grad_sl;
  if (K > 1 && grad_sd && (!grad_su || !grad_sl))
    return fail(DPNEGF_ERR_ARG, "grad_su/grad_sl required with grad_sd");
  a.work = (cx<double>*)work;
  a.status = status;
  return 0;
}

int dpnegf_rgf_transmission_grad(void* stream, int K, const int32_t* h_bsz, const void* arena,
                                 int B, const void* z, const void* sigL, int nL, const void* sigR,
                                 int nR, int ortho, const double* w, void* grad_hd, void* grad_hu,
                                 void* grad_hl, void* grad_sigL, void* grad_sigR, void* grad_sd,
                                 void* grad_su, void* grad_sl, void* work, int32_t* status) {
  negf::RgfGradArgs a;
  const int r = grad_args(a, K, h_bsz, arena, B, z, sigL, nL, sigR, nR, ortho, w, grad_hd, grad_hu,
                          grad_hl, grad_sigL, grad_sigR, grad_sd, grad_su, grad_sl, work, status);
  if (r) return r < 0 ? r : 0;
  hipError_t e = negf::rgf_grad(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rgf_transmission_grad");
  return 0;
}

int64_t dpnegf_rgf_grad_ckpt_workspace_elems(int K, const int32_t* h_bsz, int B, int stride) {
  if (!h_bsz || K < 1 || B < 0 || stride < 1) return -1;
  return negf::rgf_grad_ckpt_workspace(K, h_bsz, B, stride);
}

int dpnegf_rgf_transmission_grad_ckpt(void* stream, int K, const int32_t* h_bsz,
                                      const void* arena, int B, const void* z, const void* sigL,
                                      int nL, const void* sigR, int nR, int ortho, const double* w,
                                      void* grad_hd, void* grad_hu, void* grad_hl,
                                      void* grad_sigL, void* grad_sigR, void* grad_sd,
                                      void* grad_su, void* grad_sl, int stride, void* work,
                                      int32_t* status) {
  if (stride < 1) return fail(DPNEGF_ERR_ARG, "checkpoint stride %d < 1", stride);
  negf::RgfGradArgs a;
  const int r = grad_args(a, K, h_bsz, arena, B, z, sigL, nL, sigR, nR, ortho, w, grad_hd, grad_hu,
                          grad_hl, grad_sigL, grad_sigR, grad_sd, grad_su, grad_sl, work, status);
  if (r) return r < 0 ? r : 0;
  hipError_t e = negf::rgf_grad_ckpt(a, stride, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rgf_transmission_grad_ckpt");
  return 0;
}

int dpnegf_batch_sum_conjt(void* stream, void* out, const void* in, int m, int n, int B,
                           int accumulate) {
  if (m < 0 || n < 0 || B < 0) return fail(DPNEGF_ERR_ARG, "negative size");
  if (m == 0 || n == 0) return 0;
  if (!out || (B > 0 && !in)) return fail(DPNEGF_ERR_ARG, "null device pointer");
  hipError_t e = negf::batch_sum_conjt((cx<double>*)out, (const cx<double>*)in, m, n, B, accumulate,
                                       (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "batch_sum_conjt");
  return 0;
}

int dpnegf_weighted_re_sum(void* stream, double* out, const void* in, const void* c,
                           int64_t count, int B, int accumulate) {
  if (count < 0 || B < 0) return fail(DPNEGF_ERR_ARG, "negative size");
  if (count == 0) return 0;
  if (!out || (B > 0 && (!in || !c))) return fail(DPNEGF_ERR_ARG, "null device pointer");
  hipError_t e = negf::weighted_re_sum(out, (const cx<double>*)in, (const cx<double>*)c, count, B,
                                       accumulate, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "weighted_re_sum");
  return 0;
}

int dpnegf_zgemm_batched(void* stream, int B, int m, int n, int k, const void* A, int lda,
                         int64_t strideA, int opA, const void* Bm, int ldb, int64_t strideB, int opB,
                         double alpha_re, double alpha_im, double beta_re, double beta_im, void* C,
                         int ldc, int64_t strideC) {
  if (B < 0 || m < 0 || n < 0 || k < 0) return fail(DPNEGF_ERR_ARG, "negative size");
  if (B > 65535) return fail(DPNEGF_ERR_ARG, "B=%d exceeds 65535 per call", B);
  if (B == 0 || m == 0 || n == 0) return 0;
  if (!A || !Bm || !C) return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (opA < 0 || opA > 3 || opB < 0 || opB > 3) return fail(DPNEGF_ERR_ARG, "op must be 0..3");
  hipError_t e = negf::zgemm_batched(B, m, n, k, (const cx<double>*)A, lda, strideA, opA,
                                     (const cx<double>*)Bm, ldb, strideB, opB,
                                     cx<double>{alpha_re, alpha_im}, cx<double>{beta_re, beta_im},
                                     (cx<double>*)C, ldc, strideC, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "zgemm_batched");
  return 0;
}

int dpnegf_zinv_batched(void* stream, void* A, int n, int B, int32_t* status) {
  if (n < 1 || n > negf::GJ_NMAX_LARGE)
    return fail(DPNEGF_ERR_ARG, "n=%d outside 1..%d", n, negf::GJ_NMAX_LARGE);
  if (B < 0) return fail(DPNEGF_ERR_ARG, "negative batch");
  if (B == 0) return 0;
  if (!A) return fail(DPNEGF_ERR_ARG, "null device pointer");
  hipError_t e = negf::zinv_batched((hipStream_t)stream, (cx<double>*)A, n, B, status);
  if (e != hipSuccess) return hip_fail(e, "zinv_batched");
  return 0;
}

int dpnegf_diag_product(void* stream, void* out, const void* A, int64_t strideA, const void* Bm,
                        int64_t strideB, int m, int n, int B, int accumulate) {
  if (B < 0 || m < 0 || n < 0) return fail(DPNEGF_ERR_ARG, "negative size");
  if (B == 0 || m == 0) return 0;
  if (!out || !A || !Bm) return fail(DPNEGF_ERR_ARG, "null device pointer");
  hipError_t e = negf::diag_product((cx<double>*)out, (const cx<double>*)A, strideA,
                                    (const cx<double>*)Bm, strideB, m, n, B, accumulate,
                                    (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "diag_product");
  return 0;
}

int64_t dpnegf_rgf_transmission_blocked_workspace_elems(int K, const int32_t* h_bsz, int B) {
  if (!h_bsz || K < 1 || B < 0) return -1;
  return negf::rgf_transmission_ops_workspace(K, h_bsz, B);
}

int64_t dpnegf_autotune_query(int K, const int32_t* h_bsz, int64_t mem_bytes) {
  if (!h_bsz || K < 1 || mem_bytes < 0) return fail(DPNEGF_ERR_ARG, "bad arguments");
  for (int q = 0; q < K; ++q)
    if (h_bsz[q] < 1) return fail(DPNEGF_ERR_ARG, "block %d has size %d", q, h_bsz[q]);
  // workspace plus z, Sigma_L, Sigma_R, T and status per point, linear in B
  const int64_t w1 = negf::rgf_transmission_ops_workspace(K, h_bsz, 1);
  const int64_t w2 = negf::rgf_transmission_ops_workspace(K, h_bsz, 2);
  const int64_t n0 = h_bsz[0], nl = h_bsz[K - 1];
  const int64_t per = (w2 - w1) * 16 + (1 + n0 * n0 + nl * nl) * 16 + 8 + 4;
  const int64_t fixed = w1 * 16 - (w2 - w1) * 16;
  int64_t b = (mem_bytes - fixed) / per;
  if (b < 0) b = 0;
  return b > 65535 ? 65535 : b;
}

int dpnegf_tune_set(int kind, int m, int n, int k, int choice) {
  if (negf::tune_set(kind, m, n, k, choice))
    return fail(DPNEGF_ERR_ARG, "tune_set: bad entry kind=%d (%d, %d, %d) choice=%d", kind, m, n,
                k, choice);
  return 0;
}

int dpnegf_tune_get(int kind, int m, int n, int k) { return negf::tune_get(kind, m, n, k); }

int dpnegf_tune_builtin(int kind, int m, int n, int k) {
  return negf::tune_builtin(kind, m, n, k);
}

void dpnegf_tune_clear(void) { negf::tune_clear(); }

void dpnegf_tune_record(int on) { negf::tune_record(on != 0); }

int dpnegf_tune_recorded(int32_t* out, int cap) {
  if (cap < 0 || (cap > 0 && !out)) return fail(DPNEGF_ERR_ARG, "bad output buffer");
  return negf::tune_recorded(out, cap);
}

int dpnegf_rgf_transmission_blocked(void* stream, int K, const int32_t* h_bsz, const void* arena,
                                    int B, const void* z, const void* sigL, int nL,
                                    const void* sigR, int nR, int ortho, double* T, void* work,
                                    int32_t* status) {
  int nmax = 0;
  if (int r = check_blocks(K, h_bsz, DPNEGF_FULL_NMAX, &nmax)) return r;
  if (B < 0) return fail(DPNEGF_ERR_ARG, "B must be >= 0");
  if (B == 0) return 0;
  if (B > 65535) return fail(DPNEGF_ERR_ARG, "B=%d exceeds 65535 per call", B);
  if (!arena || !z || !sigL || !sigR || !T || !work) return fail(DPNEGF_ERR_ARG, "null device pointer");
  if (nL < 1 || nL > h_bsz[0] || nR < 1 || nR > h_bsz[K - 1])
    return fail(DPNEGF_ERR_ARG, "self-energy larger than its corner block");
  negf::RgfOpsArgs a;
  a.K = K; a.B = B; a.ortho = ortho; a.bsz = h_bsz; a.arena = (const cx<double>*)arena;
  a.z = (const cx<double>*)z; a.sigL = (const cx<double>*)sigL; a.nL = nL;
  a.sigR = (const cx<double>*)sigR; a.nR = nR; a.T = T; a.work = (cx<double>*)work;
  a.status = status;
  hipError_t e = negf::rgf_transmission_ops(a, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "rgf_transmission_blocked");
  return 0;
}

}  // extern "C"
