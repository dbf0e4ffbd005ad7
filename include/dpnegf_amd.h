/* libdpnegf_amd -- C ABI of the MI355X (gfx950) NEGF hot path.
 *
 * Every data pointer is device (HBM) memory unless its name starts with h_.
 * Complex numbers are interleaved (re, im) doubles, the storage of torch's
 * complex128, so torch tensors are passed zero-copy.  `stream` is a
 * hipStream_t (torch's current stream).  Calls are asynchronous on `stream`;
 * they allocate no device memory (workspaces are arguments).  Return 0 on
 * success, a negative DPNEGF_ERR_* code otherwise; dpnegf_last_error() gives
 * the message (thread-local).
 *
 * Device Hamiltonian "arena" of one system (one k-point), row-major blocks in
 * this order, q = 0..K-1:  hd_q [n_q x n_q], sd_q [n_q x n_q], then for
 * q < K-1:  hu_q [n_q x n_{q+1}], su_q, hl_q [n_{q+1} x n_q], sl_q.
 * These are the block lists of get_hs_device
 * (dpnegf/negf/negf_hamiltonian_init.py:745-880).
 */
#ifndef DPNEGF_AMD_H
#define DPNEGF_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPNEGF_ABI_VERSION 7

#define DPNEGF_OK 0
#define DPNEGF_ERR_ARG (-1)
#define DPNEGF_ERR_UNSUPPORTED (-2)
#define DPNEGF_ERR_HIP (-3)

#define DPNEGF_FP64 0
#define DPNEGF_MIXED 1         /* FP32 throughout (about 1e-4 on T of order 30) */
#define DPNEGF_MIXED_REFINED 2 /* FP32 block inverses + one FP64 Newton step, FP64 products */

/* size limits of the entry points */
#define DPNEGF_TRANSMISSION_NMAX 64
#define DPNEGF_LEAD_NMAX 48
#define DPNEGF_FULL_NMAX 1024

/* per-point status bits */
#define DPNEGF_STATUS_SINGULAR 1
#define DPNEGF_STATUS_LS_NOCONV 2
#define DPNEGF_STATUS_LS_RESID_FAIL 4
#define DPNEGF_STATUS_LS_RESID_WARN 8
#define DPNEGF_STATUS_REFINE_FALLBACK 16
#define DPNEGF_STATUS_BAD_INDEX 32 /* sys_idx[p] outside [0, nsys): point skipped */
#define DPNEGF_STATUS_SYNC_TIMEOUT 64 /* internal: a lead-kernel wave waited past its
                                        limit for a published panel; point given up */

int dpnegf_abi_version(void);
const char* dpnegf_last_error(void);
/* complex elements of one system's arena */
int64_t dpnegf_arena_elems(int K, const int32_t* h_bsz);

/* Transmission T(E,k) for P points, one workgroup per point.
 * Replaces recursive_gf(...)[0] (dpnegf/negf/recursive_green_cal.py:324-477,
 * forward sweep :198-207, g_trans :228) followed by DeviceProperty._cal_tc_
 * (dpnegf/negf/device_property.py:374-406).
 *   z[p]       = E + E_ref + i*eta_device                 (complex)
 *   sigL[p]    = left self-energy  [nL x nL], nL <= n_0, added to hd_0's top-left corner
 *   sigR[p]    = right self-energy [nR x nR], nR <= n_{K-1}, bottom-right of hd_{K-1}
 *   sys_idx[p] = system (k-point) of point p; arena + sys * sys_stride (null: all 0);
 *                an index outside [0, nsys) gives T[p] = NaN and DPNEGF_STATUS_BAD_INDEX
 *   ortho      = 1 if every sd = I and su = sl = 0 (S blocks are then not read)
 *   precision  = DPNEGF_FP64, DPNEGF_MIXED or DPNEGF_MIXED_REFINED (blocks <= 48;
 *                a block whose FP32 inverse leaves a residual above 1e-4 is
 *                inverted in FP64 instead: DPNEGF_STATUS_REFINE_FALLBACK)
 *   T[p]       = Re Tr[Gamma_L G_{0,K-1} Gamma_R G_{0,K-1}^dagger]
 *   work       = device buffer of dpnegf_rgf_transmission_workspace_elems(K, h_bsz, P)
 *                complex elements (per-point scratch: the corner of G_{0,K-1} of
 *                the kernels that hold the sweep's carried G_{0,q} U_q outside
 *                LDS); contents on entry ignored
 *   status[p]  = DPNEGF_STATUS_* bits (may be null)
 * Block sizes up to DPNEGF_TRANSMISSION_NMAX. */
int dpnegf_rgf_transmission(void* stream, int precision, int K, const int32_t* h_bsz,
                            const int32_t* bsz, const void* arena, int64_t sys_stride, int nsys,
                            const int32_t* sys_idx, int P, const void* z, const void* sigL, int nL,
                            const void* sigR, int nR, int ortho, double* T, void* work,
                            int32_t* status);
/* P n_0 nmax: complex elements of dpnegf_rgf_transmission's work buffer (-1 on
 * bad arguments).  No device work. */
int64_t dpnegf_rgf_transmission_workspace_elems(int K, const int32_t* h_bsz, int P);

/* Lopez-Sancho surface Green's function and lead self-energy for P points.
 * Replaces selfEnergy (dpnegf/negf/surface_green.py:9-91) and its decimation
 * cores (surface_green.py:102-159, 169-226), i.e. LeadProperty.self_energy_cal
 * (dpnegf/negf/lead_property.py:266-384) over a batch of (E, k).
 *   lead (per system, at lead + sys * lead_stride): hL, hLL, sL, sLL [n x n],
 *        then in mode 0 hDL, sDL [nred x n]   (the tuple of get_hs_lead,
 *        dpnegf/negf/negf_hamiltonian_init.py:974, with HDL_reduced applied);
 *        lead_stride >= 4 n^2 (+ 2 nred n in mode 0) when nsys > 1
 *   e[p]   = E + E_ref (complex allowed); the decimation uses e + i*eta
 *   mode 0: Sigma = (e sDL - hDL) g_s (e sDL^H - hDL^H)      [nred x nred]
 *   mode 1: Sigma = (e sL - hL) - g_s^{-1};  mode 2 (Bulk): g_s^{-1}
 *   stop   = decimation threshold on max|alpha| + |beta| (reference: 1e-40)
 *   eps_ws = workspace of dpnegf_lead_self_energy_workspace_elems(n, nred, P) complex
 *            elements; gs out [P x n x n]; sigma out [P x nred x nred]
 *   iters/status/resid per point (may be null); status bits LS_* replace the
 *   reference's RuntimeError / ArithmeticError / warning.  sys_idx[p] outside
 *   [0, nsys) sets DPNEGF_STATUS_BAD_INDEX and no lead record outside the
 *   arena is read.
 * n, nred <= DPNEGF_LEAD_NMAX run one fused workgroup per point; larger leads
 * (up to DPNEGF_FULL_NMAX, P <= 65535 per call) iterate batched block operations
 * until every point has converged. */
int64_t dpnegf_lead_self_energy_workspace_elems(int n, int nred, int P);

/* Complex128 elements of the device scratch dpnegf_lead_self_energy_noalloc
 * uses (128 MiB per GPU the library runs on, held until the process ends). */
#define DPNEGF_NOALLOC_SCRATCH_ELEMS 8388608

/* dpnegf_lead_self_energy without the eps_ws argument: the same kernels and
 * results, with the decimation workspace in a file-scope device array of
 * DPNEGF_NOALLOC_SCRATCH_ELEMS elements instead of a caller buffer, so the
 * call allocates nothing (no hipMalloc, no caller workspace).  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  Points are processed in stream-ordered chunks that fit the scratch;
 * a lead whose single point needs more than the scratch (above about 830
 * orbitals) is rejected with DPNEGF_ERR_UNSUPPORTED.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (concurrent
 * streams or host threads) race and must be serialised by the caller.  The
 * stream's device must be the one whose scratch is used; its address is
 * looked up once per device on the first call there. */
int dpnegf_lead_self_energy_noalloc(void* stream, int n, int nred, int mode, const void* lead,
                                    int64_t lead_stride, int nsys, const int32_t* sys_idx, int P,
                                    const void* e, double eta, double stop, void* gs, void* sigma,
                                    int ortho, int32_t* iters, int32_t* status, double* resid);
int dpnegf_lead_self_energy(void* stream, int n, int nred, int mode, const void* lead,
                            int64_t lead_stride, int nsys, const int32_t* sys_idx, int P,
                            const void* e, double eta, double stop, void* eps_ws, void* gs,
                            void* sigma, int ortho, int32_t* iters, int32_t* status,
                            double* resid);

/* Complex elements of the work buffer dpnegf_rgf_transmission_blocked needs. */
int64_t dpnegf_rgf_transmission_blocked_workspace_elems(int K, const int32_t* h_bsz, int B);

/* Batch shapes for the autotuner: the largest number of points B per
 * dpnegf_rgf_transmission_blocked call whose workspace and per-point buffers
 * (z, Sigma_L and Sigma_R at the corner block sizes, T, status) fit in
 * mem_bytes, capped at 65535; negative on bad arguments.  No device work.
 * The reference has no counterpart: its runner splits the energy grid into
 * chunks of a user-set size (NEGF.py:602-631). */
int64_t dpnegf_autotune_query(int K, const int32_t* h_bsz, int64_t mem_bytes);

/* Kernel-shape tuning table, process-wide, read by the block-operation path
 * (dpnegf_rgf_transmission_blocked, dpnegf_rgf_full, the batched decimation of
 * dpnegf_lead_self_energy above 48 orbitals, dpnegf_zgemm_batched,
 * dpnegf_zinv_batched) each time it launches a product or an inverse.
 *  - DPNEGF_TUNE_GEMM_TILE, key (m, n, k) of a batched product: choice
 *    DPNEGF_TILE_32x32 / _32x64 / _64x32 / _64x64, the output tile (rows x
 *    columns) of one workgroup.
 *  - DPNEGF_TUNE_INVERSE, key (n, 0, 0) of a batched inverse: choice
 *    DPNEGF_INV_REGISTER (register-resident Gauss-Jordan, n <= 128),
 *    DPNEGF_INV_LDS (blocked Gauss-Jordan in LDS with matrix-core updates,
 *    n <= 96) or DPNEGF_INV_SCHUR (inverses of the two diagonal halves joined
 *    through the Schur complement on the matrix cores, then one FP64 Newton
 *    step; 97 <= n <= 128, built in there).  Inverses above 128 always take
 *    the Schur form and recurse on their halves, keyed by the half;
 *    dpnegf_zinv_batched, which has no workspace, uses the register form
 *    where the Schur form would be chosen.
 * Choice 0 removes an entry; keys without an entry keep the built-in choice.
 * Every choice gives the same result to rounding.  Set entries before a call
 * is enqueued or captured, not while another thread launches.  While
 * recording is on (dpnegf_tune_record(1), which also forgets earlier shapes)
 * the library collects each key it looks up; dpnegf_tune_recorded writes
 * min(count, cap) records of 4 int32 (kind, m, n, k) and returns the count.
 * The reference has no counterpart: torch picks its library kernels. */
#define DPNEGF_TUNE_GEMM_TILE 0
#define DPNEGF_TUNE_INVERSE 1
#define DPNEGF_TILE_32x32 1
#define DPNEGF_TILE_32x64 2
#define DPNEGF_TILE_64x32 3
#define DPNEGF_TILE_64x64 4
#define DPNEGF_INV_REGISTER 1
#define DPNEGF_INV_LDS 2
#define DPNEGF_INV_SCHUR 3
int dpnegf_tune_set(int kind, int m, int n, int k, int choice);
int dpnegf_tune_get(int kind, int m, int n, int k);
/* The choice a key gets without a table entry; 0 for a key no driver can use. */
int dpnegf_tune_builtin(int kind, int m, int n, int k);
void dpnegf_tune_clear(void);
void dpnegf_tune_record(int on);
int dpnegf_tune_recorded(int32_t* out, int cap);

/* Transmission of B energies of one system for block sizes up to
 * DPNEGF_FULL_NMAX (the fused kernel stops at 64): the same sweep driven block
 * by block from the host with batched products and inverses.  Same result
 * contract as dpnegf_rgf_transmission (recursive_green_cal.py:198-228 and
 * device_property.py:374-406).  B <= 65535. */
int dpnegf_rgf_transmission_blocked(void* stream, int K, const int32_t* h_bsz, const void* arena,
                                    int B, const void* z, const void* sigL, int nL,
                                    const void* sigR, int nR, int ortho, double* T, void* work,
                                    int32_t* status);

/* Complex elements of the work buffer dpnegf_rgf_full needs. */
int64_t dpnegf_rgf_full_workspace_elems(int K, const int32_t* h_bsz, int B);

/* Every retarded block (and optionally lesser / greater) for B energies of one
 * system.  Replaces recursive_gf_cal (dpnegf/negf/recursive_green_cal.py:4-296)
 * including its lesser/greater quirks.  Per-block outputs are packed arenas,
 * block q at the sum of the sizes of blocks < q:
 *   gl, grd, gin, gnd, gip, gpd : B x n_q x n_q
 *   grl, gnl, gpl               : B x n_{q+1} x n_q
 *   gru, gnu, gpu               : B x n_q x n_{q+1}
 *   lc (gr_lc, may be null)     : B x n_q x n_{K-1}
 *   g_trans                     : B x n_0 x n_{K-1}
 *   s_in / s_out (may be null: no lesser / greater): B x n_q x n_q arenas
 *   seP (may be null): energy-dependent addend of every diagonal block,
 *                 B x n_q x n_q arena (the 3-D seP of recursive_gf, :411-414)
 *   s_bil (may be null): block-diagonal source S of the bilinear product
 *                 X = G S G (no counterpart in the reference: the adjoint of the
 *                 density, dJ/dH = (sum_b c_b G W G)^H); its blocks go to
 *                 gbi (left-connected), gbd, gbl (X_{q+1,q}), gbu (X_{q,q+1})
 * Block sizes up to DPNEGF_FULL_NMAX; B <= 65535 per call. */
int dpnegf_rgf_full(void* stream, int K, const int32_t* h_bsz, const void* arena, int B,
                    const void* z, const void* sigL, int nL, const void* sigR, int nR, int ortho,
                    const void* seP, void* gl, void* grd, void* grl, void* gru, void* lc, void* g_trans,
                    const void* s_in, void* gin, void* gnd, void* gnl, void* gnu,
                    const void* s_out, void* gip, void* gpd, void* gpl, void* gpu,
                    const void* s_bil, void* gbi, void* gbd, void* gbl, void* gbu, void* work,
                    int32_t* status);

/* out (+)= (sum_b in[b])^H: B contiguous m x n complex blocks summed over the
 * batch, conjugated and transposed into out [n x m] (a gradient block in
 * torch's complex convention, dJ/dRe + i dJ/dIm).  Partial sums are added in
 * a fixed order.  No counterpart in the reference (autograd does this). */
int dpnegf_batch_sum_conjt(void* stream, void* out, const void* in, int m, int n, int B,
                           int accumulate);

/* Complex elements of the work buffer dpnegf_rgf_transmission_grad needs. */
int64_t dpnegf_rgf_grad_workspace_elems(int K, const int32_t* h_bsz, int B);

/* Adjoint of the transmission: for the objective J = sum_b w[b] T(E_b),
 * grad_X = dJ/dRe(X) + i dJ/dIm(X) for every on-site block hd_q and coupling
 * block hu_q, hl_q (torch's complex gradient convention), summed over the B
 * energies.  New: the reference obtains this only by torch autograd through
 * recursive_gf (dpnegf/negf/recursive_green_cal.py:89-112).  Uses
 *   dT = 2 Re Tr[W dH],  W = G Gamma_R G^H Gamma_L G,
 *   W_ab = G_{a,K-1} C G_{0,b},  C = Gamma_R G_{0,K-1}^H Gamma_L,
 * with the last block column from the backward sweep and the first block row
 * G_{0,b} = (g_0 U_0 ... g_{b-1} U_{b-1}) G_bb.
 *   grad_hd [sum n_q^2], grad_hu [sum n_q n_{q+1}], grad_hl [sum n_{q+1} n_q]
 * and, per energy (not summed; either may be null), the gradients of the lead
 * self-energies, which enter G as H does in the corners of the first and last
 * blocks and enter T again through Gamma = i (Sigma - Sigma^H):
 *   grad_sigL [B x nL x nL] = (W_00 + i 2 w Q Gamma_R Q^H)^H on the nL corner,
 *   grad_sigR [B x nR x nR] = (W_{K-1,K-1} + i 2 w Q^H Gamma_L Q)^H on the nR
 *   corner, Q = G_{0,K-1} restricted to the two corners;
 * and (any may be null; grad_su / grad_sl required with grad_sd when K > 1)
 * the gradients of the overlap blocks, which enter only through M = z S - H:
 *   grad_sd / grad_su / grad_sl = -sum_b conj(z_b) (W_b)^H on each block,
 * laid out as grad_hd / grad_hu / grad_hl.
 * Block sizes up to DPNEGF_FULL_NMAX; B <= 65535 per call. */
int dpnegf_rgf_transmission_grad(void* stream, int K, const int32_t* h_bsz, const void* arena,
                                 int B, const void* z, const void* sigL, int nL, const void* sigR,
                                 int nR, int ortho, const double* w, void* grad_hd, void* grad_hu,
                                 void* grad_hl, void* grad_sigL, void* grad_sigR, void* grad_sd,
                                 void* grad_su, void* grad_sl, void* work, int32_t* status);

/* Complex elements of the work buffer dpnegf_rgf_transmission_grad_ckpt needs
 * for checkpoint stride `stride` (-1 for stride < 1): B times
 * (sum over checkpoints c = 0, s, 2s, ... of n_c^2 + n_0 n_c, less n_0^2)
 * + B (2 s + 18) nmax^2, plus the large-block inverse's workspace. */
int64_t dpnegf_rgf_grad_ckpt_workspace_elems(int K, const int32_t* h_bsz, int B, int stride);

/* dpnegf_rgf_transmission_grad with checkpointing: the forward sweep keeps
 * the left-connected g_c and the first-row product g_0 U_0 ... g_{c-1} U_{c-1}
 * only at c = 0, s, 2s, ... (s = stride) and the gradient sweep recomputes each
 * segment of s blocks from its checkpoint, so the work buffer holds
 * O(K / s + s) blocks per energy instead of O(K) (s near sqrt(K) minimises
 * it), for one more forward sweep.  Same arguments and outputs otherwise; the
 * checkpointing has no counterpart in the reference, whose gradients come from
 * torch autograd through recursive_gf (dpnegf/negf/recursive_green_cal.py:89-112). */
int dpnegf_rgf_transmission_grad_ckpt(void* stream, int K, const int32_t* h_bsz,
                                      const void* arena, int B, const void* z, const void* sigL,
                                      int nL, const void* sigR, int nR, int ortho, const double* w,
                                      void* grad_hd, void* grad_hu, void* grad_hl,
                                      void* grad_sigL, void* grad_sigR, void* grad_sd,
                                      void* grad_su, void* grad_sl, int stride, void* work,
                                      int32_t* status);

/* out[e] = (accumulate ? out[e] : 0) + sum_b Re(c[b] * in[b * count + e]):
 * the pole / quadrature sum of the Ozaki density, Ozaki.integrate
 * (dpnegf/negf/density.py:171-181, 189-195), over the B points of a batch.
 * out: double [count]; in: complex [B x count]; c: complex [B]. */
int dpnegf_weighted_re_sum(void* stream, double* out, const void* in, const void* c,
                           int64_t count, int B, int accumulate);

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b], complex128, row-major with
 * leading dimensions; op 0 = N, 1 = conjugate transpose, 2 = transpose,
 * 3 = conjugate; a stride of 0 shares one matrix across the batch.  The
 * products of DeviceProperty._cal_tc_ / _cal_dos_ / _cal_ldos_
 * (dpnegf/negf/device_property.py:404, 428-438, 451-461) and of the density
 * integrand (dpnegf/negf/density.py:193) run through it.  B <= 65535. */
int dpnegf_zgemm_batched(void* stream, int B, int m, int n, int k, const void* A, int lda,
                         int64_t strideA, int opA, const void* Bm, int ldb, int64_t strideB, int opB,
                         double alpha_re, double alpha_im, double beta_re, double beta_im, void* C,
                         int ldc, int64_t strideC);

/* In place: A[b] <- A[b]^-1 for B contiguous row-major n x n complex128
 * blocks, 1 <= n <= 128 (Gauss-Jordan with partial pivoting, one workgroup
 * per block).  status[b] (optional, [B]) gets bit 1 or-ed in for an exactly
 * singular pivot.  The block inverse of the reference's RGF sweep,
 * tLA.solve(-M, I) (dpnegf/negf/recursive_green_cal.py:89-93, 199-206). */
int dpnegf_zinv_batched(void* stream, void* A, int n, int B, int32_t* status);

/* out[b][i] (+)= sum_j A[b][i][j] B[b][j][i]: the diagonal of A B (A m x n,
 * B n x m, contiguous rows), for traces and per-orbital spectra
 * (dpnegf/negf/device_property.py:404, 438, 461). */
int dpnegf_diag_product(void* stream, void* out, const void* A, int64_t strideA, const void* Bm,
                        int64_t strideB, int m, int n, int B, int accumulate);

#ifdef __cplusplus
}
#endif
#endif
