// Argument blocks of the batched block-operation path (full RGF outputs).
#pragma once
#include "common.hpp"

namespace negf {

enum { OP_N = 0, OP_H = 1, OP_T = 2, OP_C = 3 };
// One-workgroup batched Gauss-Jordan inverse up to GJ_NMAX_LARGE (a 1024-thread
// register-resident sweep; beyond 128 its column stripes spill), recursive
// Schur-complement inverse over diagonal halves above that, up to BLOCK_NMAX,
// and by default from SCHUR_NMIN on wherever the caller gives a workspace.
constexpr int GJ_NMAX_LARGE = 128;
constexpr int BLOCK_NMAX = 1024;
constexpr int SCHUR_NMIN = 97;
// Workspace of the batched inverse for n x n blocks (0 below SCHUR_NMIN).
int64_t inverse_workspace(int n, int B);
// In-place inverse of B contiguous n x n blocks, 1 <= n <= GJ_NMAX_LARGE.
hipError_t zinv_batched(hipStream_t st, cx<double>* A, int n, int B, int32_t* status);

// Process-wide tuning table read by the batched product and inverse drivers
// when they pick a kernel shape.  TUNE_GEMM_TILE keys (m, n, k), choices
// TILE_32x32 .. TILE_64x64 (output tile rows x columns per workgroup);
// TUNE_INVERSE keys (n, 0, 0), choices INV_REGISTER (register-resident
// Gauss-Jordan, n <= 128), INV_LDS (blocked Gauss-Jordan in LDS, n <= 96) and
// INV_SCHUR (Schur complement over two halves plus a Newton step, 97..128).
// Choice 0 clears an entry.  With recording on, every (kind, m, n, k) a driver
// looks up is collected (in first-seen order) for the autotuner to time.
enum { TUNE_GEMM_TILE = 0, TUNE_INVERSE = 1 };
enum { TILE_32x32 = 1, TILE_32x64 = 2, TILE_64x32 = 3, TILE_64x64 = 4 };
enum { INV_REGISTER = 1, INV_LDS = 2, INV_SCHUR = 3 };
int tune_set(int kind, int m, int n, int k, int choice);  // 0, or -1 for a bad argument
int tune_get(int kind, int m, int n, int k);              // 0 when not set
int tune_builtin(int kind, int m, int n, int k);          // the choice without an entry (0: bad key)
void tune_clear();
void tune_record(bool on);                                // on: also forgets earlier shapes
int tune_recorded(int32_t* out, int cap);  // writes min(count, cap) x 4 ints, returns count

// Operand value (i, k) = op(h[b] - z_b s[b]) with s optional, row-major ld.
struct Opnd {
  const cx<double>* h;
  const cx<double>* s;
  int64_t bstride;
  int64_t sstride;
  int ld;
  int op;
};

struct BGemmArgs {
  int m, n, k, B;
  Opnd A, B_;
  cx<double> alpha, beta;
  cx<double>* C;
  const cx<double>* Cin;  // source of the beta term, laid out as C (often C itself)
  const int32_t* skip;    // [B] or null: batch entries with skip[b] != 0 are left untouched
  int64_t cstride;
  int ldc;
  const cx<double>* z;
};

struct RgfFullArgs {
  int K, B, nmax, ortho;
  const int32_t* bsz;          // host [K]
  const cx<double>* arena;     // device, one system
  const cx<double>* z;         // device [B]
  const cx<double>* sigL; int nL;
  const cx<double>* sigR; int nR;
  const cx<double>* seP;       // [sum_q B n_q^2] energy-dependent diagonal addend, or null
  cx<double>* gl;              // [sum_q B n_q^2]          left-connected g_q
  cx<double>* grd;             // [sum_q B n_q^2]
  cx<double>* grl;             // [sum_q B n_q+1 n_q]
  cx<double>* gru;             // [sum_q B n_q n_q+1]
  cx<double>* lc;              // [sum_q B n_q n_K-1] or null
  cx<double>* g_trans;         // [B n_0 n_K-1]
  const cx<double>* s_in;      // lesser source, [sum_q B n_q^2] or null
  cx<double>* gin; cx<double>* gnd; cx<double>* gnl; cx<double>* gnu;
  const cx<double>* s_out;     // greater source or null
  cx<double>* gip; cx<double>* gpd; cx<double>* gpl; cx<double>* gpu;
  const cx<double>* s_bil;     // bilinear G S G source or null
  cx<double>* gbi; cx<double>* gbd; cx<double>* gbl; cx<double>* gbu;
  cx<double>* work;            // 5 * B * nmax^2
  cx<double>* inv_ws;          // inverse_workspace(nmax, B) (null when that is 0)
  int32_t* status;             // [B] or null
};

hipError_t rgf_full(const RgfFullArgs& a, hipStream_t st);

struct RgfGradArgs {
  int K, B, ortho;
  const int32_t* bsz;          // host [K]
  const cx<double>* arena;
  const cx<double>* z;
  const cx<double>* sigL; int nL;
  const cx<double>* sigR; int nR;
  const double* w;             // [B] weights of T(E_b) in the objective
  cx<double>* grad_hd;         // [sum_q n_q^2]
  cx<double>* grad_hu;         // [sum_q n_q n_q+1]
  cx<double>* grad_hl;         // [sum_q n_q+1 n_q]
  cx<double>* grad_sigL;       // [B, nL, nL] per energy, or null
  cx<double>* grad_sigR;       // [B, nR, nR] per energy, or null
  cx<double>* grad_sd;         // [sum_q n_q^2], or null (the overlap blocks)
  cx<double>* grad_su;         // [sum_q n_q n_q+1], or null
  cx<double>* grad_sl;         // [sum_q n_q+1 n_q], or null
  cx<double>* work;            // rgf_grad_workspace elements
  int32_t* status;
};

// Adjoint with sqrt(K)-style checkpointing every `stride` blocks (same outputs
// as rgf_grad; work: rgf_grad_ckpt_workspace elements, -1 for stride < 1).
int64_t rgf_grad_ckpt_workspace(int K, const int32_t* bsz, int B, int stride);
hipError_t rgf_grad_ckpt(const RgfGradArgs& g, int stride, hipStream_t st);

hipError_t diag_product(cx<double>* out, const cx<double>* A, int64_t sa, const cx<double>* Bm,
                        int64_t sb, int m, int n, int B, int accumulate, hipStream_t st);
hipError_t zgemm_batched(int B, int m, int n, int k, const cx<double>* A, int lda, int64_t sa,
                         int opA, const cx<double>* Bm, int ldb, int64_t sb, int opB,
                         cx<double> alpha, cx<double> beta, cx<double>* C, int ldc, int64_t sc,
                         hipStream_t st);
// out (+)= (sum_b in[b])^H for B contiguous m x n blocks; out is n x m.
hipError_t batch_sum_conjt(cx<double>* out, const cx<double>* in, int m, int n, int B,
                           int accumulate, hipStream_t st);
hipError_t weighted_re_sum(double* out, const cx<double>* in, const cx<double>* c, int64_t count,
                           int B, int accumulate, hipStream_t st);
int64_t rgf_grad_workspace(int K, const int32_t* bsz, int B);

struct RgfOpsArgs {
  int K, B, ortho;
  const int32_t* bsz;          // host [K]
  const cx<double>* arena;
  const cx<double>* z;         // [B]
  const cx<double>* sigL; int nL;
  const cx<double>* sigR; int nR;
  double* T;                   // [B]
  cx<double>* work;            // rgf_transmission_ops_workspace elements
  int32_t* status;
};
struct LeadArgs;
int64_t lead_ls_batched_workspace(int n, int nred, int B);
hipError_t lead_ls_batched(const LeadArgs& a, cx<double>* work, hipStream_t st);
int64_t rgf_transmission_ops_workspace(int K, const int32_t* bsz, int B);
hipError_t rgf_transmission_ops(const RgfOpsArgs& a, hipStream_t st);
hipError_t rgf_grad(const RgfGradArgs& a, hipStream_t st);

}  // namespace negf
