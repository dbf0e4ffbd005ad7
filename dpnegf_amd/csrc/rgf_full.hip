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
#include <cstring>
#include <vector>

#include "rgf_kernels.hpp"
#include "rgf_full.hpp"

namespace negf {

__device__ __forceinline__ cx<double> opnd_at(const Opnd& o, int b, int i, int k,
                                              cx<double> zb) {
  int r = i, c = k;
  if (o.op == OP_H || o.op == OP_T) { r = k; c = i; }
  const int64_t idx = (int64_t)b * o.bstride + (int64_t)r * o.ld + c;
  cx<double> v = o.h[idx];
  if (o.s) v = v - zb * o.s[(int64_t)b * o.sstride + (int64_t)r * o.ld + c];
  if (o.op == OP_H || o.op == OP_C) v = conj(v);
  return v;
}

// 64 x 64 output tile per workgroup, 4 waves x 4 tiles of 16 x 16.
__global__ __launch_bounds__(256) void bgemm_kernel(BGemmArgs a) {
  const int tn = (a.n + 63) / 64;
  const int ti = blockIdx.x / tn, tj = blockIdx.x % tn;
  const int b = blockIdx.y;
  const int i0 = ti * 64, j0 = tj * 64;
  const int mm = min(64, a.m - i0), nn = min(64, a.n - j0);
  const cx<double> zb = a.z ? a.z[b] : cx<double>{0, 0};
  cx<double>* C = a.C + (int64_t)b * a.cstride;
  const Opnd A = a.A, Bo = a.B_;
  const cx<double> alpha = a.alpha, beta = a.beta;
  const bool use_beta = (beta.re != 0.0 || beta.im != 0.0);
  wg_gemm<double, 4, 4>(
      mm, nn, a.k, [&](int i, int k) { return opnd_at(A, b, i0 + i, k, zb); },
      [&](int k, int j) { return opnd_at(Bo, b, k, j0 + j, zb); },
      [&](int i, int j, cx<double> v) {
        cx<double>* c = C + (int64_t)(i0 + i) * a.ldc + (j0 + j);
        cx<double> out = alpha * v;
        if (use_beta) out = out + beta * (*c);
        *c = out;
      });
}

template <int NMAX, int NT>
__global__ __launch_bounds__(NT) void binv_kernel(cx<double>* A, int n, int64_t stride,
                                                  int32_t* status) {
  constexpr int MAXR = gj_maxr(NMAX, NT);
  __shared__ GJScratch<double, NMAX> gjs;
  cx<double>* M = A + (int64_t)blockIdx.x * stride;
  const int s = wg_gj_inverse<double, NT, NMAX, MAXR>(
      n, [&](int i, int k) { return M[i * n + k]; },
      [&](int i, int k, cx<double> v) { M[i * n + k] = v; }, &gjs);
  if (threadIdx.x == 0 && status && s) atomicOr(&status[blockIdx.x], STATUS_SINGULAR);
}

// D[b] = -(H - z_b S) - corner(SigmaL) - corner(SigmaR)
__global__ void bbuild_kernel(cx<double>* D, const cx<double>* H, const cx<double>* S, int n,
                              const cx<double>* z, int B, const cx<double>* sL, int nL,
                              const cx<double>* sR, int nR) {
  const int64_t tot = (int64_t)B * n * n;
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < tot;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(x / ((int64_t)n * n));
    const int e = (int)(x % ((int64_t)n * n));
    const int i = e / n, k = e % n;
    cx<double> v = H[e] - z[b] * S[e];
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


// out[i][k] = sum_b conj(in[b][k][i]);  in: [B, m, n], out: [n, m]
__global__ void breduce_conjt_kernel(cx<double>* out, const cx<double>* in, int m, int n, int B) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * n) return;
  const int k = e / n, i = e % n;  // in element (k, i)
  cx<double> acc{0, 0};
  for (int b = 0; b < B; ++b) acc = acc + in[(int64_t)b * m * n + e];
  out[(int64_t)i * m + k] = conj(acc);
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

static hipError_t gemm(hipStream_t st, int B, int m, int n, int k, Opnd A, Opnd Bo,
                       cx<double> alpha, cx<double> beta, cx<double>* C, int64_t cstride,
                       int ldc, const cx<double>* z);
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

hipError_t weighted_re_sum(double* out, const cx<double>* in, const cx<double>* c, int64_t count,
                           int B, int accumulate, hipStream_t st) {
  const int grid = (int)std::min<int64_t>((count + 255) / 256, 65535);
  hipLaunchKernelGGL(bweighted_re_sum_kernel, dim3(grid), dim3(256), 0, st, out, in, c, count, B,
                     accumulate);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// host side

static hipError_t gemm(hipStream_t st, int B, int m, int n, int k, Opnd A, Opnd Bo,
                       cx<double> alpha, cx<double> beta, cx<double>* C, int64_t cstride,
                       int ldc, const cx<double>* z) {
  if (m <= 0 || n <= 0) return hipSuccess;
  BGemmArgs a;
  a.m = m; a.n = n; a.k = k; a.B = B;
  a.A = A; a.B_ = Bo; a.alpha = alpha; a.beta = beta;
  a.C = C; a.cstride = cstride; a.ldc = ldc; a.z = z;
  dim3 grid(((m + 63) / 64) * ((n + 63) / 64), B);
  hipLaunchKernelGGL(bgemm_kernel, grid, dim3(256), 0, st, a);
  return hipGetLastError();
}

static hipError_t inverse(hipStream_t st, cx<double>* A, int n, int B, int32_t* status) {
  const int64_t s = (int64_t)n * n;
  if (n <= 16) hipLaunchKernelGGL((binv_kernel<16, 256>), dim3(B), dim3(256), 0, st, A, n, s, status);
  else if (n <= 32) hipLaunchKernelGGL((binv_kernel<32, 256>), dim3(B), dim3(256), 0, st, A, n, s, status);
  else if (n <= 64) hipLaunchKernelGGL((binv_kernel<64, 256>), dim3(B), dim3(256), 0, st, A, n, s, status);
  else if (n <= 128) hipLaunchKernelGGL((binv_kernel<128, 1024>), dim3(B), dim3(1024), 0, st, A, n, s, status);
  else if (n <= GJ_NMAX_LARGE) hipLaunchKernelGGL((binv_kernel<GJ_NMAX_LARGE, 1024>), dim3(B), dim3(1024), 0, st, A, n, s, status);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

static Opnd mat(const cx<double>* h, int ld, int64_t bstride, int op = OP_N,
                const cx<double>* s = nullptr, int64_t sstride = 0);
static Opnd mat(const cx<double>* h, int ld, int64_t bstride, int op, const cx<double>* s,
                int64_t sstride) {
  Opnd o;
  o.h = h; o.s = s; o.bstride = bstride; o.sstride = sstride; o.ld = ld; o.op = op;
  return o;
}

#define CHK(x)                                  \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) return e_;            \
  } while (0)

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
                       z, B, left ? a.sigL : nullptr, a.nL, right ? a.sigR : nullptr, a.nR);
    return hipGetLastError();
  };

  // ---- forward: left-connected g_q  (:198-207)
  CHK(build(0, true, K == 1));
  CHK(inverse(st, gl(0), n[0], B, a.status));
  for (int q = 0; q + 1 < K; ++q) {
    const int nq = n[q], n1 = n[q + 1];
    CHK(gemm(st, B, nq, n1, nq, mat(gl(q), nq, (int64_t)nq * nq), Uop(q), one, zero, w0,
             (int64_t)nq * n1, n1, z));
    CHK(build(q + 1, false, q + 1 == K - 1));
    CHK(gemm(st, B, n1, n1, nq, Lop(q), mat(w0, n1, (int64_t)nq * n1), mone, one, gl(q + 1),
             (int64_t)n1 * n1, n1, z));
    CHK(inverse(st, gl(q + 1), n1, B, a.status));
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

  // ---- lesser (:241-264) and greater (:267-289)
  for (int pass = 0; pass < 2; ++pass) {
    const cx<double>* sin = pass == 0 ? a.s_in : a.s_out;
    cx<double>* gin = pass == 0 ? a.gin : a.gip;
    cx<double>* gnd = pass == 0 ? a.gnd : a.gpd;
    cx<double>* gnl = pass == 0 ? a.gnl : a.gpl;
    cx<double>* gnu = pass == 0 ? a.gnu : a.gpu;
    if (!sin) continue;
    // lesser conjugates with ^H, the greater quirk with plain conj
    const int cj = pass == 0 ? OP_H : OP_C;
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
      CHK(gemm(st, B, n1, n1, nq, mat(w0, nq, (int64_t)n1 * nq), Uop(q, pass == 0 ? OP_N : OP_C), one,
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
      CHK(gemm(st, B, n1, nq, nq, Lop(q, pass == 0 ? OP_N : OP_C), mat(gl(q), nq, sq, cj), one, zero,
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
      CHK(gemm(st, B, nq, n1, nq, mat(gi(q), nq, sq), Uop(q, pass == 0 ? OP_N : OP_C), one, zero, w1,
               sx, n1, z));
      if (pass == 0)  // ... gru[q]^H  (n1 x nq operand from gru stored nq x n1)
        CHK(gemm(st, B, nq, nq, n1, mat(w1, n1, sx), mat(gru(q), n1, sx, OP_H), one, one, nd, sq, nq, z));
      else            // ... conj(grl[q])  (grl stored n1 x nq)
        CHK(gemm(st, B, nq, nq, n1, mat(w1, n1, sx), mat(grl(q), nq, (int64_t)n1 * nq, OP_C), one, one,
                 nd, sq, nq, z));
      CHK(gemm(st, B, nq, nq, n1, mat(gru(q), n1, sx), Lop(q), one, zero, w1, sq, nq, z));
      CHK(gemm(st, B, nq, nq, nq, mat(w1, nq, sq), mat(gi(q), nq, sq), one, one, nd, sq, nq, z));
      // gnu = gnl^H
      {
        const int64_t tot = (int64_t)B * nq * n1;
        const int grid = (int)std::min<int64_t>((tot + 255) / 256, 65535);
        hipLaunchKernelGGL(bconjt_kernel, dim3(grid), dim3(256), 0, st, gnu + ou[q], nl, n1, nq, B);
        CHK(hipGetLastError());
      }
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
  return B * (2 * nsq + 2 * std::max<int64_t>(nx, 1) + nlc + m2 + 5 * m2 + 3 * m2 + 7 * m2);
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
    hipLaunchKernelGGL(breduce_conjt_kernel, dim3((m_ * n_ + 255) / 256), dim3(256), 0, st, out,
                       in, m_, n_, B);
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
    CHK(gemm(st, B, nq, n1, nK, mat(lc(q), nK, (int64_t)nq * nK), mat(D[1 - cur], n1, (int64_t)nK * n1),
             one, zero, Wm, (int64_t)nq * n1, n1, g.z));
    CHK(reduce(g.grad_hl + gx[q], Wm, nq, n1));
    cur = 1 - cur;
  }
  return hipSuccess;
}

}  // namespace negf
