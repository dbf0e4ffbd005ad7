// Shared device-side helpers for the NEGF kernels (gfx950, wave64).
//
// Complex numbers are stored interleaved (re, im), the storage of torch's
// complex128 / complex64, so tensors are passed zero-copy.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace negf {

constexpr int WAVE = 64;

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> mk(R a, R b) { return cx<R>{a, b}; }
template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a) { return {-a.re, -a.im}; }
template <typename R>
__device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cx<R> conj(cx<R> a) { return {a.re, -a.im}; }
template <typename R>
__device__ __forceinline__ R abs1(cx<R> a) { return fabs(a.re) + fabs(a.im); }
template <typename R>
__device__ __forceinline__ R cabs(cx<R> a) { return sqrt(a.re * a.re + a.im * a.im); }
template <typename R, typename S>
__device__ __forceinline__ cx<R> cvt(cx<S> a) { return {R(a.re), R(a.im)}; }

// ---- matrix-core wrappers: D(16x16) += A(16x4) * B(4x16) ------------------------
// Operand layout (both precisions): lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Accumulator element r of lane l holds D[row(l/16, r)][l%16]:
// rows interleaved (l/16 + 4r) for the f64 instruction (measured on gfx950),
// consecutive (4(l/16) + r) for f32.
template <typename R> struct mfma;
template <> struct mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

// ---- wave / workgroup reductions -------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Maximum of a 32-bit unsigned value over a full wave (all 64 lanes active),
// returned wave-uniform.  Four DPP steps leave every lane with the maximum of
// its 16-lane row (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror); the four row maxima are then combined through readlane and
// scalar maxima (left to itself the compiler combined them with a vector
// max3 and a readfirstlane, on the pivot search's critical path).  No LDS
// round trip, unlike __shfl_xor (ds_bpermute).
template <int CTRL>
__device__ __forceinline__ unsigned dpp_max_step(unsigned v) {
  const unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
  return v > o ? v : o;
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = dpp_max_step<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_max_step<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_max_step<0x141>(v);  // row_half_mirror
  v = dpp_max_step<0x140>(v);  // row_mirror
  const unsigned a = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane((int)v, 16);
  const unsigned c = (unsigned)__builtin_amdgcn_readlane((int)v, 32);
  const unsigned d = (unsigned)__builtin_amdgcn_readlane((int)v, 48);
  unsigned m;
  asm("s_max_u32 %0, %1, %2\n\ts_max_u32 %0, %0, %3\n\ts_max_u32 %0, %0, %4"
      : "=&s"(m) : "s"(a), "s"(b), "s"(c), "s"(d) : "scc");
  return m;
}

// Sum over the workgroup; `red` is LDS scratch of >= nwaves entries.  All
// threads get the result.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
template <typename T>
__device__ __forceinline__ T block_max(T v, T* red) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int nw = (blockDim.x + WAVE - 1) / WAVE;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = red[0];
  for (int i = 1; i < nw; ++i) s = fmax(s, red[i]);
  return s;
}

}  // namespace negf
