// Shared device/host helpers for the materialite_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace mtl {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define MTL_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::mtl::Error(std::string("HIP error '") + hipGetErrorString(e_) +  \
                         "' at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define MTL_LAUNCH_CHECK() MTL_HIP(hipGetLastError())

// message returned by mtl_last_error() on this thread
void set_last_error(const std::string& msg);

// ------------------------------------------------------------------ complex
template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return {a.x * s, a.y * s};
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return {a.x, -a.y};
}

// ------------------------------------------------------------------ launch shape
// MI355X dispatches workgroups round-robin over its 8 XCDs (each with its own
// L2).  Remapping blockIdx so that XCD j receives a contiguous range of logical
// blocks keeps neighbouring tiles (which share cache lines) in one L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nblocks) {
  const int nx = 8;
  if (nblocks % nx != 0) return b;
  return (b % nx) * (nblocks / nx) + b / nx;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also waits
// for every outstanding global load of the wave; the pass kernels keep the
// next tile's loads in flight across their FFT barriers, so they synchronise
// with this one (fences restricted to the local address space).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// block-wide sum of a double, result valid in thread 0
template <int BT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BT / 64; ++i) r += red[i];
  }
  __syncthreads();
  return r;
}

template <int BT>
__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    r = red[0];
#pragma unroll
    for (int i = 1; i < BT / 64; ++i) r = fmax(r, red[i]);
  }
  __syncthreads();
  return r;
}

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// Tangent layout: the 21 packed upper-triangle entries of each point, blocked
// by 64 points: entry k of point p sits at ((p / 64) * 21 + k) * 64 + p % 64.
// A wavefront of 64 consecutive points then reads its whole tangent as one
// contiguous 21 x 256 B run (FP32) instead of 21 streams far apart.
__host__ __device__ inline long long tan_idx(int k, long long p) {
  return ((p >> 6) * 21 + k) * 64 + (p & 63);
}
// entries to allocate for n points (the last block padded to 64)
inline long long tan_elems(long long n) { return 21 * ((n + 63) / 64) * 64; }

// The same 64-point blocking for a per-point array of C components (the EVP
// state): component c of point i at ((i / 64) * C + c) * 64 + i % 64, so that
// the components of point i sit 64 apart from blk_idx(0, i, C).
__host__ __device__ inline long long blk_idx(int c, long long i, int C) {
  return ((i >> 6) * C + c) * 64 + (i & 63);
}
inline long long blk_elems(long long n, int C) { return C * ((n + 63) / 64) * 64; }

// Copy a small parameter block from device memory into LDS with all threads of
// the workgroup (the caller issues the barrier).  Large uniform parameter sets
// passed by value live in SGPRs and spill; reading them from LDS avoids that.
template <typename P>
__device__ __forceinline__ void stage_to_lds(P* dst, const P* src) {
  static_assert(sizeof(P) % sizeof(int) == 0, "parameter block must be int-sized");
  const int* s = reinterpret_cast<const int*>(src);
  int* d = reinterpret_cast<int*>(dst);
  for (int k = threadIdx.x; k < int(sizeof(P) / sizeof(int)); k += blockDim.x) d[k] = s[k];
}

}  // namespace mtl
