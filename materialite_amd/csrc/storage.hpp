// Storage views of the fields the matvec passes stream through HBM.
//
// A pass kernel reads and writes its spectrum and its Lanczos vectors through
// a view whose storage type S is float, double or f48.  f48 holds an FP32 value
// plus a bf16 word with the rounded remainder (value - FP32 value): 24 + 8
// significant bits in 6 bytes, with FP32's exponent range.  A value read back
// differs from the one stored by at most 2^-33 of it (relative), against 2^-24
// for FP32 -- enough that a MINRES solve whose Lanczos vectors and spectrum
// are kept this way follows the FP64 solve's iterates to ~1e-9.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace mtl {

struct f48 {};

// tangent and MINRES w / x vectors of a storage form: FP32 unless FP64
template <typename S>
struct work_of {
  using type = S;
};
template <>
struct work_of<f48> {
  using type = float;
};
template <typename S>
using work_t = typename work_of<S>::type;

// ---------------------------------------------------------------- f48 codec
__host__ __device__ inline uint16_t bf16_bits_rne(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf16_value(uint32_t b16) {
  const uint32_t u = b16 << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline void f48_split(double v, float& hi, uint16_t& lo) {
  hi = float(v);
  lo = bf16_bits_rne(float(v - double(hi)));
}
__host__ __device__ inline double f48_join(float hi, uint32_t lo16) {
  return double(hi) + double(bf16_value(lo16 & 0xffffu));
}

// ---------------------------------------------------------------- real vectors
template <typename S>
struct RVec {
  S* p = nullptr;
  template <typename T>
  __device__ __forceinline__ T get(long long i) const {
    return T(p[i]);
  }
  template <typename T>
  __device__ __forceinline__ void put(long long i, T v) const {
    p[i] = S(v);
  }
};

template <>
struct RVec<f48> {
  float* hi = nullptr;
  uint16_t* lo = nullptr;
  template <typename T>
  __device__ __forceinline__ T get(long long i) const {
    return T(f48_join(hi[i], lo[i]));
  }
  template <typename T>
  __device__ __forceinline__ void put(long long i, T v) const {
    float h;
    uint16_t l;
    f48_split(double(v), h, l);
    hi[i] = h;
    lo[i] = l;
  }
};

// ---------------------------------------------------------------- complex spectrum
// Unit: what one load of a complex value brings into registers (converted to
// the arithmetic type when it is put into LDS).
template <typename S>
struct CVec {
  using Unit = cplx<S>;
  cplx<S>* p = nullptr;
  static constexpr bool kPairs = sizeof(S) == 4;  // 16 B accesses of two values
  __device__ __forceinline__ CVec at(long long o) const { return {p + o}; }
  __device__ __forceinline__ Unit load(long long i) const { return p[i]; }
  template <typename T>
  __device__ __forceinline__ static cplx<T> value(const Unit& u) {
    return {T(u.x), T(u.y)};
  }
  template <typename T>
  __device__ __forceinline__ cplx<T> get(long long i) const {
    return value<T>(p[i]);
  }
  template <typename T>
  __device__ __forceinline__ void put(long long i, const cplx<T>& v) const {
    p[i] = {S(v.x), S(v.y)};
  }
  // values i and i + 1 (i even, FP32 storage only)
  template <typename T>
  __device__ __forceinline__ void get_pair(long long i, cplx<T>* dst) const {
    const float4 v = *reinterpret_cast<const float4*>(p + i);
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<float4*>(dst) = v;
    } else {
      dst[0] = {T(v.x), T(v.y)};
      dst[1] = {T(v.z), T(v.w)};
    }
  }
  template <typename T>
  __device__ __forceinline__ void put_pair(long long i, const cplx<T>* src) const {
    if constexpr (sizeof(T) == 4) {
      *reinterpret_cast<float4*>(p + i) = *reinterpret_cast<const float4*>(src);
    } else {
      *reinterpret_cast<float4*>(p + i) =
          make_float4(float(src[0].x), float(src[0].y), float(src[1].x), float(src[1].y));
    }
  }
};

template <>
struct CVec<f48> {
  struct Unit {
    cplx<float> h;
    uint32_t l;  // bf16 remainders: real part in the low half, imaginary in the high
  };
  cplx<float>* hi = nullptr;
  uint32_t* lo = nullptr;
  static constexpr bool kPairs = true;  // 16 B of values + 8 B of remainders
  __device__ __forceinline__ CVec at(long long o) const { return {hi + o, lo + o}; }
  __device__ __forceinline__ Unit load(long long i) const { return {hi[i], lo[i]}; }
  template <typename T>
  __device__ __forceinline__ static cplx<T> value(const Unit& u) {
    return {T(f48_join(u.h.x, u.l)), T(f48_join(u.h.y, u.l >> 16))};
  }
  template <typename T>
  __device__ __forceinline__ cplx<T> get(long long i) const {
    return value<T>(load(i));
  }
  template <typename T>
  __device__ __forceinline__ void put(long long i, const cplx<T>& v) const {
    float hx, hy;
    uint16_t lx, ly;
    f48_split(double(v.x), hx, lx);
    f48_split(double(v.y), hy, ly);
    hi[i] = {hx, hy};
    lo[i] = uint32_t(lx) | (uint32_t(ly) << 16);
  }
  template <typename T>
  __device__ __forceinline__ void get_pair(long long i, cplx<T>* dst) const {
    const float4 v = *reinterpret_cast<const float4*>(hi + i);
    const uint2 r = *reinterpret_cast<const uint2*>(lo + i);
    dst[0] = {T(f48_join(v.x, r.x)), T(f48_join(v.y, r.x >> 16))};
    dst[1] = {T(f48_join(v.z, r.y)), T(f48_join(v.w, r.y >> 16))};
  }
  template <typename T>
  __device__ __forceinline__ void put_pair(long long i, const cplx<T>* src) const {
    float h[4];
    uint16_t l[4];
    f48_split(double(src[0].x), h[0], l[0]);
    f48_split(double(src[0].y), h[1], l[1]);
    f48_split(double(src[1].x), h[2], l[2]);
    f48_split(double(src[1].y), h[3], l[3]);
    *reinterpret_cast<float4*>(hi + i) = make_float4(h[0], h[1], h[2], h[3]);
    *reinterpret_cast<uint2*>(lo + i) =
        make_uint2(uint32_t(l[0]) | (uint32_t(l[1]) << 16), uint32_t(l[2]) | (uint32_t(l[3]) << 16));
  }
};

// device bytes of n complex values / n real values in storage S
template <typename S>
constexpr size_t cplx_bytes(long long n) {
  if constexpr (std::is_same<S, f48>::value)
    return size_t(n) * (sizeof(cplx<float>) + sizeof(uint32_t));
  else
    return size_t(n) * sizeof(cplx<S>);
}
template <typename S>
constexpr size_t real_bytes(long long n) {
  if constexpr (std::is_same<S, f48>::value)
    return size_t(n) * (sizeof(float) + sizeof(uint16_t));
  else
    return size_t(n) * sizeof(S);
}

// views over one block of device memory (hi parts first, then remainders)
template <typename S>
inline CVec<S> cvec_over(void* base, long long n) {
  if constexpr (std::is_same<S, f48>::value) {
    CVec<f48> v;
    v.hi = static_cast<cplx<float>*>(base);
    v.lo = reinterpret_cast<uint32_t*>(v.hi + n);
    return v;
  } else {
    return CVec<S>{static_cast<cplx<S>*>(base)};
  }
}
template <typename S>
inline RVec<S> rvec_over(void* base, long long n) {
  if constexpr (std::is_same<S, f48>::value) {
    RVec<f48> v;
    v.hi = static_cast<float*>(base);
    v.lo = reinterpret_cast<uint16_t*>(v.hi + n);
    return v;
  } else {
    return RVec<S>{static_cast<S*>(base)};
  }
}

}  // namespace mtl
