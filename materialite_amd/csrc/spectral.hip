// FFT passes of the projection apply and of the MINRES matvec (gfx950).
//
// Data layout: real fields SoA [6][n]; tangent [n/64][21][64] packed upper
// triangle; spectrum [6][nx][ny][nzc] complex.  One matvec is five kernels:
//   z_forward  : D:(s r2) (or a + sgn*in), paired real->complex z-FFT
//   y_pass     : complex y-FFT of kz tiles (forward)
//   x_ghat     : complex x-FFT, Ghat applied per wave vector, inverse x-FFT
//   y_pass     : inverse y-FFT
//   z_inverse  : complex->real z-FFT, 1/n scaling, MINRES epilogue or store
// Ghat:X = a(x)n + n(x)a - (n.a) n(x)n with a = X.n, n = t/|t|,
// t_j = tan(pi q_j) (rotated finite-difference frequencies of
// small_strain_fft.py:301-328); zero on even-N Nyquist planes (:354-359) and
// diag(mask) at k = 0 (:363-367).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "profiler.hpp"
#include "spectral.hpp"

namespace mtl {

namespace {

// threads per workgroup of each pass
constexpr int kZBT = 256, kYBT = 256, kXBT = 256;
// elements of one component tile per thread in the x pass (nx * TX <= kXTileK * kXBT)
constexpr int kXTileK = 8;
// LDS bytes of the x pass's two single-component buffers (doubled for FP64)
constexpr size_t kXPairBytes = 40 * 1024;
// LDS bytes per FP32 workgroup for the two ping-pong line buffers (doubled for
// FP64 arithmetic, whatever the storage type).  Small tiles keep many workgroups resident per CU (a CU has 160 KB),
// which hides the HBM latency of the strided tile loads: measured at 128^3,
// these sizes beat 32-96 KB tiles by 25-40 %.
constexpr size_t kZBufBytes = 16 * 1024, kYBufBytes = 16 * 1024, kXBufBytes = 24 * 1024;
constexpr size_t kMaxLds = 160 * 1024;
// LDS of an in-place x-pass tile: three workgroups per CU
constexpr size_t kXInplaceBytes = 53 * 1024;
// LDS of an in-place y-pass tile: four workgroups per CU
constexpr size_t kYInplaceBytes = 40 * 1024;

// LDS row stride of the z passes: 3 interleaved lines (paired components) per
// grid line, rounded up to odd
__host__ __device__ inline int zrow_stride(int P) { return (3 * P) | 1; }
// LDS row stride of the x pass for TX interleaved lines: odd, so that the
// butterfly stores of a wavefront spread over the LDS banks
__host__ __device__ inline int odd_stride(int M) { return M | 1; }
// LDS bytes of the x pass: twiddles and two single-component buffers
inline size_t x_lds_bytes(int nx, int TX, size_t cb) {
  return (size_t(nx) + 2 * size_t(nx) * size_t(odd_stride(TX))) * cb;
}
// complex values in each of the two z-pass buffers: the FFT rows, and in
// z_inverse also the staged half spectra of the P lines (6 components)
__host__ __device__ inline int zbuf_elems(int P, int nz, int nzc) {
  const int a = zrow_stride(P) * nz, b = 6 * P * nzc;
  return a > b ? a : b;
}

enum { kModeProject = 0, kModeMatvec = 1 };


template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T>* dst, const cplx<T>* src, int L) {
  for (int i = threadIdx.x; i < L; i += blockDim.x) dst[i] = src[i];
}

// Storage type TS (the spectrum and Krylov vectors in HBM) and arithmetic type
// T (LDS, twiddles, butterflies) may differ: FP32 data transformed in FP64.
template <typename T, typename TS>
__device__ __forceinline__ cplx<T> cvt(const cplx<TS>& v) {
  return {T(v.x), T(v.y)};
}
// ----------------------------------------------------------------- z forward
template <typename T, typename S, int MODE, typename In, bool BIG>
__global__ void __launch_bounds__(kZBT) k_z_forward(Grid g, const AxisPlan* plan_g, const cplx<T>* twz,
                                                     int P, In in, const work_t<S>* tangent,
                                                     Vec6 a, double sgn,
                                                     const MinresState* st, CVec<S> spec,
                                                     const BatchSlot<S>* slots) {
  if constexpr (MODE == kModeMatvec) {
    if (slots != nullptr) {
      const BatchSlot<S>& sl = slots[blockIdx.y];
      in = sl.r2;
      tangent = sl.tangent;
      st = sl.st;
      spec = sl.spec;
    }
    if (st->done) return;
  }
  extern __shared__ unsigned char smem_raw[];
  const AxisPlan* plan = plan_g;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + g.nz;
  const int nz = g.nz, nzc = g.nzc;
  const int M = 3 * P, LS = zrow_stride(P), zb = zbuf_elems(P, nz, nzc);
  const long long nlines = (long long)g.nx * g.ny;
  const long long line0 = (long long)blockIdx.x * P;
  load_twiddles(tw, twz, nz);

  T s = T(1);
  if (MODE == kModeMatvec) s = T(st->s);
  for (int i = threadIdx.x; i < P * nz; i += kZBT) {
    const int li = i / nz, z = i - li * nz;
    const long long gl = line0 + li;
    T tau[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (gl < nlines) {
      const long long p = gl * nz + z;
      if constexpr (MODE == kModeProject) {
#pragma unroll
        for (int c = 0; c < 6; ++c) tau[c] = T(a.v[c] + sgn * double(in[c * g.n + p]));
      } else {
        T v[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) v[c] = s * in.template get<T>(blk_idx(c, p, 6));
#pragma unroll
        for (int i2 = 0; i2 < 6; ++i2) {
          T acc = T(0);
#pragma unroll
          for (int j = 0; j < 6; ++j) acc += T(tangent[tan_idx(sym6(i2, j), p)]) * v[j];
          tau[i2] = acc;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) buf[z * LS + li * 3 + q] = {tau[2 * q], tau[2 * q + 1]};
  }
  __syncthreads();
  const cplx<T>* res = block_fft<T, kZBT, BIG>(buf, buf + zb, M, LS, plan, tw, false);

  const T half = T(0.5);
  for (int i = threadIdx.x; i < P * 3 * nzc; i += kZBT) {
    const int m = i / nzc, kz = i - m * nzc;  // m = li * 3 + q
    const int li = m / 3, q = m - 3 * li;
    const long long gl = line0 + li;
    if (gl >= nlines) continue;
    const cplx<T> zf = res[kz * LS + m];
    const cplx<T> zr = cconj(res[((nz - kz) % nz) * LS + m]);
    const cplx<T> xa = {half * (zf.x + zr.x), half * (zf.y + zr.y)};
    const cplx<T> xb = {half * (zf.y - zr.y), -half * (zf.x - zr.x)};
    spec.put((2 * q) * g.nspec + gl * nzc + kz, xa);
    spec.put((2 * q + 1) * g.nspec + gl * nzc + kz, xb);
  }
}

// complex values a thread holds per in-place FFT pass (block_fft_inplace)
template <typename T>
constexpr int x_inplace_regs() {
  return sizeof(T) == 8 ? 16 : 32;
}

// ----------------------------------------------------------------- y pass
// INPLACE: the FFT passes run in the tile's one LDS buffer
// (block_fft_inplace), which allows twice the tile width for the same LDS.
template <typename T, typename S, bool BIG, bool INPLACE = false>
__global__ void __launch_bounds__(kYBT) k_y_pass(Grid g, const AxisPlan* plan_g, const cplx<T>* twy, int TY,
                                                  int ntiles, bool inv,
                                                  const MinresState* st, CVec<S> spec,
                                                  const BatchSlot<S>* slots) {
  if (slots != nullptr) {
    st = slots[blockIdx.y].st;
    spec = slots[blockIdx.y].spec;
  }
  if (st != nullptr && st->done) return;
  extern __shared__ unsigned char smem_raw[];
  const AxisPlan* plan = plan_g;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + g.ny;
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tile = b % ntiles;
  const int cx = b / ntiles;  // c * nx + x
  const int kz0 = tile * TY;
  const int width = min(TY, g.nzc - kz0);
  const int ny = g.ny;
  const CVec<S> base = spec.at((long long)cx * ny * g.nzc + kz0);  // spec[c][x][.][kz0]
  load_twiddles(tw, twy, ny);
  // FP32-stored tiles of whole, even-aligned columns move two complex values
  // (16 B) per access (even ny keeps the LDS tile after the ny twiddles 16 B
  // aligned)
  const bool pairs = CVec<S>::kPairs && TY % 2 == 0 && width == TY && g.nzc % 2 == 0 && ny % 2 == 0;
  const int hp = TY / 2;
  if constexpr (CVec<S>::kPairs) {
    if (pairs) {
      for (int i = threadIdx.x; i < hp * ny; i += kYBT) {
        const int y = i / hp, t = 2 * (i - y * hp);
        base.get_pair((long long)y * g.nzc + t, buf + y * TY + t);
      }
    }
  }
  if (!pairs) {
    for (int i = threadIdx.x; i < TY * ny; i += kYBT) {
      const int y = i / TY, t = i - y * TY;
      buf[y * TY + t] = t < width ? base.template get<T>((long long)y * g.nzc + t) : cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  const cplx<T>* res = buf;
  if constexpr (INPLACE)
    block_fft_inplace<T, kYBT, x_inplace_regs<T>()>(buf, TY, TY, plan, tw, inv);
  else
    res = block_fft<T, kYBT, BIG>(buf, buf + TY * ny, TY, TY, plan, tw, inv);
  if constexpr (CVec<S>::kPairs) {
    if (pairs) {
      for (int i = threadIdx.x; i < hp * ny; i += kYBT) {
        const int y = i / hp, t = 2 * (i - y * hp);
        base.put_pair((long long)y * g.nzc + t, res + y * TY + t);
      }
    }
  }
  if (!pairs) {
    for (int i = threadIdx.x; i < TY * ny; i += kYBT) {
      const int y = i / TY, t = i - y * TY;
      if (t < width) base.put((long long)y * g.nzc + t, res[y * TY + t]);
    }
  }
}

// ----------------------------------------------------------------- x pass + Ghat
template <typename T>
__device__ __forceinline__ void apply_ghat(cplx<T> X[6], T n0, T n1, T n2) {
  const T r2 = T(1.41421356237309504880168872420969807857);
  const T ir2 = T(0.70710678118654752440084436210484903928);
  cplx<T> a0, a1, a2;
  a0.x = X[0].x * n0 + ir2 * (X[5].x * n1 + X[4].x * n2);
  a0.y = X[0].y * n0 + ir2 * (X[5].y * n1 + X[4].y * n2);
  a1.x = X[1].x * n1 + ir2 * (X[5].x * n0 + X[3].x * n2);
  a1.y = X[1].y * n1 + ir2 * (X[5].y * n0 + X[3].y * n2);
  a2.x = X[2].x * n2 + ir2 * (X[4].x * n0 + X[3].x * n1);
  a2.y = X[2].y * n2 + ir2 * (X[4].y * n0 + X[3].y * n1);
  const cplx<T> na = {a0.x * n0 + a1.x * n1 + a2.x * n2, a0.y * n0 + a1.y * n1 + a2.y * n2};
  // S_ij = a_i n_j + n_i a_j - na n_i n_j
  X[0] = {T(2) * a0.x * n0 - na.x * n0 * n0, T(2) * a0.y * n0 - na.y * n0 * n0};
  X[1] = {T(2) * a1.x * n1 - na.x * n1 * n1, T(2) * a1.y * n1 - na.y * n1 * n1};
  X[2] = {T(2) * a2.x * n2 - na.x * n2 * n2, T(2) * a2.y * n2 - na.y * n2 * n2};
  X[3] = {r2 * (a1.x * n2 + a2.x * n1 - na.x * n1 * n2), r2 * (a1.y * n2 + a2.y * n1 - na.y * n1 * n2)};
  X[4] = {r2 * (a0.x * n2 + a2.x * n0 - na.x * n0 * n2), r2 * (a0.y * n2 + a2.y * n0 - na.y * n0 * n2)};
  X[5] = {r2 * (a0.x * n1 + a1.x * n0 - na.x * n0 * n1), r2 * (a0.y * n1 + a1.y * n0 - na.y * n0 * n1)};
}

template <typename T>
__device__ __forceinline__ T rnorm3(T a, T b, T c);
template <>
__device__ __forceinline__ float rnorm3<float>(float a, float b, float c) {
  return 1.0f / sqrtf(a * a + b * b + c * c);
}
template <>
__device__ __forceinline__ double rnorm3<double>(double a, double b, double c) {
  return 1.0 / sqrt(a * a + b * b + c * c);
}

// Tile = (y, TX consecutive kz columns) with all six components resident in
// LDS: forward x-FFT of the 6 TX interleaved lines, Ghat on each point, inverse
// x-FFT, one HBM read and one write of the tile.  Used where TX >= 2 fits the
// budget (grids up to about 128 points along x).  INPLACE runs the FFT passes
// in the tile's one buffer (block_fft_inplace), which allows twice the tile
// width for the same LDS: each (component, x) row of the tile is then a wider
// contiguous run in HBM, and the pass is bound by the number of such rows.

template <typename T, typename S, bool BIG, bool INPLACE = false>
__global__ void __launch_bounds__(kXBT) k_x_ghat_resident(Grid g, const AxisPlan* plan_g, const cplx<T>* twx, int TX,
                                                  int ntiles, const T* tanx, const T* tany,
                                                  const T* tanz, int maskbits,
                                                  const MinresState* st, CVec<S> spec,
                                                  const BatchSlot<S>* slots) {
  if (slots != nullptr) {
    st = slots[blockIdx.y].st;
    spec = slots[blockIdx.y].spec;
    maskbits = slots[blockIdx.y].maskbits;
  }
  if (st != nullptr && st->done) return;
  extern __shared__ unsigned char smem_raw[];
  const AxisPlan* plan = plan_g;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + g.nx;
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tile = b % ntiles;
  const int y = b / ntiles;
  const int kz0 = tile * TX;
  const int width = min(TX, g.nzc - kz0);
  const int nx = g.nx;
  const long long xstride = (long long)g.ny * g.nzc;
  const CVec<S> base = spec.at((long long)y * g.nzc + kz0);  // spec[c][0][y][kz0]
  load_twiddles(tw, twx, nx);
  const int tot = 6 * nx * TX;
  // FP32-stored tiles of whole, even-aligned columns move two complex values
  // (16 B) per access (even nx keeps the LDS tile after the nx twiddles 16 B
  // aligned)
  const bool pairs = CVec<S>::kPairs && TX % 2 == 0 && width == TX && g.nzc % 2 == 0 && nx % 2 == 0;
  const int hp = TX / 2;
  if constexpr (CVec<S>::kPairs) {
    if (pairs) {
      for (int i = threadIdx.x; i < tot / 2; i += kXBT) {
        const int cxp = i / hp;  // x*6 + c
        const int t = 2 * (i - cxp * hp);
        const int x = cxp / 6, c = cxp - x * 6;
        base.get_pair(c * g.nspec + x * xstride + t, buf + x * (6 * TX) + c * TX + t);
      }
    }
  }
  if (!pairs) {
    for (int i = threadIdx.x; i < tot; i += kXBT) {
      const int t = i % TX;
      const int cxp = i / TX;  // x*6 + c
      const int x = cxp / 6, c = cxp - x * 6;
      buf[x * (6 * TX) + c * TX + t] =
          t < width ? base.template get<T>(c * g.nspec + x * xstride + t) : cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  const int nyqx = (nx % 2 == 0) ? nx / 2 : -1;
  const int nyqy = (g.ny % 2 == 0) ? g.ny / 2 : -1;
  const int nyqz = (g.nz % 2 == 0) ? g.nz / 2 : -1;  // never stored, kept for clarity
  const T ty = tany[y];
  cplx<T>* buf2 = buf + 6 * TX * nx;  // the second LDS buffer (ping-pong form only)
  cplx<T>* res = buf;
  if constexpr (INPLACE)
    block_fft_inplace<T, kXBT, x_inplace_regs<T>()>(buf, 6 * TX, 6 * TX, plan, tw, false);
  else
    res = block_fft<T, kXBT, BIG>(buf, buf2, 6 * TX, 6 * TX, plan, tw, false);
  for (int i = threadIdx.x; i < nx * TX; i += kXBT) {
    const int t = i % TX, kx = i / TX;
    const int kz = kz0 + t;
    if (t >= width) continue;
    cplx<T> X[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) X[c] = res[kx * (6 * TX) + c * TX + t];
    if (kx == nyqx || y == nyqy || kz == nyqz) {
#pragma unroll
      for (int c = 0; c < 6; ++c) X[c] = {T(0), T(0)};
    } else if (kx == 0 && y == 0 && kz == 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) X[c] = (maskbits >> c) & 1 ? X[c] : cplx<T>{T(0), T(0)};
    } else {
      const T tx = tanx[kx], tz = tanz[kz];
      const T inv = rnorm3<T>(tx, ty, tz);
      apply_ghat<T>(X, tx * inv, ty * inv, tz * inv);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) res[kx * (6 * TX) + c * TX + t] = X[c];
  }
  __syncthreads();
  const cplx<T>* out = res;
  if constexpr (INPLACE)
    block_fft_inplace<T, kXBT, x_inplace_regs<T>()>(res, 6 * TX, 6 * TX, plan, tw, true);
  else
    out = block_fft<T, kXBT, BIG>(res, res == buf ? buf2 : buf, 6 * TX, 6 * TX, plan, tw, true);
  if constexpr (CVec<S>::kPairs) {
    if (pairs) {
      for (int i = threadIdx.x; i < tot / 2; i += kXBT) {
        const int cxp = i / hp;
        const int t = 2 * (i - cxp * hp);
        const int x = cxp / 6, c = cxp - x * 6;
        base.put_pair(c * g.nspec + x * xstride + t, out + x * (6 * TX) + c * TX + t);
      }
    }
  }
  if (!pairs) {
    for (int i = threadIdx.x; i < tot; i += kXBT) {
      const int t = i % TX;
      const int cxp = i / TX;
      const int x = cxp / 6, c = cxp - x * 6;
      if (t < width) base.put(c * g.nspec + x * xstride + t, out[x * (6 * TX) + c * TX + t]);
    }
  }
}

// Tile = (y, TX consecutive kz columns of 128 B where the grid allows), for
// grids too long along x for six resident components at TX >= 2.  The
// six components go through the LDS one at a time (forward x-FFT, stored back
// in place), Ghat then runs on the tile's points straight from memory (the
// tile was just written by this workgroup and is L2-resident), and the six
// components are inverse-transformed one at a time.  While component c is
// transformed, the registers already hold the loads of component c + 1.
// Keeping one component in LDS (instead of all six) keeps tiles 128 B wide and
// four workgroups per CU.
// With PAIRS (FP32, even TX dividing nzc, even nx) a thread moves two
// adjacent complex values (16 B) per access and the LDS rows have stride TX.
template <typename T, typename S, bool PAIRS, bool BIG>
__global__ void __launch_bounds__(kXBT) k_x_ghat_streamed(Grid g, const AxisPlan* plan_g, const cplx<T>* twx, int TX,
                                                  int ntiles, const T* tanx, const T* tany,
                                                  const T* tanz, int maskbits,
                                                  const MinresState* st, CVec<S> spec,
                                                  const BatchSlot<S>* slots) {
  static_assert(!PAIRS || (sizeof(T) == 4 && std::is_same<S, float>::value),
                "pairs of complex values are 16 B in FP32 only");
  if (slots != nullptr) {
    st = slots[blockIdx.y].st;
    spec = slots[blockIdx.y].spec;
    maskbits = slots[blockIdx.y].maskbits;
  }
  if (st != nullptr && st->done) return;
  extern __shared__ unsigned char smem_raw[];
  const AxisPlan* plan = plan_g;
  const int nx = g.nx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int LS = PAIRS ? TX : odd_stride(TX);
  cplx<T>* buf = tw + nx;
  cplx<T>* buf2 = buf + nx * LS;
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tile = b % ntiles;
  const int y = b / ntiles;
  const int kz0 = tile * TX;
  const int width = min(TX, g.nzc - kz0);
  const long long xstride = (long long)g.ny * g.nzc;
  const CVec<S> base = spec.at((long long)y * g.nzc + kz0);  // spec[c][0][y][kz0]
  load_twiddles(tw, twx, nx);
  // per-thread units of one component tile (a complex value, or a pair of
  // adjacent ones with PAIRS): (x, t) with t fastest
  constexpr int W = PAIRS ? 2 : 1;        // complex values per unit
  constexpr int K = kXTileK / W;          // units per thread
  using Unit = typename std::conditional<PAIRS, float4, cplx<T>>::type;   // in LDS
  using GUnit = typename std::conditional<PAIRS, float4, typename CVec<S>::Unit>::type;  // loaded
  const int TU = TX / W;                  // units per row
  const int E = nx * TU;
  int goff[K], loff[K];
  bool on[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * kXBT;
    const int x = i / TU, t = W * (i - x * TU);
    goff[k] = x * int(xstride) + t;
    loff[k] = x * LS + t;
    on[k] = i < E && t < width;
  }
  GUnit pre[K];
  auto fetch = [&](int c) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if constexpr (PAIRS) {
        const float4* src = reinterpret_cast<const float4*>(base.p + c * g.nspec);
        pre[k] = on[k] ? src[goff[k] / W] : GUnit{};
      } else {
        pre[k] = on[k] ? base.load(c * g.nspec + goff[k]) : GUnit{};
      }
    }
  };
  auto to_lds = [](const GUnit& u) -> Unit {
    if constexpr (PAIRS) return u; else return CVec<S>::template value<T>(u);
  };
  auto transform = [&](bool inv) {
    fetch(0);
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (threadIdx.x + k * kXBT < E) reinterpret_cast<Unit*>(buf)[loff[k] / W] = to_lds(pre[k]);
      lds_barrier();
      if (c < 5) fetch(c + 1);
      const cplx<T>* res = block_fft<T, kXBT, BIG>(buf, buf2, TX, LS, plan, tw, inv);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!on[k]) continue;
        if constexpr (PAIRS)
          reinterpret_cast<float4*>(base.p + c * g.nspec)[goff[k] / W] =
              reinterpret_cast<const float4*>(res)[loff[k] / W];
        else
          base.put(c * g.nspec + goff[k], res[loff[k]]);
      }
      lds_barrier();
    }
  };
  transform(false);
  __syncthreads();  // the tile's transformed components are visible to every wave
  const int nyqx = (nx % 2 == 0) ? nx / 2 : -1;
  const int nyqy = (g.ny % 2 == 0) ? g.ny / 2 : -1;
  const int nyqz = (g.nz % 2 == 0) ? g.nz / 2 : -1;  // never stored, kept for clarity
  const T ty = tany[y];
  for (int i = threadIdx.x; i < nx * TX; i += kXBT) {
    const int t = i % TX, kx = i / TX;
    const int kz = kz0 + t;
    if (t >= width) continue;
    const long long pt = kx * xstride + t;
    cplx<T> X[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) X[c] = base.template get<T>(pt + c * g.nspec);
    if (kx == nyqx || y == nyqy || kz == nyqz) {
#pragma unroll
      for (int c = 0; c < 6; ++c) X[c] = {T(0), T(0)};
    } else if (kx == 0 && y == 0 && kz == 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) X[c] = (maskbits >> c) & 1 ? X[c] : cplx<T>{T(0), T(0)};
    } else {
      const T tx = tanx[kx], tz = tanz[kz];
      const T inv = rnorm3<T>(tx, ty, tz);
      apply_ghat<T>(X, tx * inv, ty * inv, tz * inv);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) base.put(pt + c * g.nspec, X[c]);
  }
  __syncthreads();
  transform(true);
}

// ----------------------------------------------------------------- z inverse
template <typename T, typename S, int MODE, bool BIG>
__global__ void __launch_bounds__(kZBT) k_z_inverse(Grid g, const AxisPlan* plan_g, const cplx<T>* twz,
                                                     int P, CVec<S> spec, double* out,
                                                     RVec<S> r1, RVec<S> r2,
                                                     const MinresState* st, double* partials,
                                                     const BatchSlot<S>* slots) {
  if constexpr (MODE == kModeMatvec) {
    if (slots != nullptr) {
      const BatchSlot<S>& sl = slots[blockIdx.y];
      spec = sl.spec;
      r1 = sl.r1;
      r2 = sl.r2;
      st = sl.st;
      partials = sl.partials;
    }
    if (st->done) return;
  }
  extern __shared__ unsigned char smem_raw[];
  const AxisPlan* plan = plan_g;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + g.nz;
  __shared__ double red[kZBT / 64];
  const int nz = g.nz, nzc = g.nzc;
  const int M = 3 * P, LS = zrow_stride(P), zb = zbuf_elems(P, nz, nzc);
  const long long nlines = (long long)g.nx * g.ny;
  const long long line0 = (long long)blockIdx.x * P;
  load_twiddles(tw, twz, nz);
  // stage the half spectra of the block's lines (contiguous per component) in
  // the second buffer, reading each value from HBM once
  cplx<T>* stage = buf + zb;
  const int run = P * nzc;
  const long long valid = (nlines - line0 < P ? nlines - line0 : P) * nzc;
  bool staged = false;
  if constexpr (CVec<S>::kPairs) {
    if (nzc % 2 == 0 && nz % 2 == 0) {
      // FP32 storage, even nzc: two adjacent values per access (16 B; even nz
      // keeps the LDS stage 16 B aligned)
      const int hr = run / 2;
      for (int i = threadIdx.x; i < 6 * hr; i += kZBT) {
        const int c = i / hr, r = 2 * (i - c * hr);
        if (r < valid) {
          spec.get_pair(c * g.nspec + line0 * nzc + r, stage + c * run + r);
        } else {
          stage[c * run + r] = {T(0), T(0)};
          stage[c * run + r + 1] = {T(0), T(0)};
        }
      }
      staged = true;
    }
  }
  if (!staged) {
    for (int i = threadIdx.x; i < 6 * run; i += kZBT) {
      const int c = i / run, r = i - c * run;
      stage[i] = r < valid ? spec.template get<T>(c * g.nspec + line0 * nzc + r) : cplx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * 3 * nz; i += kZBT) {
    const int m = i / nz, k = i - m * nz;  // m = li * 3 + q
    const int li = m / 3, q = m - 3 * li;
    cplx<T> zv = {T(0), T(0)};
    {
      const cplx<T>* sa = stage + (2 * q) * run + li * nzc;
      const cplx<T>* sb = stage + (2 * q + 1) * run + li * nzc;
      if (k == 0) {
        zv = {sa[0].x, sb[0].x};
      } else if (k < nzc) {
        const cplx<T> xa = sa[k], xb = sb[k];
        zv = {xa.x - xb.y, xa.y + xb.x};
      } else if (!(nz % 2 == 0 && k == nz / 2)) {
        const cplx<T> xa = cconj(sa[nz - k]), xb = cconj(sb[nz - k]);
        zv = {xa.x - xb.y, xa.y + xb.x};
      }
    }
    buf[k * LS + li * 3 + q] = zv;
  }
  __syncthreads();
  const cplx<T>* res = block_fft<T, kZBT, BIG>(buf, buf + zb, M, LS, plan, tw, true);

  const T invn = T(1.0 / double(g.n));
  double acc = 0.0;
  T s = T(1), c1 = T(0);
  if constexpr (MODE == kModeMatvec) {
    s = T(st->s);
    c1 = T(st->c_r1);
  }
  for (int i = threadIdx.x; i < P * nz; i += kZBT) {
    const int li = i / nz, z = i - li * nz;
    const long long gl = line0 + li;
    if (gl >= nlines) continue;
    const long long p = gl * nz + z;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const cplx<T> zv = res[z * LS + li * 3 + q];
      const T va = zv.x * invn, vb = zv.y * invn;
      if constexpr (MODE == kModeProject) {
        out[(2 * q) * g.n + p] = double(va);
        out[(2 * q + 1) * g.n + p] = double(vb);
      } else {
        const long long ia = blk_idx(2 * q, p, 6), ib = blk_idx(2 * q + 1, p, 6);
        const T ua = va - c1 * r1.template get<T>(ia);
        const T ub = vb - c1 * r1.template get<T>(ib);
        r1.put(ia, ua);
        r1.put(ib, ub);
        acc += double(s * r2.template get<T>(ia)) * double(ua) +
               double(s * r2.template get<T>(ib)) * double(ub);
      }
    }
  }
  if constexpr (MODE == kModeMatvec) {
    const double tot = block_sum<kZBT>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
  }
}

template <typename T>
std::vector<cplx<T>> twiddle_table(int L) {
  std::vector<cplx<T>> tw(L);
  for (int m = 0; m < L; ++m) {
    const double ang = -2.0 * M_PI * double(m) / double(L);
    tw[m] = {T(std::cos(ang)), T(std::sin(ang))};
  }
  return tw;
}

template <typename T>
std::vector<T> tan_table(int n, int count) {
  // t = tan(pi q) with q = fftfreq(n) (small_strain_fft.py:303-313)
  std::vector<T> t(count);
  for (int k = 0; k < count; ++k) {
    const int kk = (k <= (n - 1) / 2) ? k : k - n;
    t[k] = T(std::tan(M_PI * double(kk) / double(n)));
  }
  return t;
}

template <typename U>
U* upload(const std::vector<U>& v) {
  U* d = nullptr;
  MTL_HIP(hipMalloc(&d, v.size() * sizeof(U)));
  MTL_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(U), hipMemcpyHostToDevice));
  return d;
}

// bit c set when strain component c is controlled (mask != 0)
int mask_bits(const Vec6& mask) {
  int b = 0;
  for (int c = 0; c < 6; ++c)
    if (mask.v[c] != 0.0) b |= 1 << c;
  return b;
}

// Raise a kernel's dynamic-LDS limit once to the largest size it is launched with.
template <typename K>
void allow_lds(K kernel, size_t bytes) {
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> granted;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lock(mu);
  auto it = granted.find(key);
  if (it != granted.end() && bytes <= it->second) return;
  MTL_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
  granted[key] = bytes;
}

}  // namespace

template <typename S, typename A>
int Spectral<S, A>::mask_bits(const Vec6& mask) {
  return ::mtl::mask_bits(mask);
}

template <typename S, typename A>
void Spectral<S, A>::plan(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) throw Error("grid dimensions must be positive");
  g_.nx = nx;
  g_.ny = ny;
  g_.nz = nz;
  g_.nzc = (nz + 1) / 2;
  g_.n = (long long)nx * ny * nz;
  g_.nspec = (long long)nx * ny * g_.nzc;
  px_ = make_axis_plan(nx);
  py_ = make_axis_plan(ny);
  pz_ = make_axis_plan(nz);
  big_x_ = uses_big_radix(px_);
  big_y_ = uses_big_radix(py_);
  big_z_ = uses_big_radix(pz_);
  const size_t cb = sizeof(cplx<A>);  // LDS element
  const size_t scale = sizeof(A) / sizeof(float);
  // FP32 data with FP64 arithmetic: the z passes do best with 1.5x the FP64
  // budget (measured at 128^3: z_forward 20.8 -> 19.7 us, z_inverse 24.5 -> 23.0 us;
  // the y and x passes are slower with larger tiles)
  constexpr bool kNarrow = !std::is_same<S, double>::value;  // storage narrower than FP64
  const size_t zscale = (sizeof(A) == 8 && kNarrow) ? 3 : scale;
  const size_t zbudget = kZBufBytes * zscale, ybudget = kYBufBytes * scale,
               xbudget = kXBufBytes * scale;
  // z passes: P grid lines (3P interleaved complex lines) per workgroup
  if ((size_t(nz) + 2 * size_t(zbuf_elems(1, nz, g_.nzc))) * cb > kMaxLds)
    throw Error("nz too large for the z pass");
  pz_lines_ = int(std::max<size_t>(1, zbudget / (2 * 3 * size_t(nz) * cb)));
  while (pz_lines_ > 1 && 2 * size_t(zbuf_elems(pz_lines_, nz, g_.nzc)) * cb > zbudget)
    --pz_lines_;
  const long long nlines = (long long)nx * ny;
  if (pz_lines_ > nlines) pz_lines_ = int(nlines);
  // y pass: ty_ kz columns per workgroup
  if ((size_t(ny) + 2 * size_t(ny)) * cb > kMaxLds) throw Error("ny too large for the y pass");
  ty_ = int(std::max<size_t>(1, std::min<size_t>({16, size_t(g_.nzc), ybudget / (2 * ny * cb)})));
  // FP64 arithmetic: in-place FFT passes, twice the tile width (up to 32
  // columns) in at most 40 KB (at 128^3: 46.7 -> 43.9 us refined, 54.5 -> 53.4
  // us FP64; with FP32 arithmetic the register-staged passes measured slower,
  // 25.3 -> 28.1 us, so the FP32 y pass keeps the ping-pong form)
  y_inplace_ = false;
  if (!big_y_ && sizeof(A) == 8) {
    for (int t = 32; t > ty_; t /= 2) {
      const int tw = std::min(t, g_.nzc);
      if ((size_t(ny) + size_t(tw) * ny) * cb <= kYInplaceBytes &&
          inplace_fits(py_, tw, kYBT, x_inplace_regs<A>())) {
        ty_ = tw;
        y_inplace_ = true;
        break;
      }
    }
  }
  // x pass: tx_ kz columns of all 6 components resident per workgroup where
  // at least two fit the budget; otherwise the components stream through LDS
  // one at a time with tx_ the widest (up to 128 B rows) tile whose two
  // single-component buffers fit
  tx_ = int(std::min<size_t>({8, size_t(g_.nzc), xbudget / (2 * 6 * nx * cb)}));
  x_resident_ = tx_ >= std::min(2, g_.nzc);
  // in-place FFT passes: the same LDS holds twice the tile width (up to 8
  // columns, three 53 KB workgroups per CU) where every pass fits the
  // registers a thread holds
  x_inplace_ = false;
  if (x_resident_ && !big_x_) {
    for (int t = 8; t > tx_; t /= 2) {
      const int tw = std::min(t, g_.nzc);
      if ((size_t(nx) + 6 * size_t(tw) * nx) * cb <= kXInplaceBytes &&
          inplace_fits(px_, 6 * tw, kXBT, x_inplace_regs<A>())) {
        tx_ = tw;
        x_inplace_ = true;
        break;
      }
    }
  }
  if (!x_resident_) {
    tx_ = 0;
    for (int t = int(128 / cb); t >= 1; t /= 2)
      if (x_lds_bytes(nx, t, cb) <= kXPairBytes * scale &&
          size_t(nx) * t <= size_t(kXTileK) * kXBT) {
        tx_ = t;
        break;
      }
    if (tx_ == 0 && x_lds_bytes(nx, 1, cb) <= kMaxLds && size_t(nx) <= size_t(kXTileK) * kXBT)
      tx_ = 1;
    if (tx_ == 0) throw Error("nx too large for the x pass");
    tx_ = std::min(tx_, g_.nzc);
    // FP32 with whole, even-width tiles: pairs of complex values per access
    x_pairs_ = std::is_same<S, float>::value && sizeof(A) == 4 && tx_ % 2 == 0 &&
               g_.nzc % tx_ == 0 && nx % 2 == 0;
  }
}

template <typename S, typename A>
Spectral<S, A>::Spectral(int nx, int ny, int nz) {
  plan(nx, ny, nz);
  MTL_HIP(hipMalloc(&spec_mem_, cplx_bytes<S>(6 * g_.nspec)));
  spec_ = cvec_over<S>(spec_mem_, 6 * g_.nspec);
  twx_ = upload(twiddle_table<A>(nx));
  twy_ = upload(twiddle_table<A>(ny));
  twz_ = upload(twiddle_table<A>(nz));
  pxd_ = upload(std::vector<AxisPlan>{px_});
  pyd_ = upload(std::vector<AxisPlan>{py_});
  pzd_ = upload(std::vector<AxisPlan>{pz_});
  tanx_ = upload(tan_table<A>(nx, nx));
  tany_ = upload(tan_table<A>(ny, ny));
  tanz_ = upload(tan_table<A>(nz, g_.nzc));
}

namespace {
// bytes of each buffer the spectral passes read, in workspace order, each
// rounded up to 256 B
template <typename S, typename A>
void workspace_layout(int nx, int ny, int nz, size_t sz[10]) {
  const long long nzc = (nz + 1) / 2;
  const long long nspec = (long long)nx * ny * nzc;
  const size_t raw[10] = {cplx_bytes<S>(6 * nspec),
                          size_t(nx) * sizeof(cplx<A>), size_t(ny) * sizeof(cplx<A>),
                          size_t(nz) * sizeof(cplx<A>), sizeof(AxisPlan), sizeof(AxisPlan),
                          sizeof(AxisPlan), size_t(nx) * sizeof(A), size_t(ny) * sizeof(A),
                          size_t(nzc) * sizeof(A)};
  for (int i = 0; i < 10; ++i) sz[i] = (raw[i] + 255) / 256 * 256;
}
}  // namespace

template <typename S, typename A>
size_t Spectral<S, A>::workspace_bytes(int nx, int ny, int nz) {
  size_t sz[10], tot = 0;
  workspace_layout<S, A>(nx, ny, nz, sz);
  for (size_t v : sz) tot += v;
  return tot;
}

template <typename S, typename A>
Spectral<S, A>::Spectral(int nx, int ny, int nz, void* workspace, size_t bytes, hipStream_t s) {
  plan(nx, ny, nz);
  external_ = true;
  if (workspace == nullptr || (reinterpret_cast<uintptr_t>(workspace) & 255) != 0)
    throw Error("spectral workspace must be a 256-byte aligned device address");
  if (bytes < workspace_bytes(nx, ny, nz)) throw Error("spectral workspace too small");
  size_t sz[10];
  workspace_layout<S, A>(nx, ny, nz, sz);
  unsigned char* w = static_cast<unsigned char*>(workspace);
  void* slot[10];
  for (int i = 0; i < 10; ++i) {
    slot[i] = w;
    w += sz[i];
  }
  spec_ = cvec_over<S>(slot[0], 6 * g_.nspec);
  twx_ = static_cast<cplx<A>*>(slot[1]);
  twy_ = static_cast<cplx<A>*>(slot[2]);
  twz_ = static_cast<cplx<A>*>(slot[3]);
  pxd_ = static_cast<AxisPlan*>(slot[4]);
  pyd_ = static_cast<AxisPlan*>(slot[5]);
  pzd_ = static_cast<AxisPlan*>(slot[6]);
  tanx_ = static_cast<A*>(slot[7]);
  tany_ = static_cast<A*>(slot[8]);
  tanz_ = static_cast<A*>(slot[9]);
  // the host tables live until the stream has copied them
  const auto tx = twiddle_table<A>(nx), ty = twiddle_table<A>(ny), tz = twiddle_table<A>(nz);
  const auto ax = tan_table<A>(nx, nx), ay = tan_table<A>(ny, ny), az = tan_table<A>(nz, g_.nzc);
  auto put = [&](void* dst, const void* src, size_t n) {
    MTL_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
  };
  put(twx_, tx.data(), tx.size() * sizeof(cplx<A>));
  put(twy_, ty.data(), ty.size() * sizeof(cplx<A>));
  put(twz_, tz.data(), tz.size() * sizeof(cplx<A>));
  put(pxd_, &px_, sizeof(AxisPlan));
  put(pyd_, &py_, sizeof(AxisPlan));
  put(pzd_, &pz_, sizeof(AxisPlan));
  put(tanx_, ax.data(), ax.size() * sizeof(A));
  put(tany_, ay.data(), ay.size() * sizeof(A));
  put(tanz_, az.data(), az.size() * sizeof(A));
  stream_sync(s);
}

template <typename S, typename A>
Spectral<S, A>::~Spectral() {
  if (external_) return;
  for (void* p : {spec_mem_, (void*)twx_, (void*)twy_, (void*)twz_, (void*)tanx_,
                  (void*)tany_, (void*)tanz_, (void*)pxd_, (void*)pyd_, (void*)pzd_})
    if (p) (void)hipFree(p);
}

template <typename S, typename A>
int Spectral<S, A>::zinv_blocks() const {
  return ceil_div((long long)g_.nx * g_.ny, pz_lines_);
}

template <typename S, typename A>
template <int MODE, typename In>
void Spectral<S, A>::run(In in, const W* tangent, double* out, RVec<S> r1, const Vec6& a,
                         double sgn, const Vec6& mask, const MinresState* st, double* partials,
                         hipStream_t s, const BatchSlot<S>* slots, int nb) {
  const int P = pz_lines_;
  const int zblocks = zinv_blocks();
  const size_t zlds = (size_t(g_.nz) + 2 * size_t(zbuf_elems(P, g_.nz, g_.nzc))) * sizeof(cplx<A>);
  auto kzf = big_z_ ? k_z_forward<A, S, MODE, In, true> : k_z_forward<A, S, MODE, In, false>;
  allow_lds(kzf, zlds);
  profiler().launch(MODE == kModeMatvec ? kKZForward : residual_id(kKZForward), kzf, dim3(zblocks, nb), dim3(kZBT), zlds, s, g_, pzd_, twz_, P, in, tangent, a, sgn, st, spec_, slots);
  MTL_LAUNCH_CHECK();

  const MinresState* stp = MODE == kModeMatvec ? st : nullptr;
  const int ytiles = ceil_div(g_.nzc, ty_);
  const size_t ylds = (size_t(g_.ny) + (y_inplace_ ? 1 : 2) * size_t(ty_) * g_.ny) * sizeof(cplx<A>);
  auto ky = big_y_       ? k_y_pass<A, S, true>
            : y_inplace_ ? k_y_pass<A, S, false, true>
                         : k_y_pass<A, S, false>;
  allow_lds(ky, ylds);
  profiler().launch(MODE == kModeMatvec ? kKYPass : residual_id(kKYPass), ky, dim3(6 * g_.nx * ytiles, nb), dim3(kYBT), ylds, s, g_, pyd_, twy_, ty_, ytiles, false, stp, spec_, slots);
  MTL_LAUNCH_CHECK();

  const int xtiles = ceil_div(g_.nzc, tx_);
  const size_t xlds =
      x_inplace_  ? (size_t(g_.nx) + size_t(6 * tx_) * g_.nx) * sizeof(cplx<A>)
      : x_resident_ ? (size_t(g_.nx) + 2 * size_t(6 * tx_) * g_.nx) * sizeof(cplx<A>)
      : x_pairs_  ? (size_t(g_.nx) + 2 * size_t(tx_) * g_.nx) * sizeof(cplx<A>)
                  : x_lds_bytes(g_.nx, tx_, sizeof(cplx<A>));
  auto kx = big_x_     ? k_x_ghat_resident<A, S, true>
            : x_inplace_ ? k_x_ghat_resident<A, S, false, true>
                         : k_x_ghat_resident<A, S, false>;
  if (!x_resident_) {
    if constexpr (std::is_same<S, float>::value && sizeof(A) == 4) {
      if (x_pairs_)
        kx = big_x_ ? k_x_ghat_streamed<A, S, true, true> : k_x_ghat_streamed<A, S, true, false>;
      else
        kx = big_x_ ? k_x_ghat_streamed<A, S, false, true> : k_x_ghat_streamed<A, S, false, false>;
    } else {
      kx = big_x_ ? k_x_ghat_streamed<A, S, false, true> : k_x_ghat_streamed<A, S, false, false>;
    }
  }
  allow_lds(kx, xlds);
  profiler().launch(MODE == kModeMatvec ? kKXGhat : residual_id(kKXGhat), kx, dim3(g_.ny * xtiles, nb), dim3(kXBT), xlds, s, g_, pxd_, twx_, tx_, xtiles, tanx_, tany_, tanz_, mask_bits(mask), stp, spec_, slots);
  MTL_LAUNCH_CHECK();

  profiler().launch(MODE == kModeMatvec ? kKYPass : residual_id(kKYPass), ky, dim3(6 * g_.nx * ytiles, nb), dim3(kYBT), ylds, s, g_, pyd_, twy_, ty_, ytiles, true, stp, spec_, slots);
  MTL_LAUNCH_CHECK();

  RVec<S> r2{};
  if constexpr (MODE == kModeMatvec) r2 = in;
  auto kzi = big_z_ ? k_z_inverse<A, S, MODE, true> : k_z_inverse<A, S, MODE, false>;
  allow_lds(kzi, zlds);
  profiler().launch(MODE == kModeMatvec ? kKZInverse : residual_id(kKZInverse), kzi, dim3(zblocks, nb), dim3(kZBT), zlds, s, g_, pzd_, twz_, P, spec_, out, r1, r2, st, partials, slots);
  MTL_LAUNCH_CHECK();
}

template <typename S, typename A>
void Spectral<S, A>::project(const double* in, double* out, const Vec6& a, double sgn,
                             const Vec6& mask, hipStream_t s) {
  run<kModeProject, const double*>(in, nullptr, out, RVec<S>{}, a, sgn, mask, nullptr, nullptr, s);
}

template <typename S, typename A>
int Spectral<S, A>::matvec_minres(RVec<S> r2, const W* tangent, RVec<S> r1, const Vec6& mask,
                                  const MinresState* st, double* partials, hipStream_t s) {
  Vec6 zero{};
  run<kModeMatvec, RVec<S>>(r2, tangent, nullptr, r1, zero, 1.0, mask, st, partials, s);
  return zinv_blocks();
}

template <typename S, typename A>
int Spectral<S, A>::matvec_batch(const BatchSlot<S>* slots, int nb, hipStream_t s) {
  // per-solve arguments come from the slots; the mask only picks the kernels'
  // shared arguments here, the slots carry each solve's mask bits
  Vec6 zero{};
  run<kModeMatvec, RVec<S>>(RVec<S>{}, nullptr, nullptr, RVec<S>{}, zero, 1.0, zero, nullptr,
                            nullptr, s, slots, nb);
  return zinv_blocks();
}

template class Spectral<float>;
template class Spectral<double>;
template class Spectral<f48, double>;

}  // namespace mtl
