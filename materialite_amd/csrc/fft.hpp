// Workgroup-level mixed-radix Stockham FFT over lines held in LDS.
//
// A batch of M lines of length L is stored element-major: element e of line m
// at buf[e * S + m] (row stride S >= M; an odd S spreads strided accesses
// over the LDS banks).  Each radix-R pass reads the R inputs of a butterfly
// (line m, index j) from one buffer, applies twiddles w^(t*k) (k = j mod Ns),
// performs an R-point DFT and writes the outputs to their autosorted Stockham
// positions (j/Ns)*Ns*R + k + u*Ns of the other buffer (ping-pong, one barrier
// per pass, no register staging).  Butterflies are numbered line-fastest, so
// the lanes of a wavefront touch runs of consecutive LDS words in every pass.
//
// Twiddles come from a per-axis table tw[m] = exp(-2 pi i m / L) copied into
// LDS; the inverse transform uses their conjugates (unnormalised).  The odd
// radices read their R-th roots of unity from the same table (tw[k L/R]):
// held as literals, their FP64 constants would occupy scalar registers in every
// inlined pass and push the kernels into scalar-register spills.
#pragma once

#include "common.hpp"

namespace mtl {

constexpr int kMaxPasses = 24;

struct AxisPlan {
  int L;
  int npass;
  int radix[kMaxPasses];
  int ns[kMaxPasses];
  // ceil(2^32 / ns): j / ns = (j * ns_magic) >> 32 exactly for j, ns < 2^16
  unsigned long long ns_magic[kMaxPasses];
};

// q = n / d for 0 <= n < 2^16 and 1 <= d <= 2^16, with magic = ceil(2^32 / d):
// n * magic / 2^32 exceeds n / d by less than n / 2^32 < 1 / d, which cannot
// carry it past the next integer.
__host__ __device__ inline unsigned long long div_magic(unsigned d) {
  return ((1ULL << 32) + d - 1) / d;
}
__device__ __forceinline__ int fast_div(int n, unsigned long long magic) {
  return int((unsigned long long)(unsigned)n * magic >> 32);
}

// --------------------------------------------------------------- small DFTs
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool, const cplx<T>*, int) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

// multiply by -i (forward) or +i (inverse)
template <typename T>
__device__ __forceinline__ cplx<T> rot_i(cplx<T> a, bool inv) {
  return inv ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

template <typename T>
struct Dft<T, 4> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv) {
    cplx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cplx<T> a2 = cadd(v[1], v[3]), a3 = rot_i(csub(v[1], v[3]), inv);
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  }
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>*, int) {
    run(v, inv);
  }
};

template <typename T>
struct Dft<T, 8> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>*, int) {
    const T h = T(0.70710678118654752440084436210484903928);
    cplx<T> b[4], c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      b[t] = cadd(v[t], v[t + 4]);
      c[t] = csub(v[t], v[t + 4]);
    }
    // c_t *= w8^t, w8 = exp(-+ i pi/4)
    {
      cplx<T> x = c[1];
      c[1] = inv ? cplx<T>{h * (x.x - x.y), h * (x.x + x.y)} : cplx<T>{h * (x.x + x.y), h * (x.y - x.x)};
      c[2] = rot_i(c[2], inv);
      x = c[3];
      c[3] = inv ? cplx<T>{-h * (x.x + x.y), h * (x.x - x.y)} : cplx<T>{h * (x.y - x.x), -h * (x.x + x.y)};
    }
    Dft<T, 4>::run(b, inv);
    Dft<T, 4>::run(c, inv);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      v[2 * m] = b[m];
      v[2 * m + 1] = c[m];
    }
  }
};

// 16 = 4 x 4: DFT4 over n1 of x[4 n1 + n2], twiddle W16^(n2 k1), DFT4 over n2;
// output k1 + 4 k2.  The W16 powers come from the L-point table (rt[m L/16]),
// W16^4 = -i exactly.
template <typename T>
struct Dft<T, 16> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>* rt, int q) {
    cplx<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cplx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      Dft<T, 4>::run(t, inv);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) {
        const int m = n2 * k1;
        if (m == 4) {
          a[n2][k1] = rot_i(a[n2][k1], inv);
        } else {
          cplx<T> w = rt[m * q];
          if (inv) w.y = -w.y;
          a[n2][k1] = cmul(a[n2][k1], w);
        }
      }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cplx<T> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<T, 4>::run(t, inv);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

template <typename T>
struct Dft<T, 3> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>* rt, int q) {
    const T s = -rt[q].y;  // sin(2 pi/3)
    cplx<T> t1 = cadd(v[1], v[2]);
    cplx<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
    cplx<T> d = csub(v[1], v[2]);
    cplx<T> t3 = inv ? cplx<T>{-s * d.y, s * d.x} : cplx<T>{s * d.y, -s * d.x};
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <typename T>
struct Dft<T, 5> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>* rt, int q) {
    const cplx<T> w1 = rt[q], w2 = rt[2 * q];  // exp(-2 pi i k/5)
    const T c1 = w1.x, c2 = w2.x, s1 = -w1.y, s2 = -w2.y;
    cplx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    cplx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    cplx<T> m1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
    cplx<T> m2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
    cplx<T> n1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    cplx<T> n2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    v[0] = {v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
    cplx<T> r1 = rot_i(n1, inv), r2 = rot_i(n2, inv);
    v[1] = cadd(m1, r1);
    v[4] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[3] = csub(m2, r2);
  }
};

template <typename T>
struct Dft<T, 7> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>* rt, int q) {
    // symmetric form: with a_k = v_k + v_{7-k}, b_k = v_k - v_{7-k},
    //   y_u, y_{7-u} = v_0 + sum_k cos(2 pi uk/7) a_k  -+ i sum_k sin(2 pi uk/7) b_k
    const cplx<T> w1 = rt[q], w2 = rt[2 * q], w3 = rt[3 * q];  // exp(-2 pi i k/7)
    const T c1 = w1.x, c2 = w2.x, c3 = w3.x;
    const T s1 = -w1.y, s2 = -w2.y, s3 = -w3.y;
    const cplx<T> a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    const cplx<T> a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    const cplx<T> a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    const cplx<T> x0 = v[0];
    const cplx<T> m1 = {x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x, x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y};
    const cplx<T> m2 = {x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x, x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y};
    const cplx<T> m3 = {x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x, x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y};
    const cplx<T> n1 = {s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y};
    const cplx<T> n2 = {s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y};
    const cplx<T> n3 = {s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y};
    v[0] = {x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y};
    const cplx<T> r1 = rot_i(n1, inv), r2 = rot_i(n2, inv), r3 = rot_i(n3, inv);
    v[1] = cadd(m1, r1);
    v[6] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[5] = csub(m2, r2);
    v[3] = cadd(m3, r3);
    v[4] = csub(m3, r3);
  }
};

// Odd prime R (11, 13): the symmetric form of the radix-7 DFT above written
// for any odd R, with the roots w_m = exp(-2 pi i m / R) read from the table:
//   y_u, y_{R-u} = v_0 + sum_k Re(w_uk) a_k  +- rot(sum_k -Im(w_uk) b_k)
template <typename T, int R>
struct DftOdd {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv, const cplx<T>* rt, int q) {
    constexpr int H = (R - 1) / 2;
    cplx<T> a[H], b[H];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      a[k - 1] = cadd(v[k], v[R - k]);
      b[k - 1] = csub(v[k], v[R - k]);
    }
    const cplx<T> x0 = v[0];
    cplx<T> y0 = x0;
#pragma unroll
    for (int k = 0; k < H; ++k) y0 = cadd(y0, a[k]);
#pragma unroll
    for (int u = 1; u <= H; ++u) {
      cplx<T> m = x0, n = {T(0), T(0)};
#pragma unroll
      for (int k = 1; k <= H; ++k) {
        const cplx<T> w = rt[((u * k) % R) * q];
        m = {m.x + w.x * a[k - 1].x, m.y + w.x * a[k - 1].y};
        n = {n.x - w.y * b[k - 1].x, n.y - w.y * b[k - 1].y};
      }
      const cplx<T> r = rot_i(n, inv);
      v[u] = cadd(m, r);
      v[R - u] = csub(m, r);
    }
    v[0] = y0;
  }
};
template <typename T>
struct Dft<T, 11> : DftOdd<T, 11> {};
template <typename T>
struct Dft<T, 13> : DftOdd<T, 13> {};

// --------------------------------------------------------------- one butterfly
template <typename T, int R>
__device__ __forceinline__ void butterfly(const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst,
                                          int S, int j, int m, int kk, int jq, int per_line,
                                          int Ns, int tstride, const cplx<T>* tw, bool inv) {
  cplx<T> v[R];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    cplx<T> x = src[(j + t * per_line) * S + m];
    if (t > 0 && kk > 0) {
      cplx<T> w = tw[t * kk * tstride];
      if (inv) w.y = -w.y;
      x = cmul(x, w);
    }
    v[t] = x;
  }
  Dft<T, R>::run(v, inv, tw, per_line);
  const int idx = jq * Ns * R + kk;
#pragma unroll
  for (int u = 0; u < R; ++u) dst[(idx + u * Ns) * S + m] = v[u];
}

// A radix-R pass for any odd prime R without a dedicated butterfly: each
// work item computes one output u of one butterfly (line m, index j) as the
// direct sum over its R inputs,
//   y_u = sum_t x_t w^(t kk tstride) w^(t u L/R)  (w = exp(-+2 pi i / L)),
// with the combined twiddle exponent stepped modulo L, and writes it to the
// autosorted position of the Stockham pass.  O(R) per output instead of
// O(log R); used only for the prime factors above 13.
template <typename T, int BT>
__device__ __forceinline__ void generic_pass(const cplx<T>* __restrict__ src,
                                             cplx<T>* __restrict__ dst, int M, int S, int L, int R,
                                             int Ns, unsigned long long ns_magic,
                                             const cplx<T>* tw, bool inv) {
  const int per_line = L / R;
  const int tstride = per_line / Ns;
  for (int w = threadIdx.x; w < L * M; w += BT) {
    const int m = w % M;
    const int q = w / M;  // u * per_line + j
    const int u = q / per_line, j = q - u * per_line;
    const int jq = fast_div(j, ns_magic);
    const int kk = j - jq * Ns;
    int step = kk * tstride + u * per_line;
    if (step >= L) step -= L;
    cplx<T> acc = src[j * S + m];
    int e = 0;
    for (int t = 1; t < R; ++t) {
      e += step;
      if (e >= L) e -= L;
      cplx<T> c = tw[e];
      if (inv) c.y = -c.y;
      acc = cadd(acc, cmul(src[(j + t * per_line) * S + m], c));
    }
    dst[(jq * Ns * R + kk + u * Ns) * S + m] = acc;
  }
}

__host__ __device__ inline bool has_butterfly(int R) {
  return R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8 || R == 11 || R == 13 ||
         R == 16;
}

// Transform M interleaved lines held in `a` (row stride S >= M complex values),
// using `b` as the second buffer.  Returns the buffer that holds the result.
// Ends with a barrier.  The radix switch sits inside the butterfly loop so that
// the per-pass index arithmetic is shared by all radices.  Butterfly
// bf = j * M + m of a thread advances by BT per step; (j, m) is stepped
// incrementally instead of divided out.
//
// BIG adds the radix-11 and radix-13 butterflies; they are compiled only into
// the kernels of axes whose plan uses them, since their registers would
// otherwise lower the occupancy of every pass.
template <typename T, int BT, bool BIG = false>
__device__ __forceinline__ cplx<T>* block_fft(cplx<T>* a, cplx<T>* b, int M, int S,
                                              const AxisPlan* __restrict__ pl, const cplx<T>* tw,
                                              bool inv) {
  const int npass = pl->npass;
  const int L = pl->L;
  const int dj = BT / M, dm = BT - dj * M;
  const int j0 = int(threadIdx.x) / M, m0 = int(threadIdx.x) - j0 * M;
  for (int p = 0; p < npass; ++p) {
    const int R = pl->radix[p];
    const int Ns = pl->ns[p];
    const unsigned long long mg = pl->ns_magic[p];
    const int per_line = L / R;
    const int tstride = per_line / Ns;
    if constexpr (BIG) {
      if (!has_butterfly(R)) {
        generic_pass<T, BT>(a, b, M, S, L, R, Ns, mg, tw, inv);
        lds_barrier();
        cplx<T>* t = a;
        a = b;
        b = t;
        continue;
      }
    }
    int m = m0;
    for (int j = j0; j < per_line;) {
      const int jq = fast_div(j, mg);
      const int kk = j - jq * Ns;
      switch (R) {
        case 16: butterfly<T, 16>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 8: butterfly<T, 8>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 4: butterfly<T, 4>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 2: butterfly<T, 2>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 3: butterfly<T, 3>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 5: butterfly<T, 5>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 7: butterfly<T, 7>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv); break;
        case 11:
          if constexpr (BIG) butterfly<T, 11>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv);
          break;
        case 13:
          if constexpr (BIG) butterfly<T, 13>(a, b, S, j, m, kk, jq, per_line, Ns, tstride, tw, inv);
          break;
        default: break;  // rejected on the host when the plan is built
      }
      j += dj;
      m += dm;
      if (m >= M) {
        m -= M;
        ++j;
      }
    }
    lds_barrier();
    cplx<T>* t = a;
    a = b;
    b = t;
  }
  return a;
}

// In-place form of block_fft for one buffer: every pass first loads the
// inputs of all of a thread's butterflies (with their twiddles) into
// registers, then a barrier, then computes them and writes the outputs to
// their Stockham positions in the same buffer, then a barrier.  It needs half
// the LDS of the ping-pong form, so a workgroup can hold twice as many lines.
// K bounds the complex values a thread holds: a radix-R pass handles up to K/R
// butterflies per thread (butterfly b = j * M + m of thread t is t + k * BT,
// as in block_fft), which inplace_fits() checks on the host.  Radices 2, 3, 4,
// 5, 7, 8 and 16 only (no BIG radices, no generic pass).
template <typename T, int BT, int K, int R>
__device__ __forceinline__ void inplace_pass(cplx<T>* a, int M, int S, int per_line, int Ns,
                                             unsigned long long mg, int tstride,
                                             const cplx<T>* tw, bool inv) {
  constexpr int NB = K / R;
  const int total = M * per_line;
  cplx<T> v[NB][R];
  int base[NB], line[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int b = int(threadIdx.x) + k * BT;
    base[k] = -1;
    if (b < total) {
      const int j = b / M, m = b - j * M;
      const int jq = fast_div(j, mg);
      const int kk = j - jq * Ns;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        cplx<T> x = a[(j + t * per_line) * S + m];
        if (t > 0 && kk > 0) {
          cplx<T> w = tw[t * kk * tstride];
          if (inv) w.y = -w.y;
          x = cmul(x, w);
        }
        v[k][t] = x;
      }
      base[k] = jq * Ns * R + kk;
      line[k] = m;
    }
  }
  lds_barrier();
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (base[k] >= 0) {
      Dft<T, R>::run(v[k], inv, tw, per_line);
#pragma unroll
      for (int u = 0; u < R; ++u) a[(base[k] + u * Ns) * S + line[k]] = v[k][u];
    }
  }
  lds_barrier();
}

template <typename T, int BT, int K>
__device__ __forceinline__ void block_fft_inplace(cplx<T>* a, int M, int S,
                                                  const AxisPlan* __restrict__ pl,
                                                  const cplx<T>* tw, bool inv) {
  const int npass = pl->npass;
  const int L = pl->L;
  for (int p = 0; p < npass; ++p) {
    const int R = pl->radix[p];
    const int Ns = pl->ns[p];
    const unsigned long long mg = pl->ns_magic[p];
    const int per_line = L / R;
    const int tstride = per_line / Ns;
    switch (R) {
      case 16: inplace_pass<T, BT, K, 16>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 8: inplace_pass<T, BT, K, 8>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 4: inplace_pass<T, BT, K, 4>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 2: inplace_pass<T, BT, K, 2>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 3: inplace_pass<T, BT, K, 3>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 5: inplace_pass<T, BT, K, 5>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      case 7: inplace_pass<T, BT, K, 7>(a, M, S, per_line, Ns, mg, tstride, tw, inv); break;
      default: break;  // rejected on the host (inplace_fits)
    }
  }
}

// Host: whether block_fft_inplace<T, BT, K> can run M lines of plan p.
inline bool inplace_fits(const AxisPlan& p, int M, int BT, int K) {
  for (int i = 0; i < p.npass; ++i) {
    const int R = p.radix[i];
    if (R != 2 && R != 3 && R != 4 && R != 5 && R != 7 && R != 8 && R != 16) return false;
    if (K / R < 1) return false;
    const long long butterflies = (long long)M * (p.L / R);
    if (butterflies > (long long)(K / R) * BT) return false;
  }
  return true;
}

// Host: factor L into radices 16,8,4,2,3,5,7,11,13 (largest powers of two
// first), then into the remaining primes, which run the generic pass.
// uses_big_radix(): the plan needs the BIG block_fft (radix 11, 13 or generic).
inline AxisPlan make_axis_plan(int L) {
  AxisPlan p{};
  p.L = L;
  p.npass = 0;
  int rem = L, ns = 1;
  const int order[9] = {16, 8, 4, 2, 3, 5, 7, 11, 13};
  for (int r : order) {
    while (rem % r == 0 && rem > 1) {
      if (p.npass >= kMaxPasses) throw Error("FFT plan: too many passes");
      p.radix[p.npass] = r;
      p.ns[p.npass] = ns;
      p.ns_magic[p.npass] = div_magic(unsigned(ns));
      ns *= r;
      rem /= r;
      ++p.npass;
    }
  }
  for (int r = 17; rem > 1; r += 2) {
    if (r > rem / r) r = rem;  // no factor up to sqrt(rem): rem is prime
    while (rem % r == 0) {
      if (p.npass >= kMaxPasses) throw Error("FFT plan: too many passes");
      p.radix[p.npass] = r;
      p.ns[p.npass] = ns;
      p.ns_magic[p.npass] = div_magic(unsigned(ns));
      ns *= r;
      rem /= r;
      ++p.npass;
    }
  }
  if (L >= (1 << 16)) throw Error("grid dimension " + std::to_string(L) + " too large");
  return p;
}

inline bool uses_big_radix(const AxisPlan& p) {
  for (int i = 0; i < p.npass; ++i)
    if (p.radix[i] > 8 && p.radix[i] != 16) return true;
  return false;
}

}  // namespace mtl
