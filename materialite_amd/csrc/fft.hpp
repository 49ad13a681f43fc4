// Workgroup-level mixed-radix Stockham FFT over lines held in LDS.
//
// A workgroup of BT threads transforms `nlines` lines of length L stored
// contiguously in LDS (line i at buf[i*L .. i*L+L)).  Each radix-R pass reads
// the R inputs of its butterflies, applies twiddles w^(t*k) (k = j mod Ns),
// performs an R-point DFT and writes the outputs to their autosorted Stockham
// positions (j/Ns)*Ns*R + k + u*Ns.  Outputs are staged in registers across a
// barrier, so the pass runs in place: every thread holds at most PTS complex
// values, and the caller guarantees nlines * L <= PTS * BT.
//
// Twiddles come from a per-axis table tw[m] = exp(-2 pi i m / L) copied into
// LDS; the inverse transform uses their conjugates (unnormalised).
#pragma once

#include "common.hpp"

namespace mtl {

constexpr int kMaxPasses = 24;

struct AxisPlan {
  int L;
  int npass;
  int radix[kMaxPasses];
  int ns[kMaxPasses];
};

// --------------------------------------------------------------- small DFTs
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool) {
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
};

template <typename T>
struct Dft<T, 8> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv) {
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

template <typename T>
struct Dft<T, 3> {
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv) {
    const T s = T(0.86602540378443864676372317075293618347);
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
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv) {
    const T c1 = T(0.30901699437494742410229341718281905886);
    const T c2 = T(-0.80901699437494742410229341718281905886);
    const T s1 = T(0.95105651629515357211643933337938214340);
    const T s2 = T(0.58778525229247312916870595463907276860);
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
  __device__ __forceinline__ static void run(cplx<T>* v, bool inv) {
    // direct 7-point DFT with exact-to-T constant roots
    const T cs[7] = {T(1), T(0.62348980185873353053), T(-0.22252093395631440429),
                     T(-0.90096886790241912624), T(-0.90096886790241912624),
                     T(-0.22252093395631440429), T(0.62348980185873353053)};
    const T sn[7] = {T(0), T(0.78183148246802980871), T(0.97492791218182360702),
                     T(0.43388373911755812048), T(-0.43388373911755812048),
                     T(-0.97492791218182360702), T(-0.78183148246802980871)};
    cplx<T> in[7];
#pragma unroll
    for (int t = 0; t < 7; ++t) in[t] = v[t];
#pragma unroll
    for (int u = 0; u < 7; ++u) {
      cplx<T> acc = {T(0), T(0)};
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const int m = (t * u) % 7;
        const T c = cs[m], s = inv ? sn[m] : -sn[m];
        acc.x += in[t].x * c - in[t].y * s;
        acc.y += in[t].x * s + in[t].y * c;
      }
      v[u] = acc;
    }
  }
};

// --------------------------------------------------------------- one pass
template <typename T, int BT, int PTS, int R>
__device__ __forceinline__ void fft_pass(cplx<T>* buf, int nlines, int L, int Ns,
                                         const cplx<T>* tw, bool inv) {
  constexpr int KMAX = (PTS + R - 1) / R;
  const int per_line = L / R;
  const int total = nlines * per_line;
  const int tstride = L / (Ns * R);
  cplx<T> st[KMAX * R];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int jj = threadIdx.x + k * BT;
    if (jj < total) {
      const int line = jj / per_line;
      const int j = jj - line * per_line;
      const int kk = j % Ns;
      const cplx<T>* a = buf + line * L;
      cplx<T> v[R];
#pragma unroll
      for (int t = 0; t < R; ++t) {
        cplx<T> x = a[j + t * per_line];
        if (t > 0 && kk > 0) {
          cplx<T> w = tw[t * kk * tstride];
          if (inv) w.y = -w.y;
          x = cmul(x, w);
        }
        v[t] = x;
      }
      Dft<T, R>::run(v, inv);
#pragma unroll
      for (int u = 0; u < R; ++u) st[k * R + u] = v[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int jj = threadIdx.x + k * BT;
    if (jj < total) {
      const int line = jj / per_line;
      const int j = jj - line * per_line;
      const int kk = j % Ns;
      cplx<T>* a = buf + line * L;
      const int idx = (j / Ns) * Ns * R + kk;
#pragma unroll
      for (int u = 0; u < R; ++u) a[idx + u * Ns] = st[k * R + u];
    }
  }
  __syncthreads();
}

// Transform all lines in LDS.  Ends with a barrier.
template <typename T, int BT, int PTS>
__device__ void block_fft(cplx<T>* buf, int nlines, const AxisPlan& pl, const cplx<T>* tw,
                          bool inv) {
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p], Ns = pl.ns[p];
    switch (R) {
      case 8: fft_pass<T, BT, PTS, 8>(buf, nlines, pl.L, Ns, tw, inv); break;
      case 4: fft_pass<T, BT, PTS, 4>(buf, nlines, pl.L, Ns, tw, inv); break;
      case 2: fft_pass<T, BT, PTS, 2>(buf, nlines, pl.L, Ns, tw, inv); break;
      case 3: fft_pass<T, BT, PTS, 3>(buf, nlines, pl.L, Ns, tw, inv); break;
      case 5: fft_pass<T, BT, PTS, 5>(buf, nlines, pl.L, Ns, tw, inv); break;
      case 7: fft_pass<T, BT, PTS, 7>(buf, nlines, pl.L, Ns, tw, inv); break;
      default: break;  // rejected on the host when the plan is built
    }
  }
}

// Host: factor L into radices 8,4,2,3,5,7 (largest first). Throws if impossible.
inline AxisPlan make_axis_plan(int L) {
  AxisPlan p{};
  p.L = L;
  p.npass = 0;
  int rem = L, ns = 1;
  const int order[6] = {8, 4, 2, 3, 5, 7};
  for (int r : order) {
    while (rem % r == 0 && rem > 1) {
      if (p.npass >= kMaxPasses) throw Error("FFT plan: too many passes");
      p.radix[p.npass] = r;
      p.ns[p.npass] = ns;
      ns *= r;
      rem /= r;
      ++p.npass;
    }
  }
  if (rem != 1)
    throw Error("grid dimension " + std::to_string(L) +
                " has a prime factor above 7; the FFT supports radices 2, 3, 4, 5, 7, 8");
  return p;
}

}  // namespace mtl
