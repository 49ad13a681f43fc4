// Host-side driver of the FFT passes that apply the rotated-scheme projection
// Ghat (small_strain_fft.py:301-382 of the reference) to 6-component fields.
#pragma once

#include "fft.hpp"
#include "krylov.hpp"
#include "storage.hpp"

namespace mtl {

struct Vec6 {
  double v[6];
};

// packed upper-triangular index of a symmetric 6x6 (row-major, i <= j)
__host__ __device__ constexpr int sym6(int i, int j) {
  return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j);
}

// Per-solve pointers of one MINRES solve in a batch of same-size solves (one
// launch of each kernel serves every solve of the batch: blockIdx.y, or
// blockIdx.x in the one-workgroup scalar kernels, selects the slot).  Roles are
// those of one iteration: r1 / r2 the Lanczos vectors before the swap that
// follows k_lanczos (the matvec reads r2 and updates r1; k_update reads r2 as
// v), w1 / w2 the previous w vectors and wf the one written.
template <typename S>
struct BatchSlot {
  CVec<S> spec;
  const work_t<S>* tangent;
  RVec<S> r1, r2;
  int maskbits;
  MinresState* st;
  double* partials;
  work_t<S>* w1;
  work_t<S>* w2;
  work_t<S>* wf;
  work_t<S>* x;
  const double* b;  // right-hand side ([6][n] FP64), read at the start
  double rtol;
};

// S: storage of the spectrum and of the Lanczos vectors in HBM (float,
// double or f48); the tangent and the MINRES w / x vectors are work_t<S>.
// A: arithmetic type of the FFT passes (LDS tiles, twiddles, Ghat).
template <typename S, typename A = S>
class Spectral {
 public:
  using W = work_t<S>;
  Spectral(int nx, int ny, int nz);
  // The same, with the spectrum, twiddle, tan and plan tables carved out of a
  // caller-provided device workspace of `bytes` bytes: nothing is allocated on
  // the device; the tables are copied to it on stream s before returning.
  Spectral(int nx, int ny, int nz, void* workspace, size_t bytes, hipStream_t s);
  // device bytes the workspace constructor needs for this grid
  static size_t workspace_bytes(int nx, int ny, int nz);
  ~Spectral();
  Spectral(const Spectral&) = delete;
  Spectral& operator=(const Spectral&) = delete;

  const Grid& grid() const { return g_; }

  // out = F^-1[ Ghat . F(a + sgn*in) ] (real part), fields SoA [6][n] FP64.
  void project(const double* in, double* out, const Vec6& a, double sgn, const Vec6& mask,
               hipStream_t s);

  // One MINRES matvec fused with its vector work:
  //   tau = D : (s * r2);  Av = F^-1 Ghat F tau;  u = Av - c_r1 * r1 -> r1;
  //   partials[b] = sum over the block of (s * r2) . u.
  // Returns the number of partial sums written.
  int matvec_minres(RVec<S> r2, const W* tangent, RVec<S> r1, const Vec6& mask,
                    const MinresState* st, double* partials, hipStream_t s);
  // The same matvec for nb solves of this grid at once, each from its slot
  // (slots: device array of nb entries).  Returns the partial sums per solve.
  int matvec_batch(const BatchSlot<S>* slots, int nb, hipStream_t s);
  CVec<S> spectrum() const { return spec_; }
  static int mask_bits(const Vec6& mask);

  int zinv_blocks() const;

 private:
  void plan(int nx, int ny, int nz);
  template <int MODE, typename In>
  void run(In in, const W* tangent, double* out, RVec<S> r1, const Vec6& a, double sgn,
           const Vec6& mask, const MinresState* st, double* partials, hipStream_t s,
           const BatchSlot<S>* slots = nullptr, int nb = 1);

  Grid g_;
  AxisPlan px_, py_, pz_;
  AxisPlan *pxd_ = nullptr, *pyd_ = nullptr, *pzd_ = nullptr;  // device copies
  int pz_lines_;  // z-lines per workgroup in the z passes
  int ty_;        // kz columns per workgroup in the y pass
  int tx_;        // kz columns per workgroup in the x pass
  bool x_resident_;  // x pass keeps all six components of its tile in LDS
  bool x_pairs_ = false;  // streamed x pass moves pairs of complex values
  bool x_inplace_ = false;  // resident x pass with in-place (one-buffer) FFT passes
  bool y_inplace_ = false;  // y pass with in-place FFT passes
  bool big_x_ = false, big_y_ = false, big_z_ = false;  // axis plans use radix 11 or 13
  void* spec_mem_ = nullptr;
  CVec<S> spec_;
  cplx<A>* twx_ = nullptr;
  cplx<A>* twy_ = nullptr;
  cplx<A>* twz_ = nullptr;
  A* tanx_ = nullptr;
  A* tany_ = nullptr;
  A* tanz_ = nullptr;
  bool external_ = false;  // buffers live in a caller's workspace
};

}  // namespace mtl
