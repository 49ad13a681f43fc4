// Host-side driver of the FFT passes that apply the rotated-scheme projection
// Ghat (small_strain_fft.py:301-382 of the reference) to 6-component fields.
#pragma once

#include "fft.hpp"
#include "krylov.hpp"

namespace mtl {

struct Vec6 {
  double v[6];
};

// packed upper-triangular index of a symmetric 6x6 (row-major, i <= j)
__host__ __device__ constexpr int sym6(int i, int j) {
  return i <= j ? i * 6 - (i * (i - 1)) / 2 + (j - i) : j * 6 - (j * (j - 1)) / 2 + (i - j);
}

// T: type of the spectrum and of the Krylov vectors and tangent in HBM;
// A: arithmetic type of the FFT passes (LDS tiles, twiddles, Ghat).  A = double
// with T = float transforms FP32 data with FP64 butterflies.
template <typename T, typename A = T>
class Spectral {
 public:
  Spectral(int nx, int ny, int nz);
  ~Spectral();
  Spectral(const Spectral&) = delete;
  Spectral& operator=(const Spectral&) = delete;

  const Grid& grid() const { return g_; }

  // out = F^-1[ Ghat . F(a + sgn*in) ] (real part), fields SoA [6][n].
  template <typename TIn, typename TOut>
  void project(const TIn* in, TOut* out, const Vec6& a, double sgn, const Vec6& mask,
               hipStream_t s);

  // One MINRES matvec fused with its vector work:
  //   tau = D : (s * r2);  Av = F^-1 Ghat F tau;  u = Av - c_r1 * r1 -> r1;
  //   partials[b] = sum over the block of (s * r2) . u.
  // Returns the number of partial sums written.
  int matvec_minres(const T* r2, const T* tangent, T* r1, const Vec6& mask,
                    const MinresState* st, double* partials, hipStream_t s);

  int zinv_blocks() const;

 private:
  template <int MODE, typename TIn, typename TOut>
  void run(const TIn* in, const T* tangent, TOut* out, T* r1, const Vec6& a, double sgn,
           const Vec6& mask, const MinresState* st, double* partials, hipStream_t s);

  Grid g_;
  AxisPlan px_, py_, pz_;
  AxisPlan *pxd_ = nullptr, *pyd_ = nullptr, *pzd_ = nullptr;  // device copies
  int pz_lines_;  // z-lines per workgroup in the z passes
  int ty_;        // kz columns per workgroup in the y pass
  int tx_;        // kz columns per workgroup in the x pass
  bool x_resident_;  // x pass keeps all six components of its tile in LDS
  bool x_pairs_ = false;  // streamed x pass moves pairs of complex values
  bool big_x_ = false, big_y_ = false, big_z_ = false;  // axis plans use radix 11 or 13
  cplx<T>* spec_ = nullptr;
  cplx<A>* twx_ = nullptr;
  cplx<A>* twy_ = nullptr;
  cplx<A>* twz_ = nullptr;
  A* tanx_ = nullptr;
  A* tany_ = nullptr;
  A* tanz_ = nullptr;
};

}  // namespace mtl
