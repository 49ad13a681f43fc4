// Device MINRES for the Galerkin system  Ghat * (D : x) = b.
#pragma once

#include "spectral.hpp"

namespace mtl {

// T: Krylov vectors and tangent; A: arithmetic type of the matvec FFT passes.
template <typename T, typename A = T>
class Minres {
 public:
  explicit Minres(Spectral<T, A>& sp);
  ~Minres();
  Minres(const Minres&) = delete;
  Minres& operator=(const Minres&) = delete;

  // Solve with x0 = 0; b is a double SoA [6][n] field, tangent T [n/64][21][64].
  // Returns the iteration count (SciPy's callback count); istop as SciPy.
  int solve(const double* b, const T* tangent, const Vec6& mask, double rtol, long long maxiter,
            hipStream_t s, int* istop_out);

  const T* x() const { return x_; }
  // Host copy of the last state (valid after solve()).
  const MinresState& state() const { return *host_state_; }

 private:
  Spectral<T, A>& sp_;
  long long n6_;
  T* vec_[5] = {};  // two Lanczos vectors and three w vectors (roles rotate)
  T* x_ = nullptr;
  double* partials_ = nullptr;
  int partial_cap_ = 0;
  MinresState* state_ = nullptr;
  MinresState* host_state_ = nullptr;  // pinned
};

}  // namespace mtl
