// Device MINRES for the Galerkin system  Ghat * (D : x) = b.
#pragma once

#include "spectral.hpp"

namespace mtl {

// S: storage of the Lanczos vectors (float, double or f48); the w vectors, the
// iterate x and the tangent are work_t<S>.  A: arithmetic of the matvec FFT
// passes and of the vector kernels.
template <typename S, typename A = S>
class Minres {
 public:
  using W = work_t<S>;
  explicit Minres(Spectral<S, A>& sp);
  ~Minres();
  Minres(const Minres&) = delete;
  Minres& operator=(const Minres&) = delete;

  // Solve with x0 = 0; b is a double SoA [6][n] field, tangent W [n/64][21][64].
  // Returns the iteration count (SciPy's callback count); istop as SciPy.
  int solve(const double* b, const W* tangent, const Vec6& mask, double rtol, long long maxiter,
            hipStream_t s, int* istop_out);

  const W* x() const { return x_; }
  // Host copy of the last state (valid after solve()).
  const MinresState& state() const { return *host_state_; }

 private:
  Spectral<S, A>& sp_;
  long long n6_;
  void* lanczos_mem_[2] = {};
  RVec<S> lanczos_[2];  // the two Lanczos vectors (roles swap)
  W* w_[3] = {};        // three w vectors (roles rotate)
  W* x_ = nullptr;
  double* partials_ = nullptr;
  int partial_cap_ = 0;
  MinresState* state_ = nullptr;
  MinresState* host_state_ = nullptr;  // pinned
};

}  // namespace mtl
