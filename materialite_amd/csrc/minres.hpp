// Device MINRES for the Galerkin system  Ghat * (D : x) = b.
#pragma once

#include <array>
#include <map>
#include <vector>

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

  // One system of a batch: this solver, its right-hand side and tangent.
  struct BatchItem {
    Minres* m;
    const double* b;
    const W* tangent;
    Vec6 mask;
    double rtol;
  };
  // Solves the systems of several solvers of the same grid and form at once
  // (x0 = 0 each): every iteration is one launch of each kernel for all of
  // them, with per-system scalars, stopping tests and iteration counts
  // (itns[i]); a system that has stopped costs its workgroups one flag read.
  // Each result is the one solve() gives for the same system.  With
  // MTL_BATCH_GRAPHS=1 (and kernel timing off) each iteration is the replay
  // of a recorded graph of its ten launches, kept by the first solver of the
  // batch per batch size.
  static void solve_batch(const std::vector<BatchItem>& items, long long maxiter, hipStream_t s,
                          int* itns);
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
  // batch slot tables (six iteration phases x capacity), device and pinned host
  BatchSlot<S>* slots_dev_ = nullptr;
  BatchSlot<S>* slots_host_ = nullptr;
  int slot_cap_ = 0;
  // one-iteration graphs of solve_batch by batch size, one per role phase;
  // they hold the slot table's address
  std::map<int, std::array<hipGraphExec_t, 6>> graphs_;
  void drop_graphs();
};

}  // namespace mtl
