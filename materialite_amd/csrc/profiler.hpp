// Optional per-kernel timing with HIP events recorded on the launching stream.
#pragma once

#include <vector>

#include "common.hpp"

namespace mtl {

enum KernelId {
  kKZForward = 0,
  kKYPass = 1,
  kKXGhat = 2,
  kKZInverse = 3,
  kKMinresVector = 4,
  kKMinresScalar = 5,
  kKEvpIter = 6,
  kKEvpFinish = 7,
  kKElastic = 8,
  kKNewton = 9,
  // the same FFT passes when they project a residual (FP64) instead of
  // applying the Krylov matvec
  kKResidualZForward = 10,
  kKResidualYPass = 11,
  kKResidualXGhat = 12,
  kKResidualZInverse = 13,
  kNumKernelIds = 14,
};

// matvec pass id -> the residual-projection id of the same pass
constexpr int residual_id(int matvec_id) { return matvec_id + kKResidualZForward; }

class Profiler {
 public:
  bool enabled = false;
  double total_ms[kNumKernelIds] = {};
  long long launches[kNumKernelIds] = {};

  void begin(hipStream_t s) {
    if (!enabled) return;
    cur_ = take();
    MTL_HIP(hipEventRecord(cur_, s));
  }
  void end(int id, hipStream_t s) {
    if (!enabled) return;
    hipEvent_t b = take();
    MTL_HIP(hipEventRecord(b, s));
    pending_.push_back({id, cur_, b});
    if (pending_.size() > 4096) flush();
  }
  void flush() {
    for (auto& r : pending_) {
      MTL_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      MTL_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms[r.id] += ms;
      launches[r.id] += 1;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    pending_.clear();
  }
  void reset() {
    flush();
    for (int i = 0; i < kNumKernelIds; ++i) {
      total_ms[i] = 0.0;
      launches[i] = 0;
    }
  }

 private:
  struct Rec {
    int id;
    hipEvent_t a, b;
  };
  hipEvent_t take() {
    if (pool_.empty()) {
      hipEvent_t e;
      MTL_HIP(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t cur_ = nullptr;
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
};

Profiler& profiler();

}  // namespace mtl
