// Optional per-kernel timing.  Kernels launched through Profiler::launch carry
// a start and a stop event on their own dispatch (hipExtLaunchKernelGGL), so a
// figure is the kernel's execution time, with no marker packets between
// kernels.
//
// After a host synchronisation the device has idled while the host ran, and
// the start event of the next dispatch would also cover its wake-up (tens of
// microseconds, which a kernel trace does not count in the dispatch); with
// timing enabled an empty kernel is launched first to take that delay.
#pragma once

#include <hip/hip_ext.h>

#include <vector>

#include "common.hpp"

namespace mtl {

// empty dispatch issued ahead of the first timed kernel after the device idled
static __global__ void profiler_wake_kernel() {}

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
  std::vector<float> durations[kNumKernelIds];  // ms of each timed launch

  template <typename F, typename... Args>
  void launch(int id, F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
    if (!enabled) {
      hipExtLaunchKernelGGL(kernel, grid, block, std::uint32_t(lds), s, nullptr, nullptr, 0u,
                            args...);
      return;
    }
    if (wake_next_) {
      wake_next_ = false;
      hipLaunchKernelGGL(profiler_wake_kernel, dim3(1), dim3(64), 0, s);
    }
    hipEvent_t a = take(), b = take();
    hipExtLaunchKernelGGL(kernel, grid, block, std::uint32_t(lds), s, a, b, 0u, args...);
    pending_.push_back({id, a, b});
    if (pending_.size() > 4096) flush();
  }

  // the next launch follows a host synchronisation
  void after_sync() { wake_next_ = enabled; }

  void flush() {
    for (auto& r : pending_) {
      MTL_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      MTL_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms[r.id] += ms;
      launches[r.id] += 1;
      durations[r.id].push_back(ms);
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
      durations[i].clear();
    }
    wake_next_ = false;
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
  bool wake_next_ = false;
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
};

Profiler& profiler();

// Host synchronisation with `s`; tells the profiler that the device idles
inline void stream_sync(hipStream_t s) {
  MTL_HIP(hipStreamSynchronize(s));
  profiler().after_sync();
}

}  // namespace mtl
