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

#include <atomic>
#include <mutex>
#include <tuple>
#include <utility>
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

// Records kernel launches as the nodes of a graph instead of launching them:
// while a recorder is active on a thread, every Profiler::launch made on that
// thread appends a kernel node that depends on the previous one, so the graph
// replays the launches in order on the stream it is launched on.  Nothing is
// captured from a stream, so other threads' stream work is unaffected.
class GraphRecorder {
 public:
  GraphRecorder() {
    MTL_HIP(hipGraphCreate(&graph_, 0));
    prev_ = active();
    active() = this;
  }
  ~GraphRecorder() {
    active() = prev_;
    if (graph_) (void)hipGraphDestroy(graph_);
  }
  GraphRecorder(const GraphRecorder&) = delete;
  GraphRecorder& operator=(const GraphRecorder&) = delete;

  static GraphRecorder*& active() {
    static thread_local GraphRecorder* r = nullptr;
    return r;
  }

  template <typename... KA, typename... Args>
  void add(void (*kernel)(KA...), dim3 grid, dim3 block, size_t lds, Args... args) {
    static_assert(sizeof...(KA) == sizeof...(Args), "kernel argument count");
    std::tuple<KA...> vals(static_cast<KA>(args)...);
    add_node(reinterpret_cast<const void*>(kernel), grid, block, lds, vals,
             std::index_sequence_for<KA...>{});
  }

  // the recorded launches as an executable graph (ends the recording)
  hipGraphExec_t instantiate() {
    active() = prev_;
    hipGraphExec_t exec = nullptr;
    MTL_HIP(hipGraphInstantiate(&exec, graph_, nullptr, nullptr, 0));
    return exec;
  }

 private:
  template <typename Tup, size_t... I>
  void add_node(const void* fn, dim3 grid, dim3 block, size_t lds, Tup& vals,
                std::index_sequence<I...>) {
    void* params[sizeof...(I) + 1] = {static_cast<void*>(&std::get<I>(vals))...};
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(fn);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = unsigned(lds);
    p.kernelParams = params;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    MTL_HIP(hipGraphAddKernelNode(&node, graph_, last_ ? &last_ : nullptr, last_ ? 1 : 0, &p));
    last_ = node;
  }
  hipGraph_t graph_ = nullptr;
  hipGraphNode_t last_ = nullptr;
  GraphRecorder* prev_ = nullptr;
};

// Handles used from several threads at once (the serving layer) share the
// profiler: with timing off a launch only reads an atomic flag; with timing on the
// records are kept under a mutex.
class Profiler {
 public:
  template <typename F, typename... Args>
  void launch(int id, F kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args) {
    if (GraphRecorder* r = GraphRecorder::active()) {
      r->add(kernel, grid, block, lds, args...);
      return;
    }
    if (!enabled_.load(std::memory_order_relaxed)) {
      hipExtLaunchKernelGGL(kernel, grid, block, std::uint32_t(lds), s, nullptr, nullptr, 0u,
                            args...);
      return;
    }
    std::lock_guard<std::mutex> lock(mu_);
    if (wake_next_) {
      wake_next_ = false;
      hipLaunchKernelGGL(profiler_wake_kernel, dim3(1), dim3(64), 0, s);
    }
    hipEvent_t a = take(), b = take();
    hipExtLaunchKernelGGL(kernel, grid, block, std::uint32_t(lds), s, a, b, 0u, args...);
    pending_.push_back({id, a, b});
    if (pending_.size() > 4096) flush_locked();
  }

  // the next launch follows a host synchronisation
  void after_sync() {
    if (!enabled_.load(std::memory_order_relaxed)) return;
    std::lock_guard<std::mutex> lock(mu_);
    wake_next_ = true;
  }

  bool enabled() const { return enabled_.load(std::memory_order_relaxed); }

  void set_enabled(bool on) {
    std::lock_guard<std::mutex> lock(mu_);
    flush_locked();
    enabled_.store(on, std::memory_order_relaxed);
    wake_next_ = false;
  }

  void flush() {
    std::lock_guard<std::mutex> lock(mu_);
    flush_locked();
  }

  void reset() {
    std::lock_guard<std::mutex> lock(mu_);
    flush_locked();
    for (int i = 0; i < kNumKernelIds; ++i) {
      total_ms_[i] = 0.0;
      launches_[i] = 0;
      durations_[i].clear();
    }
    wake_next_ = false;
  }

  // totals of kernel id `id` (after waiting for its pending records)
  void read(int id, double* total_ms, long long* launches) {
    std::lock_guard<std::mutex> lock(mu_);
    flush_locked();
    *total_ms = total_ms_[id];
    *launches = launches_[id];
  }

  // durations (ms) of the timed launches of `id`: copies up to `capacity`,
  // returns how many there are
  long long read_durations(int id, float* out, long long capacity) {
    std::lock_guard<std::mutex> lock(mu_);
    flush_locked();
    const auto& d = durations_[id];
    const long long m = capacity < (long long)d.size() ? capacity : (long long)d.size();
    for (long long i = 0; i < m; ++i) out[i] = d[size_t(i)];
    return (long long)d.size();
  }

 private:
  struct Rec {
    int id;
    hipEvent_t a, b;
  };
  void flush_locked() {
    for (auto& r : pending_) {
      MTL_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      MTL_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms_[r.id] += ms;
      launches_[r.id] += 1;
      durations_[r.id].push_back(ms);
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    pending_.clear();
  }
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
  std::atomic<bool> enabled_{false};
  std::mutex mu_;
  bool wake_next_ = false;
  double total_ms_[kNumKernelIds] = {};
  long long launches_[kNumKernelIds] = {};
  std::vector<float> durations_[kNumKernelIds];  // ms of each timed launch
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
