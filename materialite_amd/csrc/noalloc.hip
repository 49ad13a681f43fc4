// Projection apply that allocates no device memory: the spectrum and the FFT
// tables live in one file-scope device array, `scratch`, of a fixed size.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).  Each GPU holds its own
// copy of the array for the life of the process; every call on a GPU shares
// it, so calls that overlap on one GPU race (they run on the null stream and
// return after it has drained).  The same kernels as mtl_solver_project run on
// the same tables, so the results are the handle-based entry point's.
#include <mutex>

#include "../../include/materialite_amd.h"
#include "profiler.hpp"
#include "spectral.hpp"

namespace mtl {

namespace {

constexpr size_t kScratchBytes = MTL_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// device address of each GPU's `scratch` (the host-side symbol is only a
// stand-in and must never reach a kernel or a copy)
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

void* scratch_on_current_device() {
  int dev = 0;
  MTL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw Error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

template <typename T, typename A>
void project_in_scratch(int nx, int ny, int nz, const double* in, double* out,
                        const double m[6]) {
  const size_t need = Spectral<T, A>::workspace_bytes(nx, ny, nz);
  if (need > kScratchBytes)
    throw Error("grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
                " needs " + std::to_string(need) + " bytes of scratch; the no-allocation projection "
                "holds at most " + std::to_string(kScratchBytes));
  void* ws = scratch_on_current_device();
  hipStream_t s = nullptr;  // null stream: no stream object is created
  Spectral<T, A> sp(nx, ny, nz, ws, kScratchBytes, s);
  Vec6 mask, zero{};
  for (int c = 0; c < 6; ++c) mask.v[c] = m[c];
  sp.project(in, out, zero, 1.0, mask, s);
  stream_sync(s);
}

}  // namespace

}  // namespace mtl

extern "C" int mtl_project_noalloc(int nx, int ny, int nz, int precision, const double* in_dev,
                                   double* out_dev, const double stress_mask[6]) {
  if (!in_dev || !out_dev || !stress_mask) {
    mtl::set_last_error("null argument");
    return -2;
  }
  try {
    if (precision == MTL_FP64)
      mtl::project_in_scratch<double, double>(nx, ny, nz, in_dev, out_dev, stress_mask);
    else if (precision == MTL_MIXED)
      mtl::project_in_scratch<float, float>(nx, ny, nz, in_dev, out_dev, stress_mask);
    else if (precision == MTL_REFINED)
      mtl::project_in_scratch<mtl::f48, double>(nx, ny, nz, in_dev, out_dev, stress_mask);
    else {
      mtl::set_last_error("unknown precision");
      return -2;
    }
    return 0;
  } catch (const std::exception& e) {
    mtl::set_last_error(e.what());
    return -1;
  }
}
