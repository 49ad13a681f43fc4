// Newton-Krylov load increment of SmallStrainFFT on the device, and the C ABI.
//
// Restates small_strain_fft.py:106-246 of the reference: predictor, Newton loop
// (constitutive update, residual b = Ghat*(Sigma - sigma), MINRES correction,
// Abaqus-style equilibrium / strain-correction tests), commit and cutback.
// Host code keeps only scalars; every field lives in HBM as SoA [6][N] FP64
// (strain, stress, residual) or [N/64][21][64] K (tangent, K = float or double).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/materialite_amd.h"
#include "constitutive.hpp"
#include "minres.hpp"
#include "profiler.hpp"

namespace mtl {

bool inverse6_host(const double* A, double* out);

namespace {

constexpr int kNBT = 256;
constexpr int kNormBlocks = 4096;
constexpr int kNormWidth = 9;  // max|b|, sum sigma (6), max|dev dsigma|^2, max|D:dx|^2

__global__ void __launch_bounds__(kNBT) k_predictor(long long n, const double* eps_old,
                                                     double* fluct, double* eps, Vec6 de, Vec6 pred,
                                                     int reset) {
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double f = fluct[c * n + p];
      if (reset) {
        f = 0.0;
        fluct[c * n + p] = 0.0;
      }
      eps[c * n + p] = eps_old[c * n + p] + de.v[c] + f + pred.v[c];
    }
  }
}

template <typename K>
__global__ void __launch_bounds__(kNBT) k_guess(long long n, int step0, const double* sig_old,
                                                 const K* D, Vec6 ds, Vec6 de, double* guess) {
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      double g;
      if (step0) {
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) acc += double(D[tan_idx(sym6(a, b), p)]) * de.v[b];
        g = ds.v[a] + acc;
      } else {
        g = sig_old[a * n + p] + ds.v[a];
      }
      guess[a * n + p] = g;
    }
  }
}

// eps += x + shift_c, fluct += x + shift_c (shift is per component)
template <typename K>
__global__ void __launch_bounds__(kNBT) k_add_correction(long long n, const K* x, double* eps,
                                                          double* fluct, const double* shift_dev) {
  double shift[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) shift[c] = shift_dev != nullptr ? shift_dev[c] : 0.0;
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double d = double(x[blk_idx(c, p, 6)]) + shift[c];  // x blocked by 64 points
      eps[c * n + p] += d;
      fluct[c * n + p] += d;
    }
  }
}

// per-block component sums of a 6-component field blocked by 64 points
template <typename K>
__global__ void __launch_bounds__(kNBT) k_sum6(long long n, const K* x, double* partials) {
  __shared__ double red[kNBT / 64];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT)
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] += double(x[blk_idx(c, p, 6)]);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double t = block_sum<kNBT>(acc[c], red);
    if (threadIdx.x == 0) partials[blockIdx.x * 6 + c] = t;
  }
}

// shift[c] = -(sum of component c) / n for the strain-controlled components
// (mask 0), 0 otherwise, from k_sum6's per-block partials (one workgroup)
__global__ void __launch_bounds__(kNBT) k_mean_shift(int nb, const double* partials, long long n,
                                                     Vec6 mask, double* shift) {
  __shared__ double red[kNBT / 64];
  for (int c = 0; c < 6; ++c) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += kNBT) acc += partials[i * 6 + c];
    const double sum = block_sum<kNBT>(acc, red);
    if (threadIdx.x == 0) shift[c] = mask.v[c] == 0.0 ? -sum / double(n) : 0.0;
  }
}

template <typename K>
__global__ void __launch_bounds__(kNBT) k_newton_norms(long long n, const double* b,
                                                        const double* sig, const double* sig_old,
                                                        const K* D, const K* x, double* partials) {
  __shared__ double red[kNBT / 64];
  double mb = 0.0, ms = 0.0, mdx = 0.0;
  double ss[6] = {0, 0, 0, 0, 0, 0};
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT) {
    double ds[6], xv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      mb = fmax(mb, fabs(b[c * n + p]));
      const double s = sig[c * n + p];
      ss[c] += s;
      ds[c] = s - sig_old[c * n + p];
      xv[c] = x ? double(x[blk_idx(c, p, 6)]) : 0.0;
    }
    const double tr = (ds[0] + ds[1] + ds[2]) / 3.0;
    double d2 = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double v = c < 3 ? ds[c] - tr : ds[c];
      d2 += v * v;
    }
    ms = fmax(ms, d2);
    if (x) {
      double q2 = 0.0;
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc += double(D[tan_idx(sym6(a, c), p)]) * xv[c];
        q2 += acc * acc;
      }
      mdx = fmax(mdx, q2);
    }
  }
  double out[kNormWidth];
  out[0] = block_max<kNBT>(mb, red);
#pragma unroll
  for (int c = 0; c < 6; ++c) out[1 + c] = block_sum<kNBT>(ss[c], red);
  out[7] = block_max<kNBT>(ms, red);
  out[8] = block_max<kNBT>(mdx, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < kNormWidth; ++k) partials[blockIdx.x * kNormWidth + k] = out[k];
}

template <typename K>
__global__ void __launch_bounds__(kNBT) k_tangent_sum(long long n, const K* D, double* partials) {
  __shared__ double red[kNBT / 64];
  double acc[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) acc[k] = 0.0;
  for (long long p = blockIdx.x * (long long)kNBT + threadIdx.x; p < n;
       p += (long long)gridDim.x * kNBT)
#pragma unroll
    for (int k = 0; k < 21; ++k) acc[k] += double(D[tan_idx(k, p)]);
#pragma unroll
  for (int k = 0; k < 21; ++k) {
    const double t = block_sum<kNBT>(acc[k], red);
    if (threadIdx.x == 0) partials[blockIdx.x * 21 + k] = t;
  }
}

__global__ void k_scale(long long count, double* a, double f) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    a[i] *= f;
}

// SoA FP64 copy [6][n] of a 6-component field blocked by 64 points
template <typename K>
__global__ void k_unblock6(long long n, const K* a, double* b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 6 * n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long c = i / n, p = i - c * n;
    b[i] = double(a[blk_idx(int(c), p, 6)]);
  }
}

// point-major copy of an SoA field: aos[p * nc + c] = soa[c * n + p]
__global__ void k_soa_to_aos(long long n, int nc, const double* soa, double* aos) {
  const long long count = n * nc;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / nc;
    const int c = int(i - p * nc);
    aos[i] = soa[c * n + p];
  }
}

// SoA copy of a point-major field: soa[c * n + p] = aos[p * nc + c]
__global__ void k_aos_to_soa(long long n, int nc, const double* aos, double* soa) {
  const long long count = n * nc;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / nc;
    const int c = int(i - p * nc);
    soa[c * n + p] = aos[i];
  }
}

// lowest relative MINRES tolerance of the FP32 form: its FP32 Krylov vectors
// and FFT arithmetic cannot honour tighter ones on the singular pore systems
// (the refined form, whose Lanczos vectors and spectrum keep 32 significant
// bits, has no floor)
constexpr double kMixedKrylovFloor = 1.0e-5;

bool debug_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("MTL_DEBUG");
    return e != nullptr && e[0] != '\0' && e[0] != '0';
  }();
  return on;
}

int grid_for(long long count) {
  return int(std::max<long long>(1, std::min<long long>((count + kNBT - 1) / kNBT, kNormBlocks)));
}

thread_local std::string g_last_error;

}  // namespace

// The constitutive models of a handle over its n points: per-point rotations
// (FP64 SoA [9][n]) and one Model per registered description, each over the
// points whose model id names it (Multiphase dispatch, multiphase.py:31-95).
template <typename K>
class ModelSet {
 public:
  explicit ModelSet(long long n) : n_(n) { MTL_HIP(hipMalloc(&rot_, 9 * n * sizeof(double))); }
  ~ModelSet() {
    models_.clear();
    if (rot_) (void)hipFree(rot_);
  }
  ModelSet(const ModelSet&) = delete;
  ModelSet& operator=(const ModelSet&) = delete;

  int add(const ModelDesc& d) {
    if (d.kind == 1 && (d.nslip < 1 || int(d.schmid.size()) != 6 * d.nslip))
      throw Error("bad slip-system table");
    if (d.kind == 1 && !evp_slip_count_built(d.nslip))
      throw Error("elasto-viscoplastic model with " + std::to_string(d.nslip) +
                  " slip systems is not built (supported: 1, 2, 3, 4, 6, 9, 12, 15, 18, 24)");
    descs_.push_back(d);
    return int(descs_.size()) - 1;
  }

  void set_microstructure(const double* rot, const int32_t* mop, hipStream_t s) {
    // point-major rotations are uploaded as they are and reordered to SoA on
    // the device (a stream-ordered temporary: no device-wide synchronisation)
    double* aos = nullptr;
    const size_t bytes = 9 * size_t(n_) * sizeof(double);
    MTL_HIP(hipMallocAsync(reinterpret_cast<void**>(&aos), bytes, s));
    MTL_HIP(hipMemcpyAsync(aos, rot, bytes, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(9 * n_)), dim3(kNBT), 0, s, n_, 9,
                       static_cast<const double*>(aos), rot_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipFreeAsync(aos, s));
    stream_sync(s);
    std::vector<std::vector<int>> pts(descs_.size());
    for (long long p = 0; p < n_; ++p) {
      const int m = mop ? mop[p] : 0;
      if (m < 0 || m >= int(descs_.size()))
        throw Error("model id out of range at point " + std::to_string(p));
      pts[m].push_back(int(p));
    }
    models_.clear();
    for (size_t m = 0; m < descs_.size(); ++m)
      models_.push_back(make_model<K>(descs_[m], pts[m], n_));
  }

  // initial tangent of every point and zeroed model state
  void init_tangent(K* D, hipStream_t s) {
    if (models_.empty()) throw Error("set the microstructure before initialize");
    MTL_HIP(hipMemsetAsync(D, 0, tan_elems(n_) * sizeof(K), s));
    for (auto& m : models_) m->init_tangent(rot_, D, s);
  }

  // every model on its points (multiphase.py:38-51): 0 as soon as one fails,
  // else the largest iteration count
  int calculate(const double* eps, const double* guess, double dt, double* sig, K* D,
                hipStream_t s) {
    int iters = 0;
    for (auto& m : models_) {
      const int it = m->calculate(rot_, eps, guess, dt, sig, D, s);
      if (!it) return 0;
      iters = std::max(iters, it);
    }
    return iters;
  }

  void commit(hipStream_t s) {
    for (auto& m : models_) m->commit(rot_, s);
  }

  // a state field scattered to the points of the models that have it, zero
  // elsewhere (multiphase.py:75-95)
  void state_field(int which, double* out, int* ncomp, hipStream_t s) {
    int width = 0;
    std::vector<std::pair<const Model<K>*, std::vector<double>>> parts;
    for (auto& m : models_) {
      std::vector<double> v;
      int nc = 0;
      if (m->field(which, rot_, v, nc, s)) {
        width = std::max(width, nc);
        parts.emplace_back(m.get(), std::move(v));
      }
    }
    if (width == 0) throw Error("field not available in any constitutive model");
    std::fill(out, out + size_t(width) * n_, 0.0);
    for (auto& pr : parts) {
      const auto& pts = pr.first->points();
      const size_t nc = pr.second.size() / std::max<size_t>(pts.size(), 1);
      for (size_t i = 0; i < pts.size(); ++i)
        for (size_t c = 0; c < nc; ++c) out[size_t(pts[i]) * width + c] = pr.second[i * nc + c];
    }
    *ncomp = width;
  }

 private:
  long long n_;
  double* rot_ = nullptr;
  std::vector<ModelDesc> descs_;
  std::vector<std::unique_ptr<Model<K>>> models_;
};

// f(i) for every i of idx, on one host thread each (the caller's thread
// takes the first), with the caller's device current; the first exception
// raised is rethrown once all have ended.
template <typename F>
void for_each_parallel(const std::vector<int>& idx, F f) {
  const size_t n = idx.size();
  if (n <= 1) {
    for (int i : idx) f(i);
    return;
  }
  int dev = 0;
  MTL_HIP(hipGetDevice(&dev));
  std::vector<std::exception_ptr> err(n);
  std::vector<std::thread> th;
  th.reserve(n - 1);
  for (size_t k = 1; k < n; ++k)
    th.emplace_back([&, k] {
      try {
        MTL_HIP(hipSetDevice(dev));
        f(idx[k]);
      } catch (...) {
        err[k] = std::current_exception();
      }
    });
  try {
    f(idx[0]);
  } catch (...) {
    err[0] = std::current_exception();
  }
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// full Mandel 6x6 per point (N x 36) of a tangent [n/64][21][64]
template <typename K>
void tangent_to_host(const K* D, long long n, double* out) {
  std::vector<K> blk(tan_elems(n));
  MTL_HIP(hipMemcpy(blk.data(), D, blk.size() * sizeof(K), hipMemcpyDeviceToHost));
  for (long long p = 0; p < n; ++p)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) out[p * 36 + a * 6 + b] = double(blk[tan_idx(sym6(a, b), p)]);
}

struct SolverBase {
  virtual ~SolverBase() = default;
  virtual int add_model(const ModelDesc& d) = 0;
  virtual void set_microstructure(const double* rot, const int32_t* mop) = 0;
  virtual void initialize() = 0;
  virtual void step(const mtl_step_params& p, mtl_step_report& r) = 0;
  virtual void commit(double scale) = 0;
  virtual void rollback() = 0;
  virtual void get_field(int which, double* out, int* ncomp) = 0;
  virtual void project(const double* in, double* out, const double mask[6]) = 0;
  virtual int minres(const double* b, double* x, const double mask[6], double rtol) = 0;
  virtual int constitutive(const double* eps, const double* guess, double dt, double* sig) = 0;
  // constitutive() on point-major host arrays (N x 6); point handles only
  virtual int constitutive_host(const double*, const double*, double, double*) {
    throw Error("host-array constitutive calls need a point handle (mtl_points_create)");
  }
};

// S: storage of the matvec spectrum and Lanczos vectors; K = work_t<S>: the
// tangent and the MINRES w / x vectors; A: arithmetic of the matvec FFT passes.
template <typename S, typename A = S>
class Solver final : public SolverBase {
 public:
  using K = work_t<S>;
  Solver(int nx, int ny, int nz) : sp64_(nx, ny, nz), ms_(sp64_.grid().n) {
    if constexpr (std::is_same<S, double>::value) {
      spk_ = &sp64_;  // one spectrum buffer for the residual and the matvec
    } else {
      own_spk_ = std::make_unique<Spectral<S, A>>(nx, ny, nz);
      spk_ = own_spk_.get();
    }
    minres_ = std::make_unique<Minres<S, A>>(*spk_);
    n_ = sp64_.grid().n;
    MTL_HIP(hipStreamCreate(&s_));
    auto dalloc = [&](double*& p, long long c) { MTL_HIP(hipMalloc(&p, c * sizeof(double))); };
    for (double** f : {&eps_, &sig_, &sig_old_, &eps_old_, &fluct_, &guess_, &b_}) dalloc(*f, 6 * n_);
    MTL_HIP(hipMalloc(&D_, tan_elems(n_) * sizeof(K)));
    MTL_HIP(hipMalloc(&D_old_, tan_elems(n_) * sizeof(K)));
    dalloc(partials_, kNormBlocks * 21);
    dalloc(shift_, 6);
    MTL_HIP(hipHostMalloc(&hpart_, kNormBlocks * 21 * sizeof(double), hipHostMallocDefault));
  }
  ~Solver() override {
    for (double* p : {eps_, sig_, sig_old_, eps_old_, fluct_, guess_, b_, partials_, shift_})
      if (p) (void)hipFree(p);
    if (D_) (void)hipFree(D_);
    if (D_old_) (void)hipFree(D_old_);
    if (hpart_) (void)hipHostFree(hpart_);
    if (s_) (void)hipStreamDestroy(s_);
  }

  int add_model(const ModelDesc& d) override { return ms_.add(d); }

  void set_microstructure(const double* rot, const int32_t* mop) override {
    ms_.set_microstructure(rot, mop, s_);
  }

  void initialize() override {
    ms_.init_tangent(D_, s_);
    for (double* f : {eps_, sig_, sig_old_, eps_old_, fluct_, guess_, b_})
      MTL_HIP(hipMemsetAsync(f, 0, 6 * n_ * sizeof(double), s_));
    MTL_HIP(hipMemcpyAsync(D_old_, D_, tan_elems(n_) * sizeof(K), hipMemcpyDeviceToDevice, s_));
    sig_stale_ = eps_stale_ = d_stale_ = guess_is_stress_ = false;
    stream_sync(s_);
  }

  // Multiphase dispatch (multiphase.py:38-51): fails as soon as a model fails
  int constitutive_all(const double* eps, const double* guess, double dt, double* sig) {
    d_stale_ = false;  // D_ is rewritten (or restored on failure)
    return ms_.calculate(eps, guess, dt, sig, D_, s_);
  }

  struct Norms {
    double maxb, mean[6], maxdev2, maxdx2;
  };

  Norms norms(bool with_dx) {
    const int nb = grid_for(n_);
    hipLaunchKernelGGL(k_newton_norms<K>, dim3(nb), dim3(kNBT), 0, s_, n_, b_, sig_, sig_old_, D_,
                       with_dx ? minres_->x() : nullptr, partials_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipMemcpyAsync(hpart_, partials_, nb * kNormWidth * sizeof(double),
                           hipMemcpyDeviceToHost, s_));
    stream_sync(s_);
    Norms r{};
    for (int i = 0; i < nb; ++i) {
      const double* q = hpart_ + i * kNormWidth;
      r.maxb = std::max(r.maxb, q[0]);
      for (int c = 0; c < 6; ++c) r.mean[c] += q[1 + c];
      r.maxdev2 = std::max(r.maxdev2, q[7]);
      r.maxdx2 = std::max(r.maxdx2, q[8]);
    }
    for (double& m : r.mean) m /= double(n_);
    return r;
  }

  static double von_mises_of_mean(const double* m) {
    // stress.mean().dev.norm * sqrt(1.5)  (small_strain_fft.py:141-143)
    const double tr = (m[0] + m[1] + m[2]) / 3.0;
    double s = 0.0;
    for (int c = 0; c < 6; ++c) {
      const double v = c < 3 ? m[c] - tr : m[c];
      s += v * v;
    }
    return std::sqrt(s) * std::sqrt(1.5);
  }

  void tangent_mean(double out36[36]) {
    const int nb = grid_for(n_);
    hipLaunchKernelGGL(k_tangent_sum<K>, dim3(nb), dim3(kNBT), 0, s_, n_, cur_D(), partials_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipMemcpyAsync(hpart_, partials_, nb * 21 * sizeof(double), hipMemcpyDeviceToHost, s_));
    stream_sync(s_);
    double sum[21] = {};
    for (int i = 0; i < nb; ++i)
      for (int k = 0; k < 21; ++k) sum[k] += hpart_[i * 21 + k];
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) out36[a * 6 + b] = sum[sym6(a, b)] / double(n_);
  }

  void residual(const Vec6& applied, const Vec6& mask) {
    // b = Ghat * (Sigma_applied - sigma), always in FP64
    sp64_.project(sig_, b_, applied, -1.0, mask, s_);
  }

  // One load increment (small_strain_fft.py:106-211) as resumable phases, so
  // that the MINRES solves of several same-size solvers can run together:
  // step_begin (predictor), then while !step_finished(): step_pre_minres and,
  // when it returns true, one MINRES solve followed by step_post_minres.
  void step(const mtl_step_params& P, mtl_step_report& R) override {
    step_begin(P, R);
    while (!ss_.finished) {
      if (!step_pre_minres(R)) continue;
      int istop = 0;
      const int itn = minres_->solve(b_, D_, ss_.mask, ss_.lin_tol, 5LL * 6LL * n_, s_, &istop);
      if (debug_enabled())
        std::fprintf(stderr, "[mtl] step %d newton %d: minres %d iterations (istop %d)\n",
                     ss_.P.time_step_id, ss_.iteration, itn, istop);
      step_post_minres(itn, R);
    }
    step_end(R);
  }

  // The same increment for several solvers of one grid and form at once.
  // The per-solver phases (constitutive update, residual, norms: each ends
  // with a host read of a few reductions) run on one host thread per solver,
  // each on its own stream, so that their launches and synchronisations
  // overlap; only the MINRES solves are merged into shared launches.
  static void step_batch(const std::vector<Solver*>& hs, const mtl_step_params* P,
                         mtl_step_report* R) {
    const int nb = int(hs.size());
    std::vector<int> all(nb);
    for (int i = 0; i < nb; ++i) all[i] = i;
    for_each_parallel(all, [&](int i) { hs[i]->step_begin(P[i], R[i]); });
    std::vector<typename Minres<S, A>::BatchItem> items;
    std::vector<int> live, who, itns;
    std::vector<char> solve(nb);
    while (true) {
      live.clear();
      bool first = false;  // only a first Newton iteration has work before MINRES
      for (int i = 0; i < nb; ++i)
        if (!hs[i]->ss_.finished) {
          live.push_back(i);
          first = first || hs[i]->ss_.iteration == 0;
        }
      if (live.empty()) break;
      auto pre = [&](int i) { solve[i] = hs[i]->step_pre_minres(R[i]) ? 1 : 0; };
      if (first)
        for_each_parallel(live, pre);
      else
        for (int i : live) pre(i);
      items.clear();
      who.clear();
      for (int i : live)
        if (solve[i]) {
          Solver& h = *hs[i];
          items.push_back({h.minres_.get(), h.b_, h.D_, h.ss_.mask, h.ss_.lin_tol});
          who.push_back(i);
        }
      if (items.empty()) continue;
      itns.assign(nb, 0);
      std::vector<int> got(items.size(), 0);
      // every solver's stream has drained (its residual norms were read on
      // the host); the batch runs on the first solver's stream
      Minres<S, A>::solve_batch(items, 5LL * 6LL * hs[0]->n_, hs[who[0]]->s_, got.data());
      for (size_t k = 0; k < who.size(); ++k) itns[who[k]] = got[k];
      for_each_parallel(who, [&](int i) { hs[i]->step_post_minres(itns[i], R[i]); });
    }
    for (int i = 0; i < nb; ++i) hs[i]->step_end(R[i]);
  }

  const Grid& grid() const { return sp64_.grid(); }

 private:
  struct StepState {
    mtl_step_params P;
    Vec6 de, ds, applied, mask;
    double dt = 0, lin_tol = 0, eq_err = 0, vm = 0, old_eq = 0;
    int iteration = 0, max_constit = 0;
    bool converged = false, finished = false;
  };
  StepState ss_;

  void step_begin(const mtl_step_params& P, mtl_step_report& R) {
    std::memset(&R, 0, sizeof(R));
    ss_ = StepState{};
    ss_.P = P;
    Vec6 pred{};
    std::memcpy(ss_.de.v, P.strain_increment, sizeof(ss_.de.v));
    std::memcpy(ss_.ds.v, P.stress_increment, sizeof(ss_.ds.v));
    std::memcpy(ss_.applied.v, P.applied_stress, sizeof(ss_.applied.v));
    std::memcpy(ss_.mask.v, P.stress_mask, sizeof(ss_.mask.v));
    ss_.dt = P.time_increment;
    // predictor (small_strain_fft.py:117-123)
    bool all_small = true;
    for (double v : ss_.ds.v) all_small = all_small && (v < 1.0e-14);
    if (!all_small) {
      double dm[36], dinv[36];
      tangent_mean(dm);
      if (!inverse6_host(dm, dinv)) throw Error("mean tangent is singular");
      for (int a = 0; a < 6; ++a) {
        double acc = 0.0;
        for (int b = 0; b < 6; ++b) acc += dinv[a * 6 + b] * ss_.ds.v[b];
        pred.v[a] = acc;
      }
    }
    const int nb = grid_for(n_);
    hipLaunchKernelGGL(k_predictor, dim3(nb), dim3(kNBT), 0, s_, n_, eps_old_, fluct_, eps_, ss_.de,
                       pred, P.reset_fluctuation);
    MTL_LAUNCH_CHECK();
    eps_stale_ = false;
    hipLaunchKernelGGL(k_guess<K>, dim3(nb), dim3(kNBT), 0, s_, n_, P.time_step_id == 0 ? 1 : 0,
                       sig_old_, cur_D(), ss_.ds, ss_.de, guess_);
    MTL_LAUNCH_CHECK();
    guess_is_stress_ = false;
    // The FP32 form's Krylov solves cannot honour relative tolerances much
    // below 1e-5 on the singular pore systems, so its inner tolerance is
    // floored there.  The Newton residual, the equilibrium test and the
    // strain/stress fields stay FP64 in every form, so each Newton iteration
    // is an FP64-residual refinement step.
    constexpr double kFloor = std::is_same<S, float>::value ? kMixedKrylovFloor : 0.0;
    ss_.lin_tol = std::max(P.linear_solver_tolerance, kFloor);
    R.linear_solver_tolerance_used = ss_.lin_tol;
  }

  // Next Newton iteration up to its MINRES solve: false when the increment
  // ended instead (the first constitutive update failed).
  bool step_pre_minres(mtl_step_report& R) {
    StepState& q = ss_;
    ++q.iteration;
    if (q.iteration == 1) {
      const int it = constitutive_all(eps_, guess_, q.dt, sig_);
      sig_stale_ = false;
      if (!it) {
        restore_tangent();
        R.constitutive_failed = 1;
        q.finished = true;
        return false;
      }
      residual(q.applied, q.mask);
      const Norms nm = norms(false);
      q.vm = von_mises_of_mean(nm.mean);
      q.eq_err = nm.maxb / q.vm;
      // the next constitutive call's guess is this stress; it rewrites sig_
      std::swap(guess_, sig_);
      guess_is_stress_ = true;
      q.max_constit = std::max(q.max_constit, it);
    }
    q.old_eq = std::max(q.eq_err, q.P.global_tolerance);
    return true;
  }

  // The rest of the Newton iteration after its MINRES solve of itn iterations.
  void step_post_minres(int itn, mtl_step_report& R) {
    StepState& q = ss_;
    const mtl_step_params& P = q.P;
    R.krylov_iterations += itn;
    constexpr bool kMixed = !std::is_same<K, double>::value;
    const int nb = grid_for(n_);
    if (kMixed) {
      // The exact correction is compatible, with zero mean in every
      // strain-controlled component (Ghat[0] = diag(mask)); FP32 rounding
      // leaves a small relative mean there, removed here in FP64 (on the
      // device: no host round trip).
      hipLaunchKernelGGL(k_sum6<K>, dim3(nb), dim3(kNBT), 0, s_, n_, minres_->x(), partials_);
      MTL_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_mean_shift, dim3(1), dim3(kNBT), 0, s_, nb, partials_, n_, q.mask, shift_);
      MTL_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_add_correction<K>, dim3(nb), dim3(kNBT), 0, s_, n_, minres_->x(), eps_,
                       fluct_, kMixed ? shift_ : static_cast<const double*>(nullptr));
    MTL_LAUNCH_CHECK();
    const int it = constitutive_all(eps_, guess_, q.dt, sig_);
    sig_stale_ = false;
    if (!it) {
      // sig_ is left as the last accepted stress (the failed call's guess)
      if (guess_is_stress_) std::swap(guess_, sig_);
      restore_tangent();
      R.constitutive_failed = 1;
      q.finished = true;
      return;
    }
    q.max_constit = std::max(q.max_constit, it);
    residual(q.applied, q.mask);
    const Norms nm = norms(true);
    q.vm = von_mises_of_mean(nm.mean);
    const double time_avg = (P.summed_von_mises_stress + q.vm) / double(P.time_step_id + 1);
    if (q.vm / time_avg < 1.0e-12) q.vm = time_avg;
    q.eq_err = nm.maxb / q.vm;
    double max_ds = std::sqrt(nm.maxdev2) * std::sqrt(1.5);
    if (max_ds / q.vm < 1.0e-12) max_ds = q.vm;
    const double strain_err = std::sqrt(nm.maxdx2) / max_ds;
    const double projected = strain_err * q.eq_err / q.old_eq;
    if (q.eq_err < P.global_tolerance &&
        (strain_err < P.strain_correction_tolerance ||
         projected < P.strain_correction_tolerance || P.time_step_id == 0 || q.eq_err < 1e-8))
      q.converged = true;
    if (debug_enabled())
      std::fprintf(stderr, "[mtl]   eq_err %.3e strain_err %.3e projected %.3e -> %s\n", q.eq_err,
                   strain_err, projected, q.converged ? "converged" : "continue");
    constexpr int kMaxNewton = 100;  // small_strain_fft.py:125
    if (!q.converged && q.iteration < kMaxNewton) {
      std::swap(guess_, sig_);  // as above: sig_ is rewritten before it is read
      guess_is_stress_ = true;
    }
    R.scaled_strain_error = strain_err;
    R.projected_scaled_strain_error = projected;
    if (q.converged || q.iteration == kMaxNewton) q.finished = true;
  }

  void step_end(mtl_step_report& R) {
    profiler().flush();
    R.converged = ss_.converged ? 1 : 0;
    R.newton_iterations = ss_.iteration;
    R.max_constitutive_iterations = ss_.max_constit;
    R.mean_von_mises_stress = ss_.vm;
    R.equilibrium_error = ss_.eq_err;
  }

 public:
  void commit(double scale) override {
    ms_.commit(s_);
    // the current stress, strain and tangent become the committed ones by
    // swapping buffers; until the next step rewrites them the current values
    // are read from the committed buffers (cur_*)
    std::swap(sig_old_, sig_);
    std::swap(eps_old_, eps_);
    std::swap(D_old_, D_);
    sig_stale_ = eps_stale_ = d_stale_ = true;
    scale6(fluct_, scale);
    stream_sync(s_);
  }

  void rollback() override {
    scale6(fluct_, 0.5);
    stream_sync(s_);
  }

  void get_field(int which, double* out, int* ncomp) override {
    if (which == kFieldStress || which == kFieldStrain) {
      soa_to_host(which == kFieldStress ? cur_sig() : cur_eps(), 6, out);
      *ncomp = 6;
      return;
    }
    if (which == kFieldTangent) {
      stream_sync(s_);
      tangent_to_host(cur_D(), n_, out);
      *ncomp = 36;
      return;
    }
    ms_.state_field(which, out, ncomp, s_);
  }

  void project(const double* in, double* out, const double m[6]) override {
    Vec6 mask, zero{};
    std::memcpy(mask.v, m, sizeof(mask.v));
    spk_->project(in, out, zero, 1.0, mask, s_);
    stream_sync(s_);
  }

  int minres(const double* b, double* x, const double m[6], double rtol) override {
    Vec6 mask;
    std::memcpy(mask.v, m, sizeof(mask.v));
    int istop = 0;
    const int itn = minres_->solve(b, cur_D(), mask, rtol, 5LL * 6LL * n_, s_, &istop);
    hipLaunchKernelGGL(k_unblock6<K>, dim3(grid_for(6 * n_)), dim3(kNBT), 0, s_, n_, minres_->x(),
                       x);
    MTL_LAUNCH_CHECK();
    stream_sync(s_);
    return itn;
  }

  int constitutive(const double* eps, const double* guess, double dt, double* sig) override {
    const int it = constitutive_all(eps, guess, dt, sig);
    stream_sync(s_);
    return it;
  }

 private:
  void scale6(double* a, double f) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(6 * n_)), dim3(kNBT), 0, s_, 6 * n_, a, f);
    MTL_LAUNCH_CHECK();
  }
  void restore_tangent() {
    MTL_HIP(hipMemcpyAsync(D_, D_old_, tan_elems(n_) * sizeof(K), hipMemcpyDeviceToDevice, s_));
  }
  // Reorders on the device into b_ (the Newton residual, rewritten before it
  // is read in every step) and copies the point-major result to `out`.
  void soa_to_host(const double* d, int nc, double* out) {
    if (nc > 6 || d == b_) throw Error("soa_to_host: bad field");
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(nc * n_)), dim3(kNBT), 0, s_, n_, nc, d, b_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipMemcpyAsync(out, b_, size_t(nc) * n_ * sizeof(double), hipMemcpyDeviceToHost, s_));
    stream_sync(s_);
  }

  Spectral<double> sp64_;
  std::unique_ptr<Spectral<S, A>> own_spk_;
  Spectral<S, A>* spk_ = nullptr;
  std::unique_ptr<Minres<S, A>> minres_;
  long long n_ = 0;
  hipStream_t s_ = nullptr;
  double *eps_ = nullptr, *sig_ = nullptr, *sig_old_ = nullptr, *eps_old_ = nullptr,
         *fluct_ = nullptr, *guess_ = nullptr, *b_ = nullptr;
  K *D_ = nullptr, *D_old_ = nullptr;
  // guess_ holds the stress of the last accepted constitutive call (the
  // buffers were swapped instead of copied) rather than the predictor's guess
  bool guess_is_stress_ = false;
  // after a commit the current stress, strain and tangent live in the
  // committed buffers until the next step (or constitutive call) rewrites them
  bool sig_stale_ = false, eps_stale_ = false, d_stale_ = false;
  const double* cur_sig() const { return sig_stale_ ? sig_old_ : sig_; }
  const double* cur_eps() const { return eps_stale_ ? eps_old_ : eps_; }
  K* cur_D() const { return d_stale_ ? D_old_ : D_; }
  double* partials_ = nullptr;
  double* shift_ = nullptr;  // mean shift of the mixed-form correction (6)
  double* hpart_ = nullptr;
  ModelSet<K> ms_;
};

// Constitutive models over n material points and no grid: the handle behind
// the models' own protocol (initialize, calculate_stress_and_tangent,
// update_state_variables of elastic.py:23-46, elastic_viscoplastic.py:52-271,
// isotropic_elastic_plastic.py:55-179).  K: tangent storage.
template <typename K>
class Points final : public SolverBase {
 public:
  explicit Points(long long n) : n_(n), ms_(n) {
    if (n < 1) throw Error("a point handle needs at least one point");
    MTL_HIP(hipStreamCreate(&s_));
    for (double** f : {&eps_, &guess_, &sig_, &aos_})
      MTL_HIP(hipMalloc(f, 6 * size_t(n) * sizeof(double)));
    MTL_HIP(hipMalloc(&D_, tan_elems(n) * sizeof(K)));
  }
  ~Points() override {
    for (double* p : {eps_, guess_, sig_, aos_})
      if (p) (void)hipFree(p);
    if (D_) (void)hipFree(D_);
    if (s_) (void)hipStreamDestroy(s_);
  }

  int add_model(const ModelDesc& d) override { return ms_.add(d); }
  void set_microstructure(const double* rot, const int32_t* mop) override {
    ms_.set_microstructure(rot, mop, s_);
  }
  void initialize() override {
    ms_.init_tangent(D_, s_);
    MTL_HIP(hipMemsetAsync(sig_, 0, 6 * size_t(n_) * sizeof(double), s_));
    stream_sync(s_);
  }
  void step(const mtl_step_params&, mtl_step_report&) override { no_grid(); }
  void commit(double) override {
    ms_.commit(s_);
    stream_sync(s_);
  }
  void rollback() override { no_grid(); }
  void get_field(int which, double* out, int* ncomp) override {
    if (which == kFieldStress) {
      download6(sig_, out);
      *ncomp = 6;
    } else if (which == kFieldStrain) {
      download6(eps_, out);
      *ncomp = 6;
    } else if (which == kFieldTangent) {
      stream_sync(s_);
      tangent_to_host(D_, n_, out);
      *ncomp = 36;
    } else {
      ms_.state_field(which, out, ncomp, s_);
    }
  }
  void project(const double*, double*, const double*) override { no_grid(); }
  int minres(const double*, double*, const double*, double) override {
    no_grid();
    return 0;
  }
  int constitutive(const double* eps, const double* guess, double dt, double* sig) override {
    const int it = ms_.calculate(eps, guess, dt, sig, D_, s_);
    stream_sync(s_);
    return it;
  }
  int constitutive_host(const double* eps, const double* guess, double dt, double* sig) override {
    upload6(eps, eps_);
    upload6(guess, guess_);
    const int it = ms_.calculate(eps_, guess_, dt, sig_, D_, s_);
    stream_sync(s_);
    if (it) download6(sig_, sig);
    return it;
  }

 private:
  [[noreturn]] static void no_grid() {
    throw Error("a point handle has no grid: step, rollback, project and minres need "
                "mtl_solver_create");
  }
  void upload6(const double* host, double* soa) {
    const size_t bytes = 6 * size_t(n_) * sizeof(double);
    MTL_HIP(hipMemcpyAsync(aos_, host, bytes, hipMemcpyHostToDevice, s_));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(6 * n_)), dim3(kNBT), 0, s_, n_, 6,
                       static_cast<const double*>(aos_), soa);
    MTL_LAUNCH_CHECK();
  }
  void download6(const double* soa, double* host) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(6 * n_)), dim3(kNBT), 0, s_, n_, 6, soa, aos_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipMemcpyAsync(host, aos_, 6 * size_t(n_) * sizeof(double), hipMemcpyDeviceToHost, s_));
    stream_sync(s_);
  }
  long long n_;
  hipStream_t s_ = nullptr;
  double *eps_ = nullptr, *guess_ = nullptr, *sig_ = nullptr, *aos_ = nullptr;
  K* D_ = nullptr;
  ModelSet<K> ms_;
};

Profiler& profiler() {
  static Profiler p;
  return p;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

bool inverse6_host(const double* A, double* out) {
  double m[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 12; ++j) m[i][j] = j < 6 ? A[i * 6 + j] : (j - 6 == i ? 1.0 : 0.0);
  for (int c = 0; c < 6; ++c) {
    int piv = c;
    for (int r = c + 1; r < 6; ++r)
      if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
    if (m[piv][c] == 0.0) return false;
    if (piv != c)
      for (int j = 0; j < 12; ++j) std::swap(m[c][j], m[piv][j]);
    const double inv = 1.0 / m[c][c];
    for (int j = 0; j < 12; ++j) m[c][j] *= inv;
    for (int r = 0; r < 6; ++r)
      if (r != c) {
        const double f = m[r][c];
        if (f != 0.0)
          for (int j = 0; j < 12; ++j) m[r][j] -= f * m[c][j];
      }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) out[i * 6 + j] = m[i][j + 6];
  return true;
}

}  // namespace mtl

// ====================================================================== C ABI
struct mtl_solver {
  std::unique_ptr<mtl::SolverBase> impl;
};

namespace {
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    mtl::g_last_error = e.what();
    return -1;
  } catch (...) {
    mtl::g_last_error = "unknown error";
    return -1;
  }
}
int bad(const char* msg) {
  mtl::g_last_error = msg;
  return -2;
}
}  // namespace

namespace {
template <typename Sv>
bool try_step_batch(mtl_solver* const* hs, int count, const mtl_step_params* p,
                    mtl_step_report* r) {
  std::vector<Sv*> v;
  for (int i = 0; i < count; ++i) {
    auto* s = dynamic_cast<Sv*>(hs[i]->impl.get());
    if (s == nullptr) return false;
    v.push_back(s);
  }
  const mtl::Grid& g = v[0]->grid();
  for (auto* s : v)
    if (s->grid().nx != g.nx || s->grid().ny != g.ny || s->grid().nz != g.nz)
      throw mtl::Error("a step batch needs solvers of one grid");
  Sv::step_batch(v, p, r);
  return true;
}
}  // namespace

extern "C" {

const char* mtl_last_error(void) { return mtl::g_last_error.c_str(); }

int mtl_solver_create(int nx, int ny, int nz, int precision, mtl_solver** out) {
  if (!out) return bad("null output handle");
  if (precision != MTL_FP64 && precision != MTL_MIXED && precision != MTL_REFINED)
    return bad("unknown precision");
  return guarded([&] {
    auto h = std::make_unique<mtl_solver>();
    if (precision == MTL_FP64)
      h->impl = std::make_unique<mtl::Solver<double>>(nx, ny, nz);
    else if (precision == MTL_MIXED)
      h->impl = std::make_unique<mtl::Solver<float>>(nx, ny, nz);
    else
      h->impl = std::make_unique<mtl::Solver<mtl::f48, double>>(nx, ny, nz);
    *out = h.release();
  });
}

int mtl_points_create(long long num_points, int precision, mtl_solver** out) {
  if (!out) return bad("null output handle");
  if (precision != MTL_FP64 && precision != MTL_MIXED && precision != MTL_REFINED)
    return bad("unknown precision");
  if (num_points < 1 || num_points > (1LL << 31) - 1) return bad("number of points out of range");
  return guarded([&] {
    auto h = std::make_unique<mtl_solver>();
    if (precision == MTL_FP64)
      h->impl = std::make_unique<mtl::Points<double>>(num_points);
    else
      h->impl = std::make_unique<mtl::Points<float>>(num_points);
    *out = h.release();
  });
}

int mtl_solver_destroy(mtl_solver* h) {
  return guarded([&] { delete h; });
}

int mtl_solver_add_elastic(mtl_solver* h, const double stiffness[36], int* model_id) {
  if (!h || !stiffness || !model_id) return bad("null argument");
  return guarded([&] {
    mtl::ModelDesc d;
    d.kind = 0;
    std::memcpy(d.C, stiffness, sizeof(d.C));
    *model_id = h->impl->add_model(d);
  });
}

namespace {
int add_evp(mtl_solver* h, const double stiffness[36], int num_slip, const double* schmid,
            double reference_slip_rate, double rate_exponent, double slip_resistance,
            const double* per_system, int hardening, const double hardening_props[4],
            int* model_id) {
  if (!h || !stiffness || !schmid || !model_id) return bad("null argument");
  if (hardening < 0 || hardening > 3) return bad("unknown hardening law");
  if (num_slip < 1) return bad("EVP model needs at least one slip system");
  return guarded([&] {
    mtl::ModelDesc d;
    d.kind = 1;
    std::memcpy(d.C, stiffness, sizeof(d.C));
    d.nslip = num_slip;
    d.schmid.assign(schmid, schmid + 6 * num_slip);
    d.gamma0 = reference_slip_rate;
    d.n = rate_exponent;
    d.g0 = slip_resistance;
    if (per_system) d.g0_sys.assign(per_system, per_system + num_slip);
    d.hardening = hardening;
    if (hardening_props) std::memcpy(d.hp, hardening_props, sizeof(d.hp));
    *model_id = h->impl->add_model(d);
  });
}
}  // namespace

int mtl_solver_add_evp(mtl_solver* h, const double stiffness[36], int num_slip,
                       const double* schmid, double reference_slip_rate, double rate_exponent,
                       double slip_resistance, int hardening, const double hardening_props[4],
                       int* model_id) {
  return add_evp(h, stiffness, num_slip, schmid, reference_slip_rate, rate_exponent,
                 slip_resistance, nullptr, hardening, hardening_props, model_id);
}

int mtl_solver_add_evp_systems(mtl_solver* h, const double stiffness[36], int num_slip,
                               const double* schmid, double reference_slip_rate,
                               double rate_exponent, const double* slip_resistances,
                               int hardening, const double hardening_props[4], int* model_id) {
  if (!slip_resistances) return bad("null argument");
  return add_evp(h, stiffness, num_slip, schmid, reference_slip_rate, rate_exponent,
                 slip_resistances[0], slip_resistances, hardening, hardening_props, model_id);
}

int mtl_solver_add_j2(mtl_solver* h, const double stiffness[36], double shear_modulus,
                      double bulk_modulus, double yield_stress, int hardening,
                      const double hardening_props[4], int* model_id) {
  if (!h || !stiffness || !model_id) return bad("null argument");
  if (hardening < 0 || hardening > 3) return bad("unknown hardening law");
  if (!(shear_modulus > 0.0)) return bad("J2 model needs a positive shear modulus");
  return guarded([&] {
    mtl::ModelDesc d;
    d.kind = 2;
    std::memcpy(d.C, stiffness, sizeof(d.C));
    d.shear_modulus = shear_modulus;
    d.bulk_modulus = bulk_modulus;
    d.g0 = yield_stress;
    d.hardening = hardening;
    if (hardening_props) std::memcpy(d.hp, hardening_props, sizeof(d.hp));
    *model_id = h->impl->add_model(d);
  });
}

int mtl_solver_set_microstructure(mtl_solver* h, const double* rotations,
                                  const int32_t* model_of_point) {
  if (!h || !rotations) return bad("null argument");
  return guarded([&] { h->impl->set_microstructure(rotations, model_of_point); });
}

int mtl_solver_initialize(mtl_solver* h) {
  if (!h) return bad("null handle");
  return guarded([&] { h->impl->initialize(); });
}

int mtl_solver_step(mtl_solver* h, const mtl_step_params* p, mtl_step_report* r) {
  if (!h || !p || !r) return bad("null argument");
  return guarded([&] { h->impl->step(*p, *r); });
}

int mtl_solver_step_batch(mtl_solver* const* handles, int count, const mtl_step_params* params,
                          mtl_step_report* reports) {
  if (!handles || !params || !reports || count < 1) return bad("null argument or empty batch");
  for (int i = 0; i < count; ++i)
    if (!handles[i]) return bad("null handle in batch");
  return guarded([&] {
    for (int i = 0; i < count; ++i)
      for (int j = 0; j < i; ++j)
        if (handles[i] == handles[j]) throw mtl::Error("a handle appears twice in a step batch");
    if (try_step_batch<mtl::Solver<double>>(handles, count, params, reports)) return;
    if (try_step_batch<mtl::Solver<float>>(handles, count, params, reports)) return;
    if (try_step_batch<mtl::Solver<mtl::f48, double>>(handles, count, params, reports)) return;
    throw mtl::Error("a step batch needs grid solvers of one precision");
  });
}

int mtl_solver_commit(mtl_solver* h, double fluctuation_scale) {
  if (!h) return bad("null handle");
  return guarded([&] { h->impl->commit(fluctuation_scale); });
}

int mtl_solver_rollback(mtl_solver* h) {
  if (!h) return bad("null handle");
  return guarded([&] { h->impl->rollback(); });
}

int mtl_solver_get_field(mtl_solver* h, int field, double* out, int* ncomp) {
  if (!h || !out || !ncomp) return bad("null argument");
  return guarded([&] { h->impl->get_field(field, out, ncomp); });
}

int mtl_solver_project(mtl_solver* h, const double* in_dev, double* out_dev,
                       const double stress_mask[6]) {
  if (!h || !in_dev || !out_dev || !stress_mask) return bad("null argument");
  return guarded([&] { h->impl->project(in_dev, out_dev, stress_mask); });
}

int mtl_solver_minres(mtl_solver* h, const double* b_dev, double* x_dev,
                      const double stress_mask[6], double rtol, int* iterations) {
  if (!h || !b_dev || !x_dev || !stress_mask || !iterations) return bad("null argument");
  return guarded([&] { *iterations = h->impl->minres(b_dev, x_dev, stress_mask, rtol); });
}

int mtl_profile_enable(int enable) {
  return guarded([&] { mtl::profiler().set_enabled(enable != 0); });
}

int mtl_profile_reset(void) {
  return guarded([&] { mtl::profiler().reset(); });
}

int mtl_profile_read(int kernel, double* total_ms, long long* launches) {
  if (!total_ms || !launches) return bad("null argument");
  if (kernel < 0 || kernel >= mtl::kNumKernelIds) return bad("unknown kernel id");
  return guarded([&] { mtl::profiler().read(kernel, total_ms, launches); });
}

int mtl_profile_durations(int kernel, float* ms_out, long long capacity, long long* count) {
  if (!count || (capacity > 0 && !ms_out)) return bad("null argument");
  if (kernel < 0 || kernel >= mtl::kNumKernelIds) return bad("unknown kernel id");
  return guarded([&] { *count = mtl::profiler().read_durations(kernel, ms_out, capacity); });
}

int mtl_solver_constitutive_host(mtl_solver* h, const double* strain, const double* guess,
                                 double time_increment, double* stress, int* iterations) {
  if (!h || !strain || !guess || !stress || !iterations) return bad("null argument");
  return guarded(
      [&] { *iterations = h->impl->constitutive_host(strain, guess, time_increment, stress); });
}

int mtl_solver_constitutive(mtl_solver* h, const double* strain_dev, const double* guess_dev,
                            double time_increment, double* stress_dev, int* iterations) {
  if (!h || !strain_dev || !guess_dev || !stress_dev || !iterations) return bad("null argument");
  return guarded(
      [&] { *iterations = h->impl->constitutive(strain_dev, guess_dev, time_increment, stress_dev); });
}

}  // extern "C"
