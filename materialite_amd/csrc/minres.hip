// Device MINRES: the recurrence of SciPy's scipy/sparse/linalg/_isolve/minres.py
// (Paige & Saunders 1975) for x0 = 0, no preconditioner, no shift -- the call
// made at small_strain_fft.py:289-294 of the reference.  Scalars live in a
// MinresState in device memory; each iteration is
//   matvec (5 FFT kernels, epilogue u = Av - (beta/oldb) r1, partials <v,u>)
//   k_alpha      alfa = <v,u>
//   k_lanczos    y = u - (alfa/beta) r2 (in place of u), partials |y|^2
//   k_givens     beta_{k+1}, plane rotation, w/x coefficients
//   k_update     w = (v - oldeps w1 - delta w2)/gamma, x += phi w, partials |x|^2
//   k_tests      SciPy's stopping tests -> done
// and the host polls `done` once per chunk of iterations.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "minres.hpp"
#include "profiler.hpp"

namespace mtl {

namespace {

constexpr int kVBT = 256;      // vector kernels
constexpr int kSBT = 1024;     // one-workgroup scalar kernels (up to 8192 partials to sum)
constexpr int kVBlocks = 2048;  // grid-stride blocks for vector kernels
constexpr double kEps = DBL_EPSILON;

__device__ double sum_partials(const double* partials, int count, double* red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += kSBT) acc += partials[i];
  return block_sum<kSBT>(acc, red);
}

// r1 is zeroed although iteration 1 multiplies it by 0 (c_r1 = 0): stale
// memory may hold inf/NaN bit patterns and 0 * inf would poison u.  The
// vectors hold 6 components blocked by 64 points (blk_idx); b is SoA [6][n].
// Entries of the padded last block are set to zero and stay zero.
template <typename S>
__global__ void __launch_bounds__(kVBT) k_init(long long nb6, long long n, const double* b,
                                                RVec<S> r1, RVec<S> r2, work_t<S>* x,
                                                work_t<S>* w, work_t<S>* w2, double* partials,
                                                const BatchSlot<S>* slots) {
  using T = work_t<S>;
  if (slots != nullptr) {  // a batch: the slot's vectors (first iteration's roles)
    const BatchSlot<S>& sl = slots[blockIdx.y];
    b = sl.b;
    r1 = sl.r1;
    r2 = sl.r2;
    x = sl.x;
    w = sl.w1;
    w2 = sl.w2;
    partials = sl.partials;
  }
  __shared__ double red[kVBT / 64];
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)kVBT + threadIdx.x; i < nb6;
       i += (long long)gridDim.x * kVBT) {
    const long long q = i >> 6;
    const long long p = (q / 6) * 64 + (i & 63);
    const int c = int(q % 6);
    const double v = p < n ? b[c * n + p] : 0.0;
    r1.put(i, 0.0);
    r2.put(i, v);
    x[i] = T(0);
    w[i] = T(0);
    w2[i] = T(0);
    const double vs = r2.template get<double>(i);  // as stored
    acc += vs * vs;
  }
  const double t = block_sum<kVBT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <typename S>
__global__ void __launch_bounds__(kSBT) k_start(MinresState* st, const double* partials, int np,
                                                 double rtol, long long maxiter,
                                                 const BatchSlot<S>* slots) {
  __shared__ double red[kSBT / 64];
  if (slots != nullptr) {
    st = slots[blockIdx.x].st;
    partials = slots[blockIdx.x].partials;
    rtol = slots[blockIdx.x].rtol;
  }
  const double b2 = sum_partials(partials, np, red);
  if (threadIdx.x != 0) return;
  MinresState z{};
  z.rtol = rtol;
  z.maxiter = int(maxiter > 2147483647LL ? 2147483647LL : maxiter);
  z.gmin = DBL_MAX;
  z.cs = -1.0;
  z.sn = 0.0;
  if (b2 == 0.0) {  // beta1 == 0 (and |b| == 0): x = 0
    z.done = 1;
    *st = z;
    return;
  }
  const double beta1 = sqrt(b2);
  z.beta1 = beta1;
  z.beta = beta1;
  z.phibar = beta1;
  z.rhs1 = beta1;
  z.s = 1.0 / beta1;
  z.c_r1 = 0.0;
  *st = z;
}

template <typename S>
__global__ void __launch_bounds__(kSBT) k_alpha(MinresState* st, const double* partials, int np,
                                                 const BatchSlot<S>* slots) {
  __shared__ double red[kSBT / 64];
  if (slots != nullptr) {
    st = slots[blockIdx.x].st;
    partials = slots[blockIdx.x].partials;
  }
  if (st->done) return;
  const double a = sum_partials(partials, np, red);
  if (threadIdx.x != 0) return;
  st->alfa = a;
  st->c_r2 = a / st->beta;
}

// T: arithmetic (FP32 in the FP32 form, FP64 otherwise)
template <typename S, typename T>
__global__ void __launch_bounds__(kVBT) k_lanczos(long long n6, const MinresState* st, RVec<S> u,
                                                   RVec<S> r2, double* partials,
                                                   const BatchSlot<S>* slots) {
  __shared__ double red[kVBT / 64];
  if (slots != nullptr) {  // u is the slot's r1 (the matvec's output)
    const BatchSlot<S>& sl = slots[blockIdx.y];
    st = sl.st;
    u = sl.r1;
    r2 = sl.r2;
    partials = sl.partials;
  }
  if (st->done) return;
  const T c = T(st->c_r2);
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)kVBT + threadIdx.x; i < n6;
       i += (long long)gridDim.x * kVBT) {
    const T y = u.template get<T>(i) - c * r2.template get<T>(i);
    u.put(i, y);
    acc += double(y) * double(y);
  }
  const double t = block_sum<kVBT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <typename S>
__global__ void __launch_bounds__(kSBT) k_givens(MinresState* st, const double* partials, int np,
                                                  const BatchSlot<S>* slots) {
  __shared__ double red[kSBT / 64];
  if (slots != nullptr) {
    st = slots[blockIdx.x].st;
    partials = slots[blockIdx.x].partials;
  }
  if (st->done) return;
  const double b2 = sum_partials(partials, np, red);
  if (threadIdx.x != 0) return;
  MinresState z = *st;
  z.itn += 1;
  z.oldb = z.beta;
  z.beta = sqrt(b2);
  z.tnorm2 += z.alfa * z.alfa + z.oldb * z.oldb + z.beta * z.beta;
  if (z.itn == 1 && z.beta / z.beta1 <= 10.0 * kEps) z.istop = -1;
  z.oldeps = z.epsln;
  z.delta = z.cs * z.dbar + z.sn * z.alfa;
  const double gbar = z.sn * z.dbar - z.cs * z.alfa;
  z.epsln = z.sn * z.beta;
  z.dbar = -z.cs * z.beta;
  z.root = sqrt(gbar * gbar + z.dbar * z.dbar);
  const double gamma = fmax(sqrt(gbar * gbar + z.beta * z.beta), kEps);
  z.cs = gbar / gamma;
  z.sn = z.beta / gamma;
  z.phi = z.cs * z.phibar;
  z.phibar = z.sn * z.phibar;
  z.denom = 1.0 / gamma;
  z.gmax = fmax(z.gmax, gamma);
  z.gmin = fmin(z.gmin, gamma);
  const double zz = z.rhs1 / gamma;
  z.rhs1 = z.rhs2 - z.delta * zz;
  z.rhs2 = -z.epsln * zz;
  z.anorm = sqrt(z.tnorm2);
  // v_k = r2_old / beta_k feeds k_update; the next matvec uses 1/beta_{k+1}
  z.s_prev = z.s;
  z.s = 1.0 / z.beta;
  z.c_r1 = z.beta / z.oldb;
  z.pending_xnorm = 1;
  *st = z;
}

template <typename S, typename T>
__global__ void __launch_bounds__(kVBT) k_update(long long n6, const MinresState* st,
                                                  RVec<S> r2old, const work_t<S>* w1,
                                                  const work_t<S>* w2, work_t<S>* wnew,
                                                  work_t<S>* x, double* partials,
                                                  const BatchSlot<S>* slots) {
  using W = work_t<S>;
  if (slots != nullptr) {  // v is the slot's r2 (the Lanczos vector before the swap)
    const BatchSlot<S>& sl = slots[blockIdx.y];
    st = sl.st;
    r2old = sl.r2;
    w1 = sl.w1;
    w2 = sl.w2;
    wnew = sl.wf;
    x = sl.x;
    partials = sl.partials;
  }
  __shared__ double red[kVBT / 64];
  if (st->done) return;
  const T sv = T(st->s_prev), oe = T(st->oldeps), de = T(st->delta), dn = T(st->denom),
          ph = T(st->phi);
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)kVBT + threadIdx.x; i < n6;
       i += (long long)gridDim.x * kVBT) {
    const T v = sv * r2old.template get<T>(i);
    const W w = W((v - oe * T(w1[i]) - de * T(w2[i])) * dn);
    wnew[i] = w;
    const W xn = W(T(x[i]) + ph * T(w));
    x[i] = xn;
    acc += double(xn) * double(xn);
  }
  const double t = block_sum<kVBT>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <typename S>
__global__ void __launch_bounds__(kSBT) k_tests(MinresState* st, const double* partials, int np,
                                                 const BatchSlot<S>* slots) {
  __shared__ double red[kSBT / 64];
  if (slots != nullptr) {
    st = slots[blockIdx.x].st;
    partials = slots[blockIdx.x].partials;
  }
  if (st->done) return;
  const double x2 = sum_partials(partials, np, red);
  if (threadIdx.x != 0) return;
  MinresState z = *st;
  z.pending_xnorm = 0;
  z.ynorm = sqrt(x2);
  const double epsx = z.anorm * z.ynorm * kEps;
  z.rnorm = z.phibar;
  z.test1 = (z.ynorm == 0.0 || z.anorm == 0.0) ? INFINITY : z.rnorm / (z.anorm * z.ynorm);
  z.test2 = z.anorm == 0.0 ? INFINITY : z.root / z.anorm;
  z.acond = z.gmax / z.gmin;
  if (z.istop == 0) {
    if (1.0 + z.test2 <= 1.0) z.istop = 2;
    if (1.0 + z.test1 <= 1.0) z.istop = 1;
    if (z.itn >= z.maxiter) z.istop = 6;
    if (z.acond >= 0.1 / kEps) z.istop = 4;
    if (epsx >= z.beta1) z.istop = 3;
    if (z.test2 <= z.rtol) z.istop = 2;
    if (z.test1 <= z.rtol) z.istop = 1;
  }
  if (z.istop != 0) z.done = 1;
  *st = z;
}

int vec_blocks(long long n6) {
  const long long b = (n6 + kVBT - 1) / kVBT;
  return int(b < kVBlocks ? b : kVBlocks);
}

}  // namespace

template <typename S, typename A>
Minres<S, A>::Minres(Spectral<S, A>& sp) : sp_(sp) {
  n6_ = blk_elems(sp.grid().n, 6);  // 6 components blocked by 64 points
  for (int k = 0; k < 2; ++k) {
    MTL_HIP(hipMalloc(&lanczos_mem_[k], real_bytes<S>(n6_)));
    lanczos_[k] = rvec_over<S>(lanczos_mem_[k], n6_);
  }
  for (auto& v : w_) MTL_HIP(hipMalloc(&v, n6_ * sizeof(W)));
  MTL_HIP(hipMalloc(&x_, n6_ * sizeof(W)));
  partial_cap_ = std::max(kVBlocks, sp.zinv_blocks());
  MTL_HIP(hipMalloc(&partials_, partial_cap_ * sizeof(double)));
  MTL_HIP(hipMalloc(&state_, sizeof(MinresState)));
  MTL_HIP(hipHostMalloc(&host_state_, sizeof(MinresState), hipHostMallocDefault));
}

template <typename S, typename A>
Minres<S, A>::~Minres() {
  for (void* v : lanczos_mem_)
    if (v) (void)hipFree(v);
  for (auto v : w_)
    if (v) (void)hipFree(v);
  if (x_) (void)hipFree(x_);
  if (partials_) (void)hipFree(partials_);
  if (state_) (void)hipFree(state_);
  if (host_state_) (void)hipHostFree(host_state_);
  if (slots_dev_) (void)hipFree(slots_dev_);
  if (slots_host_) (void)hipHostFree(slots_host_);
  drop_graphs();
}

template <typename S, typename A>
void Minres<S, A>::drop_graphs() {
  for (auto& kv : graphs_)
    for (hipGraphExec_t g : kv.second)
      if (g) (void)hipGraphExecDestroy(g);
  graphs_.clear();
}

template <typename S, typename A>
int Minres<S, A>::solve(const double* b, const W* tangent, const Vec6& mask, double rtol,
                     long long maxiter, hipStream_t s, int* istop_out) {
  // roles: r1, r2 (Lanczos), w1, w2 (previous w), wf (free slot for the new w)
  RVec<S> r1 = lanczos_[0];
  RVec<S> r2 = lanczos_[1];
  W* w1 = w_[0];
  W* w2 = w_[1];
  W* wf = w_[2];
  const int vb = vec_blocks(n6_);
  hipLaunchKernelGGL(k_init<S>, dim3(vb), dim3(kVBT), 0, s, n6_, sp_.grid().n, b, r1, r2, x_, w2,
                     wf, partials_, static_cast<const BatchSlot<S>*>(nullptr));
  MTL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_start<S>, dim3(1), dim3(kSBT), 0, s, state_, partials_, vb, rtol, maxiter,
                     static_cast<const BatchSlot<S>*>(nullptr));
  MTL_LAUNCH_CHECK();
  // at itn 1 the previous w's are zero: w1 := wf (zeroed), w2 (zeroed)
  std::swap(w1, wf);

  int chunk = 8;
  int prev_itn = 0;
  double prev_test1 = 0.0;
  long long launched = 0;
  while (true) {
    for (int it = 0; it < chunk; ++it) {
      const int np = sp_.matvec_minres(r2, tangent, r1, mask, state_, partials_, s);
      profiler().launch(kKMinresScalar, k_alpha<S>, dim3(1), dim3(kSBT), 0, s, state_, partials_, np,
                        static_cast<const BatchSlot<S>*>(nullptr));
      MTL_LAUNCH_CHECK();
      profiler().launch(kKMinresVector, k_lanczos<S, A>, dim3(vb), dim3(kVBT), 0, s, n6_, state_, r1, r2, partials_,
                        static_cast<const BatchSlot<S>*>(nullptr));
      MTL_LAUNCH_CHECK();
      profiler().launch(kKMinresScalar, k_givens<S>, dim3(1), dim3(kSBT), 0, s, state_, partials_, vb,
                        static_cast<const BatchSlot<S>*>(nullptr));
      MTL_LAUNCH_CHECK();
      // Lanczos roles rotate: r1 <- r2 (old), r2 <- y (stored in old r1)
      std::swap(r1, r2);
      // v_k = r2_old / beta_k now lives in r1
      profiler().launch(kKMinresVector, k_update<S, A>, dim3(vb), dim3(kVBT), 0, s, n6_, state_, r1, w1, w2, wf, x_, partials_,
                        static_cast<const BatchSlot<S>*>(nullptr));
      MTL_LAUNCH_CHECK();
      profiler().launch(kKMinresScalar, k_tests<S>, dim3(1), dim3(kSBT), 0, s, state_, partials_, vb,
                        static_cast<const BatchSlot<S>*>(nullptr));
      MTL_LAUNCH_CHECK();
      // w roles rotate: w1 <- w2, w2 <- new w, free slot <- old w1
      W* old_w1 = w1;
      w1 = w2;
      w2 = wf;
      wf = old_w1;
      ++launched;
    }
    MTL_HIP(hipMemcpyAsync(host_state_, state_, sizeof(MinresState), hipMemcpyDeviceToHost, s));
    stream_sync(s);
    if (std::getenv("MTL_DEBUG") != nullptr) {
      const MinresState& z = *host_state_;
      std::fprintf(stderr, "[mtl]     minres itn %d beta %.4e alfa %.4e phibar %.4e test1 %.3e test2 %.3e\n",
                   z.itn, z.beta, z.alfa, z.phibar, z.test1, z.test2);
    }
    if (host_state_->done || launched >= maxiter) break;
    // size the next chunk from the observed geometric decay of SciPy's test1
    // so that few iterations are enqueued past convergence
    const MinresState& z = *host_state_;
    chunk = 8;
    if (prev_itn > 0 && z.itn > prev_itn && z.test1 > 0.0 && prev_test1 > z.test1) {
      const double rate = std::pow(z.test1 / prev_test1, 1.0 / double(z.itn - prev_itn));
      const double need = std::log(z.rtol / z.test1) / std::log(rate);
      if (std::isfinite(need)) chunk = int(std::min(64.0, std::max(2.0, std::ceil(need) + 1.0)));
    }
    prev_itn = z.itn;
    prev_test1 = z.test1;
  }
  if (istop_out) *istop_out = host_state_->istop;
  return host_state_->itn;
}

template <typename S, typename A>
void Minres<S, A>::solve_batch(const std::vector<BatchItem>& items, long long maxiter,
                               hipStream_t s, int* itns) {
  const int nb = int(items.size());
  if (nb == 0) return;
  Minres& lead = *items[0].m;
  for (const auto& it : items)
    if (it.m->n6_ != lead.n6_) throw Error("a MINRES batch needs solvers of one grid");
  // slot tables of the six role phases (Lanczos roles alternate, w roles
  // rotate with period 3), kept by the first solver of the batch
  if (lead.slot_cap_ < nb) {
    lead.drop_graphs();
    if (lead.slots_dev_) MTL_HIP(hipFree(lead.slots_dev_));
    if (lead.slots_host_) MTL_HIP(hipHostFree(lead.slots_host_));
    lead.slots_dev_ = nullptr;
    lead.slots_host_ = nullptr;
    MTL_HIP(hipMalloc(&lead.slots_dev_, 6 * size_t(nb) * sizeof(BatchSlot<S>)));
    MTL_HIP(hipHostMalloc(&lead.slots_host_, 6 * size_t(nb) * sizeof(BatchSlot<S>),
                          hipHostMallocDefault));
    lead.slot_cap_ = nb;
  }
  for (int ph = 0; ph < 6; ++ph)
    for (int i = 0; i < nb; ++i) {
      const BatchItem& it = items[i];
      Minres& m = *it.m;
      BatchSlot<S>& sl = lead.slots_host_[ph * nb + i];
      sl.spec = m.sp_.spectrum();
      sl.tangent = it.tangent;
      sl.r1 = m.lanczos_[ph % 2];
      sl.r2 = m.lanczos_[(ph + 1) % 2];
      sl.maskbits = Spectral<S, A>::mask_bits(it.mask);
      sl.st = m.state_;
      sl.partials = m.partials_;
      sl.w1 = m.w_[(2 - ph % 3 + 3) % 3];
      sl.w2 = m.w_[(1 - ph % 3 + 3) % 3];
      sl.wf = m.w_[(3 - ph % 3) % 3];
      sl.x = m.x_;
      sl.b = it.b;
      sl.rtol = it.rtol;
    }
  MTL_HIP(hipMemcpyAsync(lead.slots_dev_, lead.slots_host_, 6 * size_t(nb) * sizeof(BatchSlot<S>),
                         hipMemcpyHostToDevice, s));
  const BatchSlot<S>* tab = lead.slots_dev_;
  const int vb = vec_blocks(lead.n6_);
  hipLaunchKernelGGL(k_init<S>, dim3(vb, nb), dim3(kVBT), 0, s, lead.n6_, lead.sp_.grid().n,
                     static_cast<const double*>(nullptr), RVec<S>{}, RVec<S>{},
                     static_cast<work_t<S>*>(nullptr), static_cast<work_t<S>*>(nullptr),
                     static_cast<work_t<S>*>(nullptr), static_cast<double*>(nullptr), tab);
  MTL_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_start<S>, dim3(nb), dim3(kSBT), 0, s, static_cast<MinresState*>(nullptr),
                     static_cast<const double*>(nullptr), vb, 0.0, maxiter, tab);
  MTL_LAUNCH_CHECK();
  // one iteration whose roles are those of phase ph
  auto iteration = [&](int ph) {
    const BatchSlot<S>* t = tab + ph * nb;
    const int np = lead.sp_.matvec_batch(t, nb, s);
    profiler().launch(kKMinresScalar, k_alpha<S>, dim3(nb), dim3(kSBT), 0, s,
                      static_cast<MinresState*>(nullptr), static_cast<const double*>(nullptr), np, t);
    MTL_LAUNCH_CHECK();
    profiler().launch(kKMinresVector, k_lanczos<S, A>, dim3(vb, nb), dim3(kVBT), 0, s, lead.n6_,
                      static_cast<const MinresState*>(nullptr), RVec<S>{}, RVec<S>{},
                      static_cast<double*>(nullptr), t);
    MTL_LAUNCH_CHECK();
    profiler().launch(kKMinresScalar, k_givens<S>, dim3(nb), dim3(kSBT), 0, s,
                      static_cast<MinresState*>(nullptr), static_cast<const double*>(nullptr), vb, t);
    MTL_LAUNCH_CHECK();
    profiler().launch(kKMinresVector, k_update<S, A>, dim3(vb, nb), dim3(kVBT), 0, s, lead.n6_,
                      static_cast<const MinresState*>(nullptr), RVec<S>{},
                      static_cast<const work_t<S>*>(nullptr), static_cast<const work_t<S>*>(nullptr),
                      static_cast<work_t<S>*>(nullptr), static_cast<work_t<S>*>(nullptr),
                      static_cast<double*>(nullptr), t);
    MTL_LAUNCH_CHECK();
    profiler().launch(kKMinresScalar, k_tests<S>, dim3(nb), dim3(kSBT), 0, s,
                      static_cast<MinresState*>(nullptr), static_cast<const double*>(nullptr), vb, t);
    MTL_LAUNCH_CHECK();
  };
  // Replaying recorded one-iteration graphs instead of launching kernels
  // gives the same results; on MI355X (ROCm 7) it measured 5-10 % slower
  // than direct launches for 32^3-64^3 batches, so it is opt-in.
  const char* env = std::getenv("MTL_BATCH_GRAPHS");
  const bool graphs = !profiler().enabled() && env != nullptr && env[0] == '1';
  std::array<hipGraphExec_t, 6>* cycle = nullptr;
  if (graphs) {
    auto found = lead.graphs_.find(nb);
    if (found == lead.graphs_.end()) {
      std::array<hipGraphExec_t, 6> made{};
      try {
        for (int p = 0; p < 6; ++p) {
          GraphRecorder rec;
          iteration(p);
          made[p] = rec.instantiate();
        }
      } catch (...) {
        for (hipGraphExec_t g : made)
          if (g) (void)hipGraphExecDestroy(g);
        throw;
      }
      found = lead.graphs_.emplace(nb, made).first;
    }
    cycle = &found->second;
  }
  int chunk = 8;
  long long launched = 0;
  std::vector<int> prev_itn(nb, 0);
  std::vector<double> prev_test1(nb, 0.0);
  while (true) {
    if (graphs) {
      for (int k = 0; k < chunk; ++k) MTL_HIP(hipGraphLaunch((*cycle)[launched++ % 6], s));
    } else {
      for (int k = 0; k < chunk; ++k) iteration(int(launched++ % 6));
    }
    for (const auto& it : items)
      MTL_HIP(hipMemcpyAsync(it.m->host_state_, it.m->state_, sizeof(MinresState),
                             hipMemcpyDeviceToHost, s));
    stream_sync(s);
    bool all_done = true;
    chunk = 2;
    for (int i = 0; i < nb; ++i) {
      const MinresState& z = *items[i].m->host_state_;
      if (z.done) continue;
      all_done = false;
      // the slowest system sizes the next chunk (SciPy test1's observed decay)
      int need = 8;
      if (prev_itn[i] > 0 && z.itn > prev_itn[i] && z.test1 > 0.0 && prev_test1[i] > z.test1) {
        const double rate = std::pow(z.test1 / prev_test1[i], 1.0 / double(z.itn - prev_itn[i]));
        const double est = std::log(z.rtol / z.test1) / std::log(rate);
        if (std::isfinite(est)) need = int(std::min(64.0, std::max(2.0, std::ceil(est) + 1.0)));
      }
      chunk = std::max(chunk, need);
      prev_itn[i] = z.itn;
      prev_test1[i] = z.test1;
    }
    if (all_done || launched >= maxiter) break;
  }
  for (int i = 0; i < nb; ++i) itns[i] = items[i].m->host_state_->itn;
}

template class Minres<float>;
template class Minres<double>;
template class Minres<f48, double>;

}  // namespace mtl
