// Elastic and elasto-viscoplastic (crystal plasticity) updates on gfx950.
//
// Elastic (small_strain_fft/elastic.py:23-31): sigma = Rm^T C Rm : eps.
//
// Elasto-viscoplastic (small_strain_fft/elastic_viscoplastic.py:91-271), per
// point in the crystal frame, one implicit Newton step per iteration:
//   tau_s = P_s : sig,  pref_s = g0dot |tau_s/g_s|^(n-1) / g_s,  rate_s = pref_s tau_s
//   ep    = ep_old + dt sum_s rate_s P_s
//   J     = I + dt C M,  M = sum_s n pref_s P_s (x) P_s
//   sig  <- sig - J^-1 (sig - C (eps - ep)),   g <- hardening(...)
// Since C is SPD and M is PSD, J^-1 = S^-1 C^-1 with S = C^-1 + dt M SPD, so the
// update is computed as sig - S^-1 (C^-1 sig - (eps - ep)) with a Cholesky
// inverse of S, and the consistent tangent J^-1 C equals S^-1 (symmetric).
//
// Iteration order follows the reference exactly: every point iterates until at
// least half of the model's points satisfy the convergence test (iteration
// K*), after which only the unconverged points iterate, each independently, up
// to 60 iterations in total.  Each iteration is one kernel over the state in
// HBM (ping-pong buffers and a device converged-count).  The iterate's slip
// resistances are not stored: every hardening law gives g_s as a function of
// the committed g_old_s, the accumulated slip and the scalar slip increment
// d = dt sum_s |rate_s| of the previous iteration, so one double per point
// (d) replaces NS read and NS written values per iteration; recomputing g_s
// from it repeats the same operations and is bitwise identical.
//
// Slip resistances are stored as one value g per point.  With one initial
// slip resistance for every system (elastic_viscoplastic.py:66-69 adds a
// scalar to every system) each hardening law (hardening_laws.py:18-90)
// applies the same operations to every system of a point (the increment
// depends on the point's total slip only, Armstrong-Frederick's also on
// g_old_s itself), so a point's NS values stay bitwise equal for the whole
// run.  With one initial value per system, g is that of a reference system r
// and the others follow from it exactly: linear, Voce and perfect plasticity
// add the same increment to every system, so g_s - g0_s = g - g0_r; the
// Armstrong-Frederick step g' = (g + h d) / (1 + r d) scales every
// (h/r - g_s) by the same factor, so g_s - g0_s = c_s (g - g0_r) with
// c_s = (h/r - g0_s) / (h/r - g0_r) (r = 0 is the linear case).  Either way
// g_s = g0_s + c_s (g - g0_r), which returns every g0_s exactly at the start.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "constitutive.hpp"
#include "profiler.hpp"

namespace mtl {

namespace {

constexpr int kCBT = 256;

// Mandel form of a rotation R (v_crystal = R v_specimen):
// (R A R^T) in Mandel = Rm A_mandel  (tensor.py:1668-1700).
__device__ __forceinline__ void mandel_rotation(const double* R, double Rm[36]) {
  const double r2 = 1.41421356237309504880168872420969807857;
#define R_(i, j) R[(i)*3 + (j)]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Rm[i * 6 + j] = R_(i, j) * R_(i, j);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Rm[i * 6 + 3] = r2 * R_(i, 1) * R_(i, 2);
    Rm[i * 6 + 4] = r2 * R_(i, 0) * R_(i, 2);
    Rm[i * 6 + 5] = r2 * R_(i, 0) * R_(i, 1);
  }
  // rows 3,4,5 pair the (1,2), (0,2), (0,1) rows of R
  const int pa[3] = {1, 0, 0}, pb[3] = {2, 2, 1};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int a = pa[m], b = pb[m];
#pragma unroll
    for (int j = 0; j < 3; ++j) Rm[(3 + m) * 6 + j] = r2 * R_(a, j) * R_(b, j);
    Rm[(3 + m) * 6 + 3] = R_(a, 1) * R_(b, 2) + R_(a, 2) * R_(b, 1);
    Rm[(3 + m) * 6 + 4] = R_(a, 0) * R_(b, 2) + R_(a, 2) * R_(b, 0);
    Rm[(3 + m) * 6 + 5] = R_(a, 0) * R_(b, 1) + R_(a, 1) * R_(b, 0);
  }
#undef R_
}

__device__ __forceinline__ void load_rotation(const double* rot9, long long npts, long long p,
                                              double Rm[36]) {
  double R[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = rot9[k * npts + p];
  mandel_rotation(R, Rm);
}

// v_c = Rm v_s
__device__ __forceinline__ void to_crystal(const double Rm[36], const double vs[6], double vc[6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) acc += Rm[a * 6 + b] * vs[b];
    vc[a] = acc;
  }
}

// v_s = Rm^T v_c
__device__ __forceinline__ void to_specimen(const double Rm[36], const double vc[6], double vs[6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) acc += Rm[b * 6 + a] * vc[b];
    vs[a] = acc;
  }
}

// Ds = Rm^T Dc Rm for a packed symmetric Dc  (tensor.py:1620-1634)
__device__ __forceinline__ void tangent_to_specimen(const double Rm[36], const double Dc[21],
                                                    double Ds[21]) {
  double T[36];  // Dc Rm
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int b = 0; b < 6; ++b) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += Dc[sym6(i, j)] * Rm[j * 6 + b];
      T[i * 6 + b] = acc;
    }
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) acc += Rm[i * 6 + a] * T[i * 6 + b];
      Ds[sym6(a, b)] = acc;
    }
}

template <typename K>
__device__ __forceinline__ void store_tangent(K* tangent, long long npts, long long p,
                                              const double Ds[21]) {
#pragma unroll
  for (int k = 0; k < 21; ++k) tangent[tan_idx(k, p)] = K(Ds[k]);
}

// tangent <- Rm^T Dc Rm as tangent_to_specimen + store_tangent (the same sums in
// the same order), one column of Dc Rm at a time, each packed entry stored as
// soon as it is formed: 6 intermediate values live instead of 36 + 21
template <typename K>
__device__ __forceinline__ void store_tangent_to_specimen(const double Rm[36], const double Dc[21],
                                                          K* tangent, long long npts,
                                                          long long p) {
#pragma unroll
  for (int b = 0; b < 6; ++b) {
    double Tb[6];  // column b of Dc Rm
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += Dc[sym6(i, j)] * Rm[j * 6 + b];
      Tb[i] = acc;
    }
#pragma unroll
    for (int a = 0; a <= b; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < 6; ++i) acc += Rm[i * 6 + a] * Tb[i];
      tangent[tan_idx(sym6(a, b), p)] = K(acc);
    }
  }
}

struct Stiff21 {
  double c[21];
};

// ------------------------------------------------------------------ elastic
template <typename K>
__global__ void __launch_bounds__(kCBT) k_elastic(int n, const int* idx, long long npts,
                                                   const double* rot9, Stiff21 C,
                                                   const double* eps, double* sig, K* tangent,
                                                   int write_stress) {
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  const long long p = idx[i];
  double Rm[36], Ds[21];
  load_rotation(rot9, npts, p, Rm);
  tangent_to_specimen(Rm, C.c, Ds);
  store_tangent(tangent, npts, p, Ds);
  if (write_stress) {
    double e[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) e[c] = eps[c * npts + p];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      double acc = 0.0;
#pragma unroll
      for (int b = 0; b < 6; ++b) acc += Ds[sym6(a, b)] * e[b];
      sig[a * npts + p] = acc;
    }
  }
}

// ------------------------------------------------------------------ EVP
template <int NS>
struct EvpParams {
  double P[NS][6];
  double PP[NS][21];  // P_s (x) P_s packed (the reference's schmid_outer_schmid)
  double Cinv[21];
  double gamma0, n, dt, ref_modulus, tol;
  int nint;  // n-1 when it is an integer in [0, 64], else -1
  int hard;
  double hp[4];
  int persys;  // slip resistances differ by system: g_s = g0_s + c_s (g - gref)
  int pad;
  double gref, g0s[NS], cs[NS];
};

// The model parameters (Schmid tensors, compliance: ~100 doubles) are staged in
// LDS rather than passed by value: as kernel arguments they are held in SGPRs
// and spill heavily, which produced wrong results in this kernel.
template <typename P>
__device__ __forceinline__ void load_params(P* dst, const P* src) {
  static_assert(sizeof(P) % sizeof(int) == 0, "parameter block must be int-sized");
  const int* s = reinterpret_cast<const int*>(src);
  int* d = reinterpret_cast<int*>(dst);
  for (int k = threadIdx.x; k < int(sizeof(P) / sizeof(int)); k += blockDim.x) d[k] = s[k];
}

struct EvpControl {
  int count[kEvpMaxIterations + 4];  // points converging in iteration it
  long long total;                   // points converged after the last iteration
  int masked;                        // masked phase active
  int kstar;                         // iteration that started it
  int done;
  int fail;
  int iterations;
  int pad;
};

// |x|^ef; e = ef when ef is an integer in [0, 64] (INTPOW), then by squaring.
// Separate instantiations keep the registers of the general pow() out of the
// integer-exponent kernels.
template <bool INTPOW>
__device__ __forceinline__ double powi_abs(double x, int e, double ef) {
  x = fabs(x);
  if (!INTPOW) return pow(x, ef);
  double r = 1.0, b = x;
  while (e) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

// In-place Cholesky factor of a packed SPD 6x6: on return S holds the lower
// factor, L[i][j] (i > j) at sym6(i, j), with 1 / L[j][j] on the diagonal.
__device__ __forceinline__ void spd_factor6(double S[21]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double dg = S[sym6(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) dg -= S[sym6(j, k)] * S[sym6(j, k)];
    const double il = 1.0 / sqrt(dg);
    S[sym6(j, j)] = il;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double v = S[sym6(i, j)];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= S[sym6(i, k)] * S[sym6(j, k)];
      S[sym6(i, j)] = v * il;
    }
  }
}

// x = (L L^T)^-1 r for a factor from spd_factor6
__device__ __forceinline__ void spd_solve6(const double L[21], const double r[6], double x[6]) {
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double v = r[i];
#pragma unroll
    for (int k = 0; k < i; ++k) v -= L[sym6(i, k)] * y[k];
    y[i] = v * L[sym6(i, i)];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double v = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) v -= L[sym6(k, i)] * x[k];
    x[i] = v * L[sym6(i, i)];
  }
}

// (L L^T)^-1, packed, for a factor from spd_factor6
__device__ __forceinline__ void spd_inverse6(const double L[21], double Sinv[21]) {
  double Li[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    Li[i][i] = L[sym6(i, i)];
#pragma unroll
    for (int j = 0; j < i; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = j; k < i; ++k) acc += L[sym6(i, k)] * Li[k][j];
      Li[i][j] = -acc * Li[i][i];
    }
  }
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = a; b < 6; ++b) {
      double acc = 0.0;
#pragma unroll
      for (int k = b; k < 6; ++k) acc += Li[k][a] * Li[k][b];
      Sinv[sym6(a, b)] = acc;
    }
}

// One iteration of the per-point implicit update (elastic_viscoplastic.py:190-253),
// accumulated one slip system at a time (gs: the point's slip resistance).
// The per-system slip increments (when asked for) go to LDS columns (system s
// at dgam[s * kCBT]) rather than registers, so the loop over systems need not
// be unrolled and the kernel keeps enough waves resident to overlap its HBM
// traffic with the FP64 arithmetic.  On return S holds the
// Cholesky factor of C^-1 + dt M, sig_new the updated stress, d the slip
// increment dt sum_s |rate_s|, dgam (when non-null) the slip increments
// rate_s dt, and ep (when ep_old_p is non-null) the plastic strain.  e0 is the
// crystal strain minus the committed plastic strain (fixed during a call, set
// up once per call), read (components strided by ln) only after the loop over
// systems, which keeps it out of the registers live during it.
template <int NS, bool IP>
__device__ __forceinline__ void evp_iterate(const EvpParams<NS>& q, const double sig[6],
                                            const double gs, const double* e0_p,
                                            const double* ep_old_p, long long ln,
                                            double sig_new[6], double ep[6], double S[21],
                                            double* dgam, double& diff, double& d) {
  double epr[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) epr[m] = 0.0;
#pragma unroll
  for (int k = 0; k < 21; ++k) S[k] = 0.0;
  double dsum = 0.0;
  // one reciprocal for all systems when the slip resistance is the same for
  // every system of a point, instead of two FP64 divisions per system; the
  // rate differs from the divided form by a few ulps
  const double ginv = 1.0 / gs;
#pragma unroll 1
  for (int s = 0; s < NS; ++s) {
    double tau = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) tau += q.P[s][m] * sig[m];
    const double gi = q.persys ? 1.0 / (q.g0s[s] + q.cs[s] * (gs - q.gref)) : ginv;
    const double pref = q.gamma0 * powi_abs<IP>(tau * gi, q.nint, q.n - 1.0) * gi;
    const double rate = pref * tau;
    const double w = q.n * pref;
#pragma unroll
    for (int m = 0; m < 6; ++m) epr[m] += rate * q.P[s][m];
#pragma unroll
    for (int k = 0; k < 21; ++k) S[k] += w * q.PP[s][k];
    if (dgam != nullptr) dgam[s * kCBT] = rate * q.dt;
    dsum += fabs(rate);
  }
  if (ep_old_p != nullptr) {
#pragma unroll
    for (int m = 0; m < 6; ++m) ep[m] = ep_old_p[m * ln] + epr[m] * q.dt;
  }
#pragma unroll
  for (int k = 0; k < 21; ++k) S[k] = q.Cinv[k] + q.dt * S[k];
  double rc[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    double acc = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) acc += q.Cinv[sym6(a, b)] * sig[b];
    rc[a] = acc - (e0_p[a * ln] - epr[a] * q.dt);
  }
  spd_factor6(S);
  double x[6];
  spd_solve6(S, rc, x);
  double d2 = 0.0;
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    sig_new[a] = sig[a] - x[a];
    d2 += x[a] * x[a];
  }
  diff = sqrt(d2);
  d = q.dt * dsum;
}

// Slip resistance (every system of the point) after a slip increment d from
// the committed state (hardening_laws.py:18-90).
template <int NS>
__device__ __forceinline__ double evp_harden(const EvpParams<NS>& q, double g_old, double acc_old,
                                             double d) {
  if (q.hard == kLinear) return g_old + q.hp[0] * d;
  if (q.hard == kVoce) {
    const double theta0 = q.hp[0], theta1 = q.hp[1], tau1 = q.hp[2];
    const double ratio = fabs(theta0 / tau1);
    const double e0 = exp(-ratio * acc_old);
    const double ed = exp(-ratio * d);
    const double inc = theta1 * d - (ratio * tau1 - theta1) / ratio * e0 * (ed - 1.0) -
                       theta1 / ratio * e0 * (ed * (1.0 + ratio * (acc_old + d)) - 1.0 - ratio * acc_old);
    return g_old + inc;
  }
  if (q.hard == kArmstrongFrederick) {
    const double h = q.hp[0], r = q.hp[1];
    return g_old + (h - r * g_old) * d / (1.0 + r * d);
  }
  return g_old;
}

template <int NS>
__device__ __forceinline__ bool evp_converged(const EvpParams<NS>& q, double diff, double gnorm) {
  // elastic_viscoplastic.py:160-164 (a 0/0 or x/0 ratio compares false)
  return (diff / gnorm <= q.tol) || (diff / q.ref_modulus < 1.0e-12);
}

struct EvpBuffers {
  // 6- and NS-component arrays are blocked by 64 points (blk_idx)
  double* e0;        // [6] crystal strain minus committed plastic strain
  double* gnorm;     // [n]
  double* sig[2];    // [6] ping-pong
  double* dsl[2];    // [n] ping-pong slip increment d of the iteration
  const double* g_trial;  // [n] slip resistance of the first iteration
  unsigned char* conv;
  int* last;         // [n] last iteration each point ran
  double* ep_old;    // [6]
  double* g_old;     // [n] (every system of a point has the same value)
  double* gam_old;   // [NS]
  double* acc_old;   // [n]
  double* ep_pend;   // outputs of the call, adopted as trial state on success
  double* g_pend;
  double* gam_pend;
  double* acc_pend;
};

template <int NS>
__global__ void __launch_bounds__(kCBT) k_evp_setup(int n, const int* idx, long long npts,
                                                     const double* rot9, const double* eps,
                                                     const double* guess, EvpBuffers B) {
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  const long long p = idx[i];
  double Rm[36], v[6], vc[6];
  load_rotation(rot9, npts, p, Rm);
#pragma unroll
  for (int c = 0; c < 6; ++c) v[c] = eps[c * npts + p];
  to_crystal(Rm, v, vc);
#pragma unroll
  for (int c = 0; c < 6; ++c) B.e0[blk_idx(c, i, 6)] = vc[c] - B.ep_old[blk_idx(c, i, 6)];
  double nrm = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    v[c] = guess[c * npts + p];
    nrm += v[c] * v[c];
  }
  B.gnorm[i] = sqrt(nrm);
  to_crystal(Rm, v, vc);
#pragma unroll
  for (int c = 0; c < 6; ++c) B.sig[0][blk_idx(c, i, 6)] = vc[c];
}

__global__ void k_evp_ctrl_reset(EvpControl* c) {
  int* w = reinterpret_cast<int*>(c);
  for (int i = threadIdx.x; i < int(sizeof(EvpControl) / sizeof(int)); i += blockDim.x) w[i] = 0;
}

// State entering iteration it (1-based): the stress written by iteration it-1
// (the initial guess for it = 1) and the slip resistance that iteration
// produced (the trial value for it = 1).
template <int NS>
__device__ __forceinline__ void load_point(const EvpParams<NS>& q, const EvpBuffers& B, int n,
                                           int i, int it, double sig[6], double& g) {
  const int buf = (it - 1) & 1;
#pragma unroll
  for (int c = 0; c < 6; ++c) sig[c] = B.sig[buf][blk_idx(c, i, 6)];
  g = it == 1 ? B.g_trial[i] : evp_harden<NS>(q, B.g_old[i], B.acc_old[i], B.dsl[buf][i]);
}

// Iteration `it` (1-based) of every point that still iterates: all points
// before the masked phase, only unconverged ones after it
// (elastic_viscoplastic.py:116-158).  A point's state after iteration j lives
// in ping-pong buffer j&1; last[i] records the last iteration it ran.
// Workgroups per CU the iteration kernel is compiled for: four (128 VGPRs) with
// the integer power, which fits without spilling for every instantiated system
// count except 9; three otherwise.
constexpr int evp_iter_min_blocks(int ns, bool ip) { return ip && ns != 9 ? 4 : 3; }

template <int NS, bool IP>
__global__ void __launch_bounds__(kCBT, evp_iter_min_blocks(NS, IP)) k_evp_iter(int it, int n, const EvpParams<NS>* qg,
                                                    EvpBuffers B, EvpControl* ctrl) {
  if (ctrl->done) return;
  const int masked = ctrl->masked;
  __shared__ EvpParams<NS> q;
  __shared__ int cnt;
  const int i = blockIdx.x * kCBT + threadIdx.x;
  const bool active = i < n && !(masked && B.conv[i]);
  // in the masked phase most workgroups have no point left: they skip the
  // parameter load (a few KB per workgroup) and return
  if (!__syncthreads_or(active)) return;
  load_params(&q, qg);
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if (active) {
    const int dst = it & 1;
    double sig[6], g;
    load_point<NS>(q, B, n, i, it, sig, g);
    double sig_new[6], ep[6], S[21], diff, d;
    evp_iterate<NS, IP>(q, sig, g, B.e0 + blk_idx(0, i, 6), nullptr, 64, sig_new, ep, S, nullptr,
                        diff, d);
#pragma unroll
    for (int c = 0; c < 6; ++c) B.sig[dst][blk_idx(c, i, 6)] = sig_new[c];
    B.dsl[dst][i] = d;
    const bool cv = evp_converged<NS>(q, diff, B.gnorm[i]);
    B.conv[i] = cv ? 1 : 0;
    B.last[i] = it;
    if (cv) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&ctrl->count[it], cnt);
}

// Convergence bookkeeping after iteration `it` (elastic_viscoplastic.py:160-170).
__global__ void k_evp_control(int it, int n, EvpControl* ctrl) {
  if (threadIdx.x != 0 || ctrl->done) return;
  // unmasked: every point was tested; masked: converged points kept their flag
  const long long total = ctrl->masked ? ctrl->total + ctrl->count[it] : ctrl->count[it];
  ctrl->total = total;
  if (total == n) {
    ctrl->done = 1;
    ctrl->iterations = it;
    return;
  }
  if (it >= kEvpMaxIterations) {
    ctrl->fail = 1;
    ctrl->done = 1;
    return;
  }
  if (!ctrl->masked && 2 * total >= n) {  // converged fraction >= 0.5
    ctrl->masked = 1;
    ctrl->kstar = it;
  }
}

// Outputs of each point's last iteration, recomputed from the state it started
// from (bitwise the same stress and slip resistances as that iteration wrote).
template <int NS, typename K, bool IP>
__global__ void __launch_bounds__(kCBT) k_evp_outputs(int n, const int* idx, long long npts,
                                                       const double* rot9,
                                                       const EvpParams<NS>* qg, EvpBuffers B,
                                                       const EvpControl* ctrl, double* sig_out,
                                                       K* tangent) {
  if (ctrl->fail || !ctrl->done) return;
  __shared__ EvpParams<NS> q;
  __shared__ double dgsh[NS * kCBT];
  load_params(&q, qg);
  __syncthreads();
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  double sig[6];
  double sig_new[6], ep[6], S[21], Sinv[21], diff, d;
  double g;
  double* dgam = dgsh + threadIdx.x;
  load_point<NS>(q, B, n, i, B.last[i], sig, g);
  evp_iterate<NS, IP>(q, sig, g, B.e0 + blk_idx(0, i, 6), B.ep_old + blk_idx(0, i, 6), 64,
                      sig_new, ep, S, dgam, diff, d);
  // trial state first (frees its registers before the rotations)
  const double acc_old = B.acc_old[i];
  B.g_pend[i] = evp_harden<NS>(q, B.g_old[i], acc_old, d);
#pragma unroll
  for (int c = 0; c < 6; ++c) B.ep_pend[blk_idx(c, i, 6)] = ep[c];
  double asum = 0.0;
#pragma unroll 1
  for (int s = 0; s < NS; ++s) {
    const double dg = dgam[s * kCBT];
    B.gam_pend[blk_idx(s, i, NS)] = B.gam_old[blk_idx(s, i, NS)] + dg;
    asum += fabs(dg);
  }
  B.acc_pend[i] = acc_old + asum;
  spd_inverse6(S, Sinv);
  const long long p = idx[i];
  double Rm[36], vs[6];
  load_rotation(rot9, npts, p, Rm);
  to_specimen(Rm, sig_new, vs);
#pragma unroll
  for (int c = 0; c < 6; ++c) sig_out[c * npts + p] = vs[c];
  store_tangent_to_specimen(Rm, Sinv, tangent, npts, p);
}


// crystal plastic strain -> specimen frame, AoS output for the host
__global__ void __launch_bounds__(kCBT) k_ep_specimen(int n, const int* idx, long long npts,
                                                       const double* rot9, const double* ep,
                                                       double* out) {
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  double Rm[36], vc[6], vs[6];
  load_rotation(rot9, npts, idx[i], Rm);
#pragma unroll
  for (int c = 0; c < 6; ++c) vc[c] = ep[blk_idx(c, i, 6)];
  to_specimen(Rm, vc, vs);
#pragma unroll
  for (int c = 0; c < 6; ++c) out[i * 6LL + c] = vs[c];
}

__global__ void k_fill(double* a, long long count, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

void fill(double* a, long long count, double v, hipStream_t s) {
  const int blocks = int(std::min<long long>((count + 255) / 256, 4096));
  hipLaunchKernelGGL(k_fill, dim3(std::max(blocks, 1)), dim3(256), 0, s, a, count, v);
  MTL_LAUNCH_CHECK();
}

Stiff21 pack21(const double* C36) {
  Stiff21 s{};
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) s.c[sym6(i, j)] = 0.5 * (C36[i * 6 + j] + C36[j * 6 + i]);
  return s;
}

// 6x6 inverse with partial pivoting (host)
bool inverse6(const double* A, double* out) {
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

int* upload_idx(const std::vector<int>& v) {
  int* d = nullptr;
  MTL_HIP(hipMalloc(&d, std::max<size_t>(v.size(), 1) * sizeof(int)));
  if (!v.empty()) MTL_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  return d;
}

// ------------------------------------------------------------------ models
template <typename K>
class ElasticModel final : public Model<K> {
 public:
  ElasticModel(const ModelDesc& d, const std::vector<int>& pts, long long npts) {
    this->n_ = int(pts.size());
    this->npts_ = npts;
    this->host_idx_ = pts;
    this->idx_ = upload_idx(pts);
    C_ = pack21(d.C);
  }
  ~ElasticModel() override { (void)hipFree(this->idx_); }
  void init_tangent(const double* rot9, K* tangent, hipStream_t s) override {
    launch(rot9, nullptr, nullptr, tangent, 0, s);
  }
  int calculate(const double* rot9, const double* eps, const double*, double, double* sig,
                K* tangent, hipStream_t s) override {
    launch(rot9, eps, sig, tangent, 1, s);
    return 1;
  }
  void commit(const double*, hipStream_t) override {}
  bool field(int, const double*, std::vector<double>&, int&, hipStream_t) override { return false; }

 private:
  void launch(const double* rot9, const double* eps, double* sig, K* tangent, int ws,
              hipStream_t s) {
    if (this->n_ == 0) return;
    profiler().launch(kKElastic, k_elastic<K>, dim3(ceil_div(this->n_, kCBT)), dim3(kCBT), 0, s, this->n_, this->idx_, this->npts_, rot9, C_, eps, sig, tangent, ws);
    MTL_LAUNCH_CHECK();
  }
  Stiff21 C_;
};

// Three buffer sets for a model's internal state: the committed set (old),
// the trial set (the last successful call's outputs, tr) and a free set the
// next call writes (pend).  Adopting a call's outputs and committing relabel
// sets instead of copying them; after a commit old and tr name the same set.
struct SlotRing {
  int old = 0, tr = 0, pend = 1;
  void adopt() {  // the call's outputs become the trial state
    tr = pend;
    pend = 3 - old - tr;  // the set that is neither committed nor trial
  }
  void commit() { old = tr; }
};

template <typename K, int NS>
class EvpModel final : public Model<K> {
 public:
  EvpModel(const ModelDesc& d, const std::vector<int>& pts, long long npts) : d_(d) {
    this->n_ = int(pts.size());
    this->npts_ = npts;
    this->host_idx_ = pts;
    this->idx_ = upload_idx(pts);
    C_ = pack21(d.C);
    double Cinv[36];
    if (!inverse6(d.C, Cinv)) throw Error("elasto-viscoplastic stiffness must be invertible");
    // reference modulus E[100] = 1 / (e1 e1 : C^-1 : e1 e1)  (elastic_viscoplastic.py:32-34)
    q_.ref_modulus = 1.0 / Cinv[0];
    const Stiff21 ci = pack21(Cinv);
    for (int k = 0; k < 21; ++k) q_.Cinv[k] = ci.c[k];
    for (int s = 0; s < NS; ++s)
      for (int m = 0; m < 6; ++m) q_.P[s][m] = d.schmid[s * 6 + m];
    for (int s = 0; s < NS; ++s)
      for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) q_.PP[s][sym6(a, b)] = q_.P[s][a] * q_.P[s][b];
    q_.gamma0 = d.gamma0;
    q_.n = d.n;
    const double e = d.n - 1.0;
    q_.nint = (e >= 0.0 && e <= 64.0 && std::floor(e) == e) ? int(e) : -1;
    q_.tol = kEvpTolerance;
    q_.hard = d.hardening;
    for (int k = 0; k < 4; ++k) q_.hp[k] = d.hp[k];
    per_system_slip_resistances(d);
    const long long n = std::max(this->n_, 1);
    auto alloc = [&](double*& ptr, long long cnt) { MTL_HIP(hipMalloc(&ptr, cnt * sizeof(double))); };
    alloc(B_.e0, blk_elems(n, 6));
    alloc(B_.gnorm, n);
    for (int k = 0; k < 2; ++k) {
      alloc(B_.sig[k], blk_elems(n, 6));
      alloc(B_.dsl[k], n);
    }
    MTL_HIP(hipMalloc(&B_.conv, n));
    MTL_HIP(hipMalloc(&B_.last, n * sizeof(int)));
    for (auto& st : slot_) {
      alloc(st.ep, blk_elems(n, 6));
      alloc(st.g, n);
      alloc(st.gam, blk_elems(n, NS));
      alloc(st.acc, n);
    }
    bind();
    MTL_HIP(hipMalloc(&ctrl_, sizeof(EvpControl)));
    MTL_HIP(hipMalloc(&qd_, sizeof(EvpParams<NS>)));
    MTL_HIP(hipHostMalloc(&qh_, sizeof(EvpParams<NS>), hipHostMallocDefault));
    MTL_HIP(hipHostMalloc(&hctrl_, sizeof(EvpControl), hipHostMallocDefault));
  }
  ~EvpModel() override {
    for (double* p : {B_.e0, B_.gnorm, B_.sig[0], B_.sig[1], B_.dsl[0], B_.dsl[1]})
      if (p) (void)hipFree(p);
    for (auto& st : slot_)
      for (double* p : {st.ep, st.g, st.gam, st.acc})
        if (p) (void)hipFree(p);
    if (B_.conv) (void)hipFree(B_.conv);
    if (B_.last) (void)hipFree(B_.last);
    if (ctrl_) (void)hipFree(ctrl_);
    if (qd_) (void)hipFree(qd_);
    if (qh_) (void)hipHostFree(qh_);
    if (hctrl_) (void)hipHostFree(hctrl_);
    (void)hipFree(this->idx_);
  }

  void init_tangent(const double* rot9, K* tangent, hipStream_t s) override {
    // elastic_viscoplastic.py:52-89: zero plastic state, g = g0
    const long long n = this->n_;
    if (n == 0) return;
    ring_ = SlotRing{};
    bind();
    const EvpState& st = slot_[ring_.old];  // committed and trial state
    fill(st.ep, blk_elems(n, 6), 0.0, s);
    fill(st.g, n, d_.g0, s);
    fill(st.gam, blk_elems(n, NS), 0.0, s);
    fill(st.acc, n, 0.0, s);
    hipLaunchKernelGGL(k_elastic<K>, dim3(ceil_div(this->n_, kCBT)), dim3(kCBT), 0, s, this->n_,
                       this->idx_, this->npts_, rot9, C_, nullptr, nullptr, tangent, 0);
    MTL_LAUNCH_CHECK();
  }

  int calculate(const double* rot9, const double* eps, const double* guess, double dt,
                double* sig, K* tangent, hipStream_t s) override {
    const int n = this->n_;
    if (n == 0) return 1;
    q_.dt = dt;
    *qh_ = q_;
    MTL_HIP(hipMemcpyAsync(qd_, qh_, sizeof(EvpParams<NS>), hipMemcpyHostToDevice, s));
    const int blocks = ceil_div(n, kCBT);
    bind();
    hipLaunchKernelGGL(k_evp_setup<NS>, dim3(blocks), dim3(kCBT), 0, s, n, this->idx_, this->npts_,
                       rot9, eps, guess, B_);
    MTL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_evp_ctrl_reset, dim3(1), dim3(64), 0, s, ctrl_);
    MTL_LAUNCH_CHECK();
    // iterations are enqueued in chunks; the kernels return at once after
    // convergence, and the host reads the control block once per chunk
    int it = 0;
    int chunk = 6;
    while (true) {
      for (int k = 0; k < chunk && it < kEvpMaxIterations; ++k) {
        ++it;
        profiler().launch(kKEvpIter, q_.nint >= 0 ? k_evp_iter<NS, true> : k_evp_iter<NS, false>,
                          dim3(blocks), dim3(kCBT), 0, s, it, n, qd_, B_, ctrl_);
        MTL_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_evp_control, dim3(1), dim3(64), 0, s, it, n, ctrl_);
        MTL_LAUNCH_CHECK();
      }
      MTL_HIP(hipMemcpyAsync(hctrl_, ctrl_, sizeof(EvpControl), hipMemcpyDeviceToHost, s));
      stream_sync(s);
      if (hctrl_->done || it >= kEvpMaxIterations) break;
      // a host round trip costs about as much as 8 no-op iterations past
      // convergence (workgroups with no unconverged point return at once)
      chunk = 8;
    }
    if (std::getenv("MTL_DEBUG") != nullptr) {
      std::fprintf(stderr, "[mtl] evp n=%d converging per iteration:", n);
      for (int k = 1; k <= std::min(it, kEvpMaxIterations); ++k)
        std::fprintf(stderr, " %d", hctrl_->count[k]);
      std::fprintf(stderr, " | K*=%d iterations=%d fail=%d\n", hctrl_->kstar, hctrl_->iterations,
                   hctrl_->fail);
    }
    if (hctrl_->fail || !hctrl_->done) return 0;
    profiler().launch(kKEvpFinish,
                      q_.nint >= 0 ? k_evp_outputs<NS, K, true> : k_evp_outputs<NS, K, false>,
                      dim3(blocks), dim3(kCBT), 0, s, n, this->idx_, this->npts_, rot9, qd_, B_,
                      ctrl_, sig, tangent);
    MTL_LAUNCH_CHECK();
    // adopt the call's outputs as the trial state (elastic_viscoplastic.py:175-182)
    ring_.adopt();
    bind();
    return hctrl_->iterations;
  }

  void commit(const double*, hipStream_t) override {
    // elastic_viscoplastic.py:255-271: the trial state becomes the committed one
    ring_.commit();
    bind();
  }

  bool field(int which, const double* rot9, std::vector<double>& out, int& ncomp,
             hipStream_t s) override {
    const int n = this->n_;
    std::vector<double> soa;
    // per-point arrays: one value per point, or nc components blocked by 64
    // points (blk_idx; for nc = 1 the two layouts coincide)
    auto fetch = [&](const double* src, int nc) {
      soa.resize(nc == 1 ? size_t(n) : size_t(blk_elems(n, nc)));
      if (n) MTL_HIP(hipMemcpyAsync(soa.data(), src, soa.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
      stream_sync(s);
      out.resize(size_t(nc) * n);
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < nc; ++c) out[size_t(i) * nc + c] = soa[size_t(blk_idx(c, i, nc))];
      ncomp = nc;
    };
    switch (which) {
      case kFieldPlasticStrain: {
        // committed crystal plastic strain, rotated to the specimen frame
        // stream-ordered temporary: no device-wide synchronisation
        double* tmp = nullptr;
        MTL_HIP(hipMallocAsync(reinterpret_cast<void**>(&tmp), std::max(6 * n, 1) * sizeof(double), s));
        if (n)
          hipLaunchKernelGGL(k_ep_specimen, dim3(ceil_div(n, kCBT)), dim3(kCBT), 0, s, n,
                             this->idx_, this->npts_, rot9, slot_[ring_.old].ep, tmp);
        out.resize(size_t(6) * n);
        if (n) MTL_HIP(hipMemcpyAsync(out.data(), tmp, out.size() * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
        MTL_HIP(hipFreeAsync(tmp, s));
        stream_sync(s);
        ncomp = 6;
        return true;
      }
      case kFieldSlipResistance: {
        // one value per point on the device (that of every system, or of the
        // reference system when they differ)
        fetch(slot_[ring_.tr].g, 1);
        std::vector<double> per_point(out);
        out.resize(size_t(NS) * n);
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < NS; ++k)
            out[size_t(i) * NS + k] =
                q_.persys ? q_.g0s[k] + q_.cs[k] * (per_point[size_t(i)] - q_.gref)
                          : per_point[size_t(i)];
        ncomp = NS;
        return true;
      }
      case kFieldSlipShear: fetch(slot_[ring_.tr].gam, NS); return true;
      case kFieldAccumulatedSlip: fetch(slot_[ring_.tr].acc, 1); return true;
      default: return false;
    }
  }

 private:
  // g0_s given per system: pick the reference system and the constants of
  // g_s = g0_s + c_s (g - g0_r) (see the top of this file); d_.g0 becomes the
  // reference system's initial value
  void per_system_slip_resistances(const ModelDesc& d) {
    q_.persys = 0;
    if (d.g0_sys.empty()) return;
    if (int(d.g0_sys.size()) != NS) throw Error("one initial slip resistance per slip system");
    const std::vector<double>& g0 = d.g0_sys;
    const bool af = d.hardening == kArmstrongFrederick && d.hp[1] != 0.0;
    const double sat = af ? d.hp[0] / d.hp[1] : 0.0;
    int ref = 0;
    if (af)
      for (int s = 1; s < NS; ++s)
        if (std::fabs(sat - g0[s]) > std::fabs(sat - g0[ref])) ref = s;
    d_.g0 = g0[ref];
    bool uniform = true;
    for (int s = 0; s < NS; ++s) uniform = uniform && g0[s] == g0[ref];
    if (uniform) return;
    q_.persys = 1;
    q_.gref = g0[ref];
    for (int s = 0; s < NS; ++s) {
      q_.g0s[s] = g0[s];
      q_.cs[s] = af && s != ref ? (sat - g0[s]) / (sat - g0[ref]) : 1.0;
    }
  }

  ModelDesc d_;
  Stiff21 C_;
  EvpParams<NS> q_{};
  EvpBuffers B_{};
  struct EvpState {
    double *ep = nullptr, *g = nullptr, *gam = nullptr, *acc = nullptr;
  };
  EvpState slot_[3];
  SlotRing ring_;
  // the kernels' view of the committed, free and trial sets
  void bind() {
    const EvpState &o = slot_[ring_.old], &p = slot_[ring_.pend], &t = slot_[ring_.tr];
    B_.ep_old = o.ep;
    B_.g_old = o.g;
    B_.gam_old = o.gam;
    B_.acc_old = o.acc;
    B_.ep_pend = p.ep;
    B_.g_pend = p.g;
    B_.gam_pend = p.gam;
    B_.acc_pend = p.acc;
    B_.g_trial = t.g;
  }
  EvpControl* ctrl_ = nullptr;
  EvpControl* hctrl_ = nullptr;
  EvpParams<NS>* qd_ = nullptr;  // device copy read by the kernels
  EvpParams<NS>* qh_ = nullptr;  // pinned staging copy
};

// ------------------------------------------------------------------ J2 plasticity
// IsotropicElasticPlastic (isotropic_elastic_plastic.py:17-192): elastic
// predictor s = 2G dev(eps - ep_old), p = K tr(eps - ep_old); where the von
// Mises predictor exceeds the committed yield stress, the equivalent plastic
// strain increment follows the fixed-point iteration
//   r = max(mises - 3G de - Y, 0),  de <- de + r / (3G),  Y <- h(de / dt)
// for every point together until |de_k - de_{k-1}| < 1e-7 at every point
// (at most 100 iterations).  The iteration count is global, so the device
// first runs every point's sequence once, OR-ing into a 128-bit mask the
// iterations at which the point has not converged (k_j2_scan); the first
// iteration with no bit set is the count the reference stops at, and
// k_j2_finish replays each point to exactly that iteration.
constexpr int kJ2MaxIterations = 100;    // isotropic_elastic_plastic.py:80
constexpr double kJ2Tolerance = 1.0e-7;  // :79

struct J2Params {
  double C[21];  // isotropic stiffness, packed Mandel
  double G, K, dt;
  double hp[4];
  int hard;
  int pad;
};

struct J2Control {
  unsigned int mask[4];  // bit k-1: some point not converged at iteration k
  int plastic;           // some point's predictor exceeds its yield stress
  int pad[3];
};

struct J2Buffers {
  double* ep_old;  // [6][n] committed plastic strain
  double* eq_old;  // [n] committed equivalent plastic strain
  double* y_old;   // [n] committed yield stress
  double* ep_pend; // outputs of the call
  double* eq_pend;
  double* y_pend;
};

// hardening law of one point: h(eq_old, y_old, rate = de / dt, dt), the
// reference's Scalar branch (hardening_laws.py:18-90, slip increment dt |rate|)
__device__ __forceinline__ double j2_harden(const J2Params& q, double eq_old, double y_old,
                                            double de) {
  const double d = q.dt * fabs(de / q.dt);
  if (q.hard == kLinear) return y_old + q.hp[0] * d;
  if (q.hard == kVoce) {
    const double theta0 = q.hp[0], theta1 = q.hp[1], tau1 = q.hp[2];
    const double ratio = fabs(theta0 / tau1);
    const double e0 = exp(-ratio * eq_old);
    const double ed = exp(-ratio * d);
    return y_old + theta1 * d - (ratio * tau1 - theta1) / ratio * e0 * (ed - 1.0) -
           theta1 / ratio * e0 * (ed * (1.0 + ratio * (eq_old + d)) - 1.0 - ratio * eq_old);
  }
  if (q.hard == kArmstrongFrederick) {
    const double h = q.hp[0], r = q.hp[1];
    return y_old + (h - r * y_old) * d / (1.0 + r * d);
  }
  return y_old;
}

struct J2Predictor {
  double s[6];   // deviatoric predictor stress
  double mean;   // K tr(elastic predictor strain)
  double mises;  // sqrt(3/2) |s|
};

__device__ __forceinline__ J2Predictor j2_predict(const J2Params& q, const double* eps,
                                                  long long npts, long long p,
                                                  const double* ep_old, long long n, int i) {
  J2Predictor r;
  double ee[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) ee[c] = eps[c * npts + p] - ep_old[c * n + i];
  const double tr = ee[0] + ee[1] + ee[2];
  const double vol = 1.0 / 3.0 * tr;
  double nrm = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double dv = c < 3 ? ee[c] - vol : ee[c];
    r.s[c] = 2.0 * q.G * dv;
    nrm += r.s[c] * r.s[c];
  }
  r.mean = q.K * tr;
  r.mises = sqrt(1.5) * sqrt(nrm);
  return r;
}

// one fixed-point iteration of a point (isotropic_elastic_plastic.py:108-126)
__device__ __forceinline__ void j2_step(const J2Params& q, double mises, double eq_old,
                                        double y_old, double& de, double& y, double& diff) {
  double r = mises - 3.0 * q.G * de - y;
  if (r <= 0.0) r = 0.0;
  const double prev = de;
  de = de - (-1.0 / (3.0 * q.G)) * r;
  y = j2_harden(q, eq_old, y_old, de);
  diff = de - prev;
}

__global__ void __launch_bounds__(kCBT) k_j2_scan(int n, const int* idx, long long npts,
                                                   const J2Params* qg, const double* eps,
                                                   J2Buffers B, J2Control* ctl) {
  __shared__ J2Params q;
  __shared__ unsigned int smask[4];
  __shared__ int splastic;
  load_params(&q, qg);
  if (threadIdx.x < 4) smask[threadIdx.x] = 0u;
  if (threadIdx.x == 0) splastic = 0;
  __syncthreads();
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i < n) {
    const long long p = idx[i];
    const J2Predictor pr = j2_predict(q, eps, npts, p, B.ep_old, n, i);
    const double y_old = B.y_old[i], eq_old = B.eq_old[i];
    if (!(pr.mises <= y_old)) {
      atomicOr(&splastic, 1);
      unsigned int m[4] = {0u, 0u, 0u, 0u};
      double de = 0.0, y = y_old, diff = 0.0;
      for (int k = 1; k <= kJ2MaxIterations; ++k) {
        j2_step(q, pr.mises, eq_old, y_old, de, y, diff);
        if (!(fabs(diff) < kJ2Tolerance)) m[(k - 1) >> 5] |= 1u << ((k - 1) & 31);
        else if (diff == 0.0) break;  // fixed point: every later iterate is identical
      }
#pragma unroll
      for (int w = 0; w < 4; ++w)
        if (m[w]) atomicOr(&smask[w], m[w]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 4 && smask[threadIdx.x]) atomicOr(&ctl->mask[threadIdx.x], smask[threadIdx.x]);
  if (threadIdx.x == 0 && splastic) atomicOr(&ctl->plastic, 1);
}

// Stress, tangent and (when `plastic`) pending state after `iters` iterations
// (isotropic_elastic_plastic.py:89-168); all-elastic calls return the
// predictor and the stiffness and leave the state alone (:89-95).
template <typename K>
__global__ void __launch_bounds__(kCBT) k_j2_finish(int n, const int* idx, long long npts,
                                                     const J2Params* qg, const double* eps,
                                                     int iters, int plastic, J2Buffers B,
                                                     double* sig, K* tangent) {
  __shared__ J2Params q;
  load_params(&q, qg);
  __syncthreads();
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  const long long p = idx[i];
  const long long ln = n;
  const J2Predictor pr = j2_predict(q, eps, npts, p, B.ep_old, n, i);
  const double y_old = B.y_old[i], eq_old = B.eq_old[i];
  const bool elastic = pr.mises <= y_old;
  double Ds[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) Ds[k] = q.C[k];
  if (!plastic) {
#pragma unroll
    for (int c = 0; c < 6; ++c) sig[c * npts + p] = pr.s[c] + (c < 3 ? pr.mean : 0.0);
    store_tangent(tangent, npts, p, Ds);
    return;
  }
  double de = 0.0, y = y_old, diff = 0.0;
  for (int k = 1; k <= iters; ++k) j2_step(q, pr.mises, eq_old, y_old, de, y, diff);
#pragma unroll
  for (int c = 0; c < 6; ++c)
    B.ep_pend[c * ln + i] = B.ep_old[c * ln + i] + 1.5 * (pr.s[c] / pr.mises) * de;
  B.eq_pend[i] = eq_old + de;
  B.y_pend[i] = y;
  const double ratio = 3.0 * q.G / pr.mises;
#pragma unroll
  for (int c = 0; c < 6; ++c)
    sig[c * npts + p] = (1.0 - ratio * de) * pr.s[c] + (c < 3 ? pr.mean : 0.0);
  if (!elastic) {
    const double term1 = 2.0 * q.G * ratio * de;
    const double term2 = ratio * ratio * (de / pr.mises - 1.0 / (3.0 * q.G));
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int b = a; b < 6; ++b) {
        const double idev = (a == b ? 1.0 : 0.0) - (a < 3 && b < 3 ? 1.0 / 3.0 : 0.0);
        Ds[sym6(a, b)] = q.C[sym6(a, b)] - term1 * idev + term2 * (pr.s[a] * pr.s[b]);
      }
  }
  store_tangent(tangent, npts, p, Ds);
}

// the stiffness as every point's tangent (isotropic: the same in every frame)
template <typename K>
__global__ void __launch_bounds__(kCBT) k_const_tangent(int n, const int* idx, long long npts,
                                                         Stiff21 C, K* tangent) {
  const int i = blockIdx.x * kCBT + threadIdx.x;
  if (i >= n) return;
  store_tangent(tangent, npts, idx[i], C.c);
}

__global__ void k_j2_ctrl_reset(J2Control* c) {
  int* w = reinterpret_cast<int*>(c);
  for (int i = threadIdx.x; i < int(sizeof(J2Control) / sizeof(int)); i += blockDim.x) w[i] = 0;
}

template <typename K>
class J2Model final : public Model<K> {
 public:
  J2Model(const ModelDesc& d, const std::vector<int>& pts, long long npts) : d_(d) {
    this->n_ = int(pts.size());
    this->npts_ = npts;
    this->host_idx_ = pts;
    this->idx_ = upload_idx(pts);
    const Stiff21 c = pack21(d.C);
    for (int k = 0; k < 21; ++k) q_.C[k] = c.c[k];
    q_.G = d.shear_modulus;
    q_.K = d.bulk_modulus;
    q_.hard = d.hardening;
    for (int k = 0; k < 4; ++k) q_.hp[k] = d.hp[k];
    const long long n = std::max(this->n_, 1);
    auto alloc = [&](double*& ptr, long long cnt) { MTL_HIP(hipMalloc(&ptr, cnt * sizeof(double))); };
    for (auto& st : slot_) {
      alloc(st.ep, 6 * n);
      alloc(st.eq, n);
      alloc(st.y, n);
    }
    bind();
    MTL_HIP(hipMalloc(&ctl_, sizeof(J2Control)));
    MTL_HIP(hipMalloc(&qd_, sizeof(J2Params)));
    MTL_HIP(hipHostMalloc(&qh_, sizeof(J2Params), hipHostMallocDefault));
    MTL_HIP(hipHostMalloc(&hctl_, sizeof(J2Control), hipHostMallocDefault));
  }
  ~J2Model() override {
    for (auto& st : slot_)
      for (double* p : {st.ep, st.eq, st.y})
        if (p) (void)hipFree(p);
    if (ctl_) (void)hipFree(ctl_);
    if (qd_) (void)hipFree(qd_);
    if (qh_) (void)hipHostFree(qh_);
    if (hctl_) (void)hipHostFree(hctl_);
    (void)hipFree(this->idx_);
  }

  void init_tangent(const double* rot9, K* tangent, hipStream_t s) override {
    // isotropic_elastic_plastic.py:55-75: zero plastic state, yield stress Y0
    const long long n = this->n_;
    if (n == 0) return;
    ring_ = SlotRing{};
    bind();
    const J2State& st = slot_[ring_.old];  // committed and trial state
    fill(st.ep, 6 * n, 0.0, s);
    fill(st.eq, n, 0.0, s);
    fill(st.y, n, d_.g0, s);
    hipLaunchKernelGGL(k_const_tangent<K>, dim3(ceil_div(this->n_, kCBT)), dim3(kCBT), 0, s,
                       this->n_, this->idx_, this->npts_, pack21(d_.C), tangent);
    MTL_LAUNCH_CHECK();
  }

  int calculate(const double* rot9, const double* eps, const double*, double dt, double* sig,
                K* tangent, hipStream_t s) override {
    const int n = this->n_;
    if (n == 0) return 1;
    q_.dt = dt;
    *qh_ = q_;
    MTL_HIP(hipMemcpyAsync(qd_, qh_, sizeof(J2Params), hipMemcpyHostToDevice, s));
    const int blocks = ceil_div(n, kCBT);
    hipLaunchKernelGGL(k_j2_ctrl_reset, dim3(1), dim3(64), 0, s, ctl_);
    MTL_LAUNCH_CHECK();
    profiler().launch(kKElastic, k_j2_scan, dim3(blocks), dim3(kCBT), 0, s, n, this->idx_, this->npts_, qd_, eps, B_, ctl_);
    MTL_LAUNCH_CHECK();
    MTL_HIP(hipMemcpyAsync(hctl_, ctl_, sizeof(J2Control), hipMemcpyDeviceToHost, s));
    stream_sync(s);
    int iters = 1;
    if (hctl_->plastic) {
      iters = 0;
      for (int k = 1; k <= kJ2MaxIterations; ++k)
        if (!((hctl_->mask[(k - 1) >> 5] >> ((k - 1) & 31)) & 1u)) {
          iters = k;
          break;
        }
      if (iters == 0) return 0;  // "Too many iterations in material loop" (:128-130)
    }
    profiler().launch(kKElastic, k_j2_finish<K>, dim3(blocks), dim3(kCBT), 0, s, n, this->idx_, this->npts_, qd_, eps, iters, hctl_->plastic, B_, sig, tangent);
    MTL_LAUNCH_CHECK();
    if (hctl_->plastic) {
      // the call's state becomes the trial state (isotropic_elastic_plastic.py:131-137)
      ring_.adopt();
      bind();
    }
    return iters;
  }

  void commit(const double*, hipStream_t) override {
    // isotropic_elastic_plastic.py:170-179: the trial state becomes the committed one
    ring_.commit();
    bind();
  }

  bool field(int which, const double*, std::vector<double>& out, int& ncomp,
             hipStream_t s) override {
    const int n = this->n_;
    auto fetch = [&](const double* src, int nc) {
      std::vector<double> soa(size_t(nc) * n);
      if (n) MTL_HIP(hipMemcpyAsync(soa.data(), src, soa.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
      stream_sync(s);
      out.resize(size_t(nc) * n);
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < nc; ++c) out[size_t(i) * nc + c] = soa[size_t(c) * n + i];
      ncomp = nc;
    };
    switch (which) {
      case kFieldPlasticStrain: fetch(slot_[ring_.tr].ep, 6); return true;
      case kFieldEqPlasticStrain: fetch(slot_[ring_.tr].eq, 1); return true;
      case kFieldYieldStress: fetch(slot_[ring_.tr].y, 1); return true;
      default: return false;
    }
  }

 private:
  ModelDesc d_;
  J2Params q_{};
  J2Buffers B_{};
  struct J2State {
    double *ep = nullptr, *eq = nullptr, *y = nullptr;
  };
  J2State slot_[3];
  SlotRing ring_;
  // the kernels' view of the committed and free sets
  void bind() {
    const J2State &o = slot_[ring_.old], &p = slot_[ring_.pend];
    B_.ep_old = o.ep;
    B_.eq_old = o.eq;
    B_.y_old = o.y;
    B_.ep_pend = p.ep;
    B_.eq_pend = p.eq;
    B_.y_pend = p.y;
  }
  J2Control* ctl_ = nullptr;
  J2Control* hctl_ = nullptr;
  J2Params* qd_ = nullptr;
  J2Params* qh_ = nullptr;
};

}  // namespace

template <typename K>
std::unique_ptr<Model<K>> make_model(const ModelDesc& d, const std::vector<int>& points,
                                     long long npts) {
  if (d.kind == 0) return std::make_unique<ElasticModel<K>>(d, points, npts);
  if (d.kind == 2) return std::make_unique<J2Model<K>>(d, points, npts);
  // FCC octahedral (12); HCP basal / prismatic (3), pyramidal <a> (6),
  // pyramidal <c+a> (12) and their concatenations (6, 9, 15, 18, 24)
  switch (d.nslip) {
    case 12: return std::make_unique<EvpModel<K, 12>>(d, points, npts);
    case 1: return std::make_unique<EvpModel<K, 1>>(d, points, npts);
    case 2: return std::make_unique<EvpModel<K, 2>>(d, points, npts);
    case 3: return std::make_unique<EvpModel<K, 3>>(d, points, npts);
    case 4: return std::make_unique<EvpModel<K, 4>>(d, points, npts);
    case 6: return std::make_unique<EvpModel<K, 6>>(d, points, npts);
    case 9: return std::make_unique<EvpModel<K, 9>>(d, points, npts);
    case 15: return std::make_unique<EvpModel<K, 15>>(d, points, npts);
    case 18: return std::make_unique<EvpModel<K, 18>>(d, points, npts);
    case 24: return std::make_unique<EvpModel<K, 24>>(d, points, npts);
    default:
      throw Error("elasto-viscoplastic model with " + std::to_string(d.nslip) +
                  " slip systems is not built (supported: 1, 2, 3, 4, 6, 9, 12, 15, 18, 24)");
  }
}

template std::unique_ptr<Model<float>> make_model<float>(const ModelDesc&, const std::vector<int>&,
                                                         long long);
template std::unique_ptr<Model<double>> make_model<double>(const ModelDesc&,
                                                           const std::vector<int>&, long long);

}  // namespace mtl
