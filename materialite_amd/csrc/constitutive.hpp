// Per-voxel constitutive models of the SmallStrainFFT path.
#pragma once

#include <memory>
#include <vector>

#include "spectral.hpp"

namespace mtl {

enum HardeningKind { kPerfectPlasticity = 0, kLinear = 1, kVoce = 2, kArmstrongFrederick = 3 };

constexpr int kEvpMaxIterations = 60;  // elastic_viscoplastic.py:93
constexpr double kEvpTolerance = 2.0e-5;  // elastic_viscoplastic.py:92

struct ModelDesc {
  int kind = 0;  // 0 elastic, 1 elasto-viscoplastic, 2 isotropic elastic-plastic (J2)
  double C[36] = {};  // crystal-frame stiffness, Mandel
  // elasto-viscoplastic
  int nslip = 0;
  std::vector<double> schmid;  // nslip*6, Mandel sym(d (x) n)
  double gamma0 = 0, n = 0, g0 = 0;  // g0: initial slip resistance / J2 yield stress
  std::vector<double> g0_sys;        // nslip initial slip resistances (empty: g0 for all)
  int hardening = kPerfectPlasticity;
  double hp[4] = {};
  // J2: shear and bulk moduli
  double shear_modulus = 0, bulk_modulus = 0;
};

// Device state and kernels of one model over its subset of points.
template <typename K>
class Model {
 public:
  virtual ~Model() = default;
  virtual void init_tangent(const double* rot9, K* tangent, hipStream_t s) = 0;
  // Returns the iteration count, or 0 when the update failed to converge.
  virtual int calculate(const double* rot9, const double* eps, const double* guess, double dt,
                        double* sig, K* tangent, hipStream_t s) = 0;
  virtual void commit(const double* rot9, hipStream_t s) = 0;
  // Host copy of a state field in point order (AoS, ncomp per point) into
  // `out` rows of this model's points; returns false if not available.
  virtual bool field(int which, const double* rot9, std::vector<double>& out, int& ncomp,
                     hipStream_t s) = 0;
  int count() const { return n_; }
  const std::vector<int>& points() const { return host_idx_; }

 protected:
  int n_ = 0;
  long long npts_ = 0;
  int* idx_ = nullptr;  // device, model-local -> global point
  std::vector<int> host_idx_;
};

template <typename K>
std::unique_ptr<Model<K>> make_model(const ModelDesc& d, const std::vector<int>& points,
                                     long long npts);

// slip-system counts the elasto-viscoplastic kernels are compiled for
inline bool evp_slip_count_built(int ns) {
  switch (ns) {
    case 1: case 2: case 3: case 4: case 6: case 9: case 12: case 15: case 18: case 24:
      return true;
    default:
      return false;
  }
}

enum FieldId {
  kFieldStress = 0,
  kFieldStrain = 1,
  kFieldPlasticStrain = 2,
  kFieldSlipResistance = 3,
  kFieldSlipShear = 4,
  kFieldAccumulatedSlip = 5,
  kFieldTangent = 6,
  kFieldEqPlasticStrain = 7,
  kFieldYieldStress = 8,
};

}  // namespace mtl
