/* materialite_amd -- C ABI of the MI355X SmallStrainFFT solver.
 *
 * The reference (nasa/materialite v0.1.3) is pure Python and has no FFI; its
 * boundary for this path is the Python API
 *   SmallStrainFFT(load_schedule, ...)(material, ...)
 *   (materialite/models/small_strain_fft/small_strain_fft.py:29-264).
 * Each entry point below replaces the reference code cited beside it; the
 * Python package materialite_amd binds them with ctypes and keeps the
 * reference's Python API on top.
 *
 * Conventions: every function returns 0 on success and a negative status on
 * failure, with the message in mtl_last_error() (thread-local).  Host arrays
 * are plain C arrays; "dev" pointers are device (HBM) addresses, e.g. a torch
 * tensor's data_ptr().  Point order is x slowest, z fastest
 * (materialite/util/util.py:40-42).  Mandel order [xx, yy, zz, r2*yz, r2*xz,
 * r2*xy] (materialite/basis_operations.py:37-48).  A handle is used by one host
 * thread at a time; separate handles are independent.
 */
#ifndef MATERIALITE_AMD_H
#define MATERIALITE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mtl_solver mtl_solver;

/* precision of a solver handle */
enum { MTL_FP64 = 0, /* every field, FFT and Krylov vector in FP64 */
       MTL_MIXED = 1, /* Krylov vectors, tangents and matvec FFTs in FP32;
                         strain/stress, residual FFT and all norms in FP64 */
       MTL_REFINED = 2 /* FP32 tangent and MINRES w / x vectors; the matvec
                          spectrum and the Lanczos vectors stored as an FP32 value
                          plus a bf16 remainder (32 significant bits in 6 bytes);
                          FFT butterflies, twiddles and Ghat in FP64; strain/stress,
                          residual FFT and all norms in FP64 */ };

/* hardening laws (small_strain_fft/hardening_laws.py) */
enum { MTL_HARDENING_PERFECT = 0, /* :18-25 */
       MTL_HARDENING_LINEAR = 1,  /* :28-39, props[0] = hardening_rate */
       MTL_HARDENING_VOCE = 2,    /* :42-72, props = {theta_0, theta_1, tau_1} */
       MTL_HARDENING_ARMSTRONG_FREDERICK = 3 /* :75-90, props = {direct_hardening,
                                                dynamic_recovery} */ };

/* fields readable with mtl_solver_get_field (AoS, point-major) */
enum { MTL_FIELD_STRESS = 0,            /* 6 per point */
       MTL_FIELD_STRAIN = 1,            /* 6 */
       MTL_FIELD_PLASTIC_STRAIN = 2,    /* 6, committed, specimen frame */
       MTL_FIELD_SLIP_RESISTANCE = 3,   /* nslip */
       MTL_FIELD_SLIP_SHEAR = 4,        /* nslip */
       MTL_FIELD_ACCUMULATED_SLIP = 5,  /* 1 */
       MTL_FIELD_TANGENT = 6,           /* 36 (full Mandel 6x6) */
       MTL_FIELD_EQ_PLASTIC_STRAIN = 7, /* 1, J2 trial state */
       MTL_FIELD_YIELD_STRESS = 8       /* 1, J2 trial state */ };

const char* mtl_last_error(void);

/* Solver for one periodic grid.  Replaces the per-run setup of
 * SmallStrainFFT.run (small_strain_fft.py:72-104) including
 * _get_projection_operator (:301-368), which is never materialised here. */
int mtl_solver_create(int nx, int ny, int nz, int precision, mtl_solver** out);
int mtl_solver_destroy(mtl_solver* h);

/* A handle over num_points material points and no grid, for the constitutive
 * models' own protocol: initialize, calculate_stress_and_tangent and
 * update_state_variables of Elastic (elastic.py:23-46), ElasticViscoplastic
 * (elastic_viscoplastic.py:52-271) and IsotropicElasticPlastic
 * (isotropic_elastic_plastic.py:55-179), several models at once as Multiphase
 * (multiphase.py:31-62).  The model, microstructure, initialize, commit
 * (update_state_variables; the scale is ignored), get_field and constitutive
 * entry points apply; step, rollback, project and minres return an error.
 * precision: MTL_FP64 keeps the tangent in FP64, the other forms in FP32.
 * Destroy with mtl_solver_destroy. */
int mtl_points_create(long long num_points, int precision, mtl_solver** out);

/* Constitutive models (crystal-frame Mandel 6x6 stiffness, row-major).
 * Elastic: small_strain_fft/elastic.py:16-49. */
int mtl_solver_add_elastic(mtl_solver* h, const double stiffness[36], int* model_id);
/* ElasticViscoplastic: small_strain_fft/elastic_viscoplastic.py:20-294.
 * schmid: num_slip x 6 Mandel sym(d (x) n) (slip_system.py:139-141). */
int mtl_solver_add_evp(mtl_solver* h, const double stiffness[36], int num_slip,
                       const double* schmid, double reference_slip_rate,
                       double rate_exponent, double slip_resistance, int hardening,
                       const double hardening_props[4], int* model_id);
/* The same model with one initial slip resistance per slip system
 * (slip_resistances: num_slip values), the per-system Scalar that
 * elastic_viscoplastic.py:66-69 broadcasts onto every point. */
int mtl_solver_add_evp_systems(mtl_solver* h, const double stiffness[36], int num_slip,
                               const double* schmid, double reference_slip_rate,
                               double rate_exponent, const double* slip_resistances,
                               int hardening, const double hardening_props[4], int* model_id);

/* IsotropicElasticPlastic (J2 radial return with isotropic hardening):
 * small_strain_fft/isotropic_elastic_plastic.py:17-192.  stiffness: the
 * model's isotropic Mandel 6x6 (from_isotropic_constants, tensor.py:1452-1467);
 * bulk modulus as computed at isotropic_elastic_plastic.py:29-31; hardening
 * MTL_HARDENING_* with the props of mtl_solver_add_evp, applied to the
 * equivalent plastic strain. */
int mtl_solver_add_j2(mtl_solver* h, const double stiffness[36], double shear_modulus,
                      double bulk_modulus, double yield_stress, int hardening,
                      const double hardening_props[4], int* model_id);

/* Microstructure: per-point rotation matrices (N x 9, v_crystal = R v_specimen,
 * tensor.py:1855-1896) and per-point model ids (N int32; Multiphase dispatch,
 * small_strain_fft.py:384-398 and multiphase.py:18-51). */
int mtl_solver_set_microstructure(mtl_solver* h, const double* rotations,
                                  const int32_t* model_of_point);

/* Initial tangent and zeroed state (constitutive initialize(), elastic.py:23-27,
 * elastic_viscoplastic.py:52-89; fields of small_strain_fft.py:93-104). */
int mtl_solver_initialize(mtl_solver* h);

typedef struct {
  double strain_increment[6]; /* load_schedule.strain_increment(t, dt) */
  double stress_increment[6]; /* load_schedule.stress_increment(t, dt) */
  double applied_stress[6];   /* load_schedule.stress(t, dt) */
  double stress_mask[6];      /* load_schedule.stress_mask */
  double time_increment;
  double global_tolerance;
  double strain_correction_tolerance;
  double linear_solver_tolerance;
  double summed_von_mises_stress;
  int time_step_id;
  int reset_fluctuation; /* load reversal (small_strain_fft.py:111-116) */
} mtl_step_params;

typedef struct {
  int converged;
  int constitutive_failed;
  int newton_iterations;
  long long krylov_iterations;
  int max_constitutive_iterations;
  double mean_von_mises_stress;
  double equilibrium_error;
  double scaled_strain_error;
  double projected_scaled_strain_error;
  double linear_solver_tolerance_used; /* the requested one, or the FP32 form's floor
                                          (1e-5) when the request is tighter */
} mtl_step_report;

/* One load increment: predictor, Newton loop with MINRES corrections and
 * convergence tests (small_strain_fft.py:106-211). */
int mtl_solver_step(mtl_solver* h, const mtl_step_params* p, mtl_step_report* r);
/* One load increment of several solvers of the same grid and precision at
 * once -- the same-size requests of a serving stream (no counterpart in the
 * reference, which solves one material at a time): each solver runs the
 * Newton loop of mtl_solver_step on its own parameters, and the MINRES solves
 * of the solvers still iterating run together, one launch of each kernel per
 * iteration serving all of them.  reports[i] is what mtl_solver_step would
 * report for handles[i] with params[i]. */
int mtl_solver_step_batch(mtl_solver* const* handles, int count, const mtl_step_params* params,
                          mtl_step_report* reports);
/* Accept the increment (small_strain_fft.py:213-241): old fields <- new,
 * constitutive commit, fluctuation strain scaled by dt_new/dt. */
int mtl_solver_commit(mtl_solver* h, double fluctuation_scale);
/* Cut back (small_strain_fft.py:242-246): fluctuation strain halved. */
int mtl_solver_rollback(mtl_solver* h);
/* Copy a field to a host array (N x ncomp). ncomp receives the width. */
int mtl_solver_get_field(mtl_solver* h, int field, double* out, int* ncomp);

/* ---- kernel-level entry points (device pointers, stream-ordered) ---- */

/* out = F^-1[Ghat . F(in)] (real part) for a [6][N] SoA FP64 field, computed
 * in the handle's FFT precision (_apply_projection_tensor,
 * small_strain_fft.py:376-382). */
int mtl_solver_project(mtl_solver* h, const double* in_dev, double* out_dev,
                       const double stress_mask[6]);
/* MINRES solve of Ghat*(D:x) = b with the handle's current tangent
 * (_get_fluctuation_strain, small_strain_fft.py:281-299); b, x are [6][N]
 * FP64 device arrays. */
int mtl_solver_minres(mtl_solver* h, const double* b_dev, double* x_dev,
                      const double stress_mask[6], double rtol, int* iterations);
/* Constitutive update of every model (calculate_stress_and_tangent) on
 * [6][N] FP64 device strain and guess; stress is written to stress_dev and the
 * tangent to the handle.  iterations = 0 on non-convergence. */
int mtl_solver_constitutive(mtl_solver* h, const double* strain_dev, const double* guess_dev,
                            double time_increment, double* stress_dev, int* iterations);

/* calculate_stress_and_tangent on host arrays (a point handle): strain and
 * guess stress N x 6 point-major, stress written N x 6 when the update
 * converged; the tangent is then read with mtl_solver_get_field(h,
 * MTL_FIELD_TANGENT).  iterations = 0 when a model did not converge (the
 * reference returns (None, None, False) there, elastic_viscoplastic.py:166-168). */
int mtl_solver_constitutive_host(mtl_solver* h, const double* strain, const double* guess,
                                 double time_increment, double* stress, int* iterations);

/* ---- kernel timing (no counterpart in the reference, which has no profiler
 * hooks; used by bench.py for the roofline figures) ----
 * While enabled, HIP start/stop events ride on the dispatch of each library
 * kernel (on its launch stream).  kernel ids: 0 z-forward FFT pass, 1 y FFT pass, 2 x FFT pass
 * with Ghat, 3 z-inverse FFT pass, 4 MINRES vector updates, 5 MINRES scalar
 * steps, 6 EVP iteration, 7 EVP outputs, 8 elastic update, 9 Newton
 * reductions, 10-13 the four FFT passes (ids 0-3) when they project a
 * residual instead of applying the Krylov matvec.  The records are shared by
 * all handles and safe to use from several threads; while timing is on, the
 * kernels of concurrent handles are recorded together. */
int mtl_profile_enable(int enable);
int mtl_profile_reset(void);
int mtl_profile_read(int kernel, double* total_ms, long long* launches);
/* Milliseconds of each timed launch of `kernel` since the last reset, in
 * launch order: writes min(capacity, count) values and sets *count to the
 * number recorded.  No reference counterpart (measurement only). */
int mtl_profile_durations(int kernel, float* ms_out, long long capacity, long long* count);

/* ---- projection without device allocation ----
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Computes what mtl_solver_project computes for a handle of
 * the same grid and precision, with the spectrum and FFT tables in a fixed
 * device array of MTL_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of a handle's buffers.  Grids whose spectrum and tables need
 * more return an error (FP64: up to about 5.5e5 points, e.g. 80^3; the FP32
 * forms about twice that).  Runs on the null stream of the current device and
 * returns when done; calls that overlap on one GPU share the array and race. */
#define MTL_NOALLOC_SCRATCH_BYTES (64u << 20)
int mtl_project_noalloc(int nx, int ny, int nz, int precision, const double* in_dev,
                        double* out_dev, const double stress_mask[6]);

#ifdef __cplusplus
}
#endif

#endif /* MATERIALITE_AMD_H */
