"""ctypes binding of libmaterialite_amd.so (declared in include/materialite_amd.h).

The library is built in-tree by ``materialite_amd/build.py``.  There is no
CPU fallback: if the library is missing, cannot be loaded, or no HIP device is
present, ``lib()`` raises.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MATERIALITE_AMD_LIB selects another in-tree build of the same library
LIB_PATH = os.environ.get("MATERIALITE_AMD_LIB", os.path.join(HERE, "libmaterialite_amd.so"))

MTL_FP64 = 0
MTL_MIXED = 1
MTL_REFINED = 2
PRECISIONS = {"fp64": MTL_FP64, "mixed": MTL_MIXED, "refined": MTL_REFINED}

HARDENING = {
    "perfect_plasticity": 0,
    "linear": 1,
    "voce": 2,
    "armstrong_frederick": 3,
}

FIELD = {
    "stress": 0,
    "strain": 1,
    "plastic_strains": 2,
    "slip_resistances": 3,
    "slip_system_shear_strains": 4,
    "accumulated_slip": 5,
    "tangent": 6,
    "eq_plastic_strains": 7,
    "yield_stresses": 8,
}

_D6 = ctypes.c_double * 6


class StepParams(ctypes.Structure):
    _fields_ = [
        ("strain_increment", _D6),
        ("stress_increment", _D6),
        ("applied_stress", _D6),
        ("stress_mask", _D6),
        ("time_increment", ctypes.c_double),
        ("global_tolerance", ctypes.c_double),
        ("strain_correction_tolerance", ctypes.c_double),
        ("linear_solver_tolerance", ctypes.c_double),
        ("summed_von_mises_stress", ctypes.c_double),
        ("time_step_id", ctypes.c_int),
        ("reset_fluctuation", ctypes.c_int),
    ]


class StepReport(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("constitutive_failed", ctypes.c_int),
        ("newton_iterations", ctypes.c_int),
        ("krylov_iterations", ctypes.c_longlong),
        ("max_constitutive_iterations", ctypes.c_int),
        ("mean_von_mises_stress", ctypes.c_double),
        ("equilibrium_error", ctypes.c_double),
        ("scaled_strain_error", ctypes.c_double),
        ("projected_scaled_strain_error", ctypes.c_double),
        ("linear_solver_tolerance_used", ctypes.c_double),
    ]


EXPORTS = {
    "mtl_last_error": (ctypes.c_char_p, []),
    "mtl_solver_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "mtl_solver_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "mtl_solver_add_elastic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_add_evp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_add_evp_systems": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_add_j2": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_set_microstructure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
    "mtl_solver_initialize": (ctypes.c_int, [ctypes.c_void_p]),
    "mtl_solver_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(StepParams),
                                       ctypes.POINTER(StepReport)]),
    "mtl_solver_step_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(StepParams), ctypes.POINTER(StepReport)]),
    "mtl_solver_commit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "mtl_solver_rollback": (ctypes.c_int, [ctypes.c_void_p]),
    "mtl_solver_get_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_project": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "mtl_solver_minres": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int)]),
    "mtl_solver_constitutive": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_double, ctypes.c_void_p,
                                               ctypes.POINTER(ctypes.c_int)]),
    "mtl_points_create": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "mtl_solver_constitutive_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p, ctypes.c_double,
                                                    ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "mtl_project_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mtl_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "mtl_profile_reset": (ctypes.c_int, []),
    "mtl_profile_read": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_longlong)]),
    "mtl_profile_durations": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong,
                                             ctypes.POINTER(ctypes.c_longlong)]),
}

KERNELS = ["z_forward", "y_pass", "x_ghat", "z_inverse", "minres_vector", "minres_scalar",
           "evp_iteration", "evp_outputs", "elastic", "newton", "residual_z_forward",
           "residual_y_pass", "residual_x_ghat", "residual_z_inverse"]


def profile_enable(on=True):
    check(load_library().mtl_profile_enable(1 if on else 0))


def profile_reset():
    check(load_library().mtl_profile_reset())


def profile_read():
    """{kernel name: (total ms, launches)} since the last reset."""
    lib = load_library()
    out = {}
    for i, name in enumerate(KERNELS):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        check(lib.mtl_profile_read(i, ctypes.byref(ms), ctypes.byref(n)))
        out[name] = (ms.value, n.value)
    return out

def profile_durations(kernel):
    """Milliseconds of each timed launch of `kernel` (a name of KERNELS) since
    the last reset, as a float32 array in launch order."""
    import numpy as np

    lib = load_library()
    i = KERNELS.index(kernel)
    n = ctypes.c_longlong()
    check(lib.mtl_profile_durations(i, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.float32)
    if n.value:
        check(lib.mtl_profile_durations(i, out.ctypes.data, n.value, ctypes.byref(n)))
    return out


_LIB = None


def load_library(path=LIB_PATH):
    """Load and type the shared library (no device calls)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m materialite_amd.build` "
            "(the solver has no CPU fallback)"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(status):
    if status != 0:
        msg = load_library().mtl_last_error().decode(errors="replace")
        raise RuntimeError(f"materialite_amd: {msg}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class NativeSolver:
    """Owning wrapper of an ``mtl_solver`` handle."""

    def __init__(self, dims, precision="fp64"):
        self._lib = load_library()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, not {precision!r}")
        prec = PRECISIONS[precision]
        h = ctypes.c_void_p()
        nx, ny, nz = (int(d) for d in dims)
        check(self._lib.mtl_solver_create(nx, ny, nz, prec, ctypes.byref(h)))
        self._h = h
        self.dims = (nx, ny, nz)
        self.num_points = nx * ny * nz

    def close(self):
        if getattr(self, "_h", None):
            check(self._lib.mtl_solver_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_elastic(self, stiffness):
        c = np.ascontiguousarray(stiffness, dtype=np.float64).reshape(36)
        mid = ctypes.c_int()
        check(self._lib.mtl_solver_add_elastic(self._h, _ptr(c), ctypes.byref(mid)))
        return mid.value

    def add_evp(self, stiffness, schmid, reference_slip_rate, rate_exponent, slip_resistance,
                hardening, props):
        c = np.ascontiguousarray(stiffness, dtype=np.float64).reshape(36)
        s = np.ascontiguousarray(schmid, dtype=np.float64).reshape(-1, 6)
        p = np.zeros(4)
        p[: len(props)] = props
        mid = ctypes.c_int()
        g0 = np.asarray(slip_resistance, dtype=np.float64)
        if g0.ndim == 0:
            check(self._lib.mtl_solver_add_evp(
                self._h, _ptr(c), s.shape[0], _ptr(s), float(reference_slip_rate),
                float(rate_exponent), float(g0), int(hardening), _ptr(p), ctypes.byref(mid)))
        else:
            g0 = np.ascontiguousarray(g0.reshape(-1))
            if g0.size != s.shape[0]:
                raise ValueError(f"{g0.size} initial slip resistances for {s.shape[0]} slip systems")
            check(self._lib.mtl_solver_add_evp_systems(
                self._h, _ptr(c), s.shape[0], _ptr(s), float(reference_slip_rate),
                float(rate_exponent), _ptr(g0), int(hardening), _ptr(p), ctypes.byref(mid)))
        return mid.value

    def add_j2(self, stiffness, shear_modulus, bulk_modulus, yield_stress, hardening, props):
        c = np.ascontiguousarray(stiffness, dtype=np.float64).reshape(36)
        p = np.zeros(4)
        p[: len(props)] = props
        mid = ctypes.c_int()
        check(self._lib.mtl_solver_add_j2(self._h, _ptr(c), float(shear_modulus),
                                          float(bulk_modulus), float(yield_stress),
                                          int(hardening), _ptr(p), ctypes.byref(mid)))
        return mid.value

    def set_microstructure(self, rotations, model_of_point):
        r = np.ascontiguousarray(rotations, dtype=np.float64).reshape(self.num_points, 9)
        m = np.ascontiguousarray(model_of_point, dtype=np.int32).reshape(self.num_points)
        self._keep = (r, m)
        check(self._lib.mtl_solver_set_microstructure(self._h, _ptr(r), _ptr(m)))

    def initialize(self):
        check(self._lib.mtl_solver_initialize(self._h))

    def step(self, params):
        rep = StepReport()
        check(self._lib.mtl_solver_step(self._h, ctypes.byref(params), ctypes.byref(rep)))
        return rep

    @staticmethod
    def step_batch(solvers, params):
        """One increment of several solvers of one grid and precision together
        (their MINRES solves share launches); returns one report per solver."""
        lib = load_library()
        n = len(solvers)
        handles = (ctypes.c_void_p * n)(*[s._h.value for s in solvers])
        ps = (StepParams * n)(*params)
        reps = (StepReport * n)()
        check(lib.mtl_solver_step_batch(handles, n, ps, reps))
        return list(reps)

    def commit(self, fluctuation_scale):
        check(self._lib.mtl_solver_commit(self._h, float(fluctuation_scale)))

    def rollback(self):
        check(self._lib.mtl_solver_rollback(self._h))

    def get_field(self, name, ncomp_hint=36):
        # stress, strain and plastic strain always have 6 columns; other fields
        # get room for the widest one (36, the tangent) and are cut to size
        width = 6 if name in ("stress", "strain", "plastic_strains") else max(ncomp_hint, 36)
        out = np.empty((self.num_points, width), dtype=np.float64)
        nc = ctypes.c_int()
        check(self._lib.mtl_solver_get_field(self._h, FIELD[name], _ptr(out), ctypes.byref(nc)))
        if nc.value == width:
            return out
        flat = out.reshape(-1)[: self.num_points * nc.value]
        return flat.reshape(self.num_points, nc.value).copy()

    # kernel-level entry points: device pointers (ints)
    def project_dev(self, in_ptr, out_ptr, stress_mask):
        m = np.ascontiguousarray(stress_mask, dtype=np.float64)
        check(self._lib.mtl_solver_project(self._h, ctypes.c_void_p(in_ptr),
                                           ctypes.c_void_p(out_ptr), _ptr(m)))

    def minres_dev(self, b_ptr, x_ptr, stress_mask, rtol):
        m = np.ascontiguousarray(stress_mask, dtype=np.float64)
        it = ctypes.c_int()
        check(self._lib.mtl_solver_minres(self._h, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr),
                                          _ptr(m), float(rtol), ctypes.byref(it)))
        return it.value

    def constitutive_dev(self, strain_ptr, guess_ptr, dt, stress_ptr):
        it = ctypes.c_int()
        check(self._lib.mtl_solver_constitutive(self._h, ctypes.c_void_p(strain_ptr),
                                                ctypes.c_void_p(guess_ptr), float(dt),
                                                ctypes.c_void_p(stress_ptr), ctypes.byref(it)))
        return it.value


class NativePoints(NativeSolver):
    """Owning wrapper of a point handle (``mtl_points_create``): constitutive
    models over ``num_points`` material points and no grid."""

    def __init__(self, num_points, precision="fp64"):
        self._lib = load_library()
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, not {precision!r}")
        h = ctypes.c_void_p()
        check(self._lib.mtl_points_create(int(num_points), PRECISIONS[precision], ctypes.byref(h)))
        self._h = h
        self.dims = (int(num_points), 1, 1)
        self.num_points = int(num_points)

    def constitutive_host(self, strain, guess, dt):
        """(stress N x 6, iterations); iterations 0 when a model did not converge."""
        e = np.ascontiguousarray(strain, dtype=np.float64).reshape(self.num_points, 6)
        g = np.ascontiguousarray(guess, dtype=np.float64).reshape(self.num_points, 6)
        out = np.empty((self.num_points, 6))
        it = ctypes.c_int()
        check(self._lib.mtl_solver_constitutive_host(self._h, _ptr(e), _ptr(g), float(dt), _ptr(out),
                                                     ctypes.byref(it)))
        return out, it.value
