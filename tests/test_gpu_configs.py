"""The five benchmark workloads on the device, each against its oracle.

* 16^3 elastic, 20 grains, 10 steps, tolerances 1e-8 / 1e-10: every form's
  full stress and strain fields against the NumPy FP64 oracle.
* 64^3 elastic, 200 grains: the first load step of uniaxial x and of pure xy
  shear against the oracle.
* 128^3 FCC EVP, 1000 grains (the bench workload), the first 1.6 s of load:
  the FP32-storage forms against the FP64 form on the device (the NumPy path
  cannot hold this size's EVP temporaries at useful speed): the same Newton
  and Krylov counts in every increment, the stress within STRESS_TOL and the
  slip resistances within 1e-4 of the largest value.
* Serving stream requests of 32^3-96^3: the same result alone and in a mix.
* 256^3 two-phase EVP, the first 1.0 s of load: as the 128^3 case.

Microstructures are the seeded periodic-Voronoi polycrystals of the bench
(materialite_amd.serving.grain_material / two_phase_material).
"""

import numpy as np
import pytest

from materialite_amd import serving

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# The refined form is held to 1e-6 of the largest stress (it follows the FP64
# iterates to a few 1e-7, inside the 1e-5 target).  The FP32 form is outside
# that target at these sizes (measured 1.27e-5 after 13 increments at 128^3);
# it is held to the 5e-5 it is documented at.
STRESS_TOL = {"refined": 1e-6, "mixed": 5e-5}
SLIP_TOL = 1e-4


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _solve(material, schedule, model, precision, end_time, dt, **kw):
    """Device solve; returns (stress, strain, outputs, per-increment counts)."""
    from materialite_amd.models.small_strain_fft import SmallStrainFFT

    counts = []

    def cb(info):
        counts.append((info["newton_iterations"], info["krylov_iterations"]))
        return False

    solver = SmallStrainFFT(schedule, end_time=end_time, initial_time_increment=dt,
                            max_time_increment=dt, constitutive_model=model, precision=precision)
    out = solver.run(material, increment_callback=cb, **kw)
    return out, counts


def _oracle_elastic(n, grains, seed, load, end_time, dt):
    from oracle import constitutive as OC
    from oracle import tensors as OT
    from oracle.solver import LoadSchedule as OLS
    from oracle.solver import SmallStrainFFT as OSolver

    m = serving.grain_material(n, grains, seed)
    rot = m.extract("orientation").rotation_matrix.reshape(-1, 3, 3)
    if load == "x":
        ls = OLS.from_constant_uniaxial_strain_rate(1e-3, "x")
    else:  # pure xy shear: Mandel slot 5 at sqrt(2) 1e-3, the other stresses zero
        rate = np.zeros(6)
        rate[5] = np.sqrt(2.0) * 1e-3
        mask = np.ones(6)
        mask[5] = 0.0
        ls = OLS.from_constant_rates(rate, np.zeros(6), mask)
    solver = OSolver(ls, end_time=end_time, initial_time_increment=dt, max_time_increment=dt)
    ref = solver.run([n] * 3, rot, OC.Elastic(OT.cubic_stiffness(250e3, 150e3, 120e3)),
                     global_tolerance=1e-8, linear_solver_tolerance=1e-10)
    return m, ref


# ------------------------------------------------------------------ 16^3
@pytest.fixture(scope="module")
def config1():
    print("16^3 elastic: oracle", flush=True)
    return _oracle_elastic(16, 20, 20260001, "x", 10.0, 1.0)


@pytest.mark.parametrize("precision", ["fp64", "refined", "mixed"])
def test_config1_16_elastic_fields_match_oracle(config1, precision):
    from materialite_amd.models.small_strain_fft import LoadSchedule

    material, ref = config1
    print(f"16^3 elastic {precision}", flush=True)
    out, counts = _solve(material, LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x"),
                         serving.cubic_elastic(), precision, 10.0, 1.0,
                         global_tolerance=1e-8, linear_solver_tolerance=1e-10)
    assert len(counts) == 10
    es = _rel(out.extract("stress").components, ref["stress"])
    ee = _rel(out.extract("strain").components, ref["strain"])
    print(f"  stress {es:.3e} strain {ee:.3e}", flush=True)
    tol = {"fp64": 1e-9, "refined": 1e-6, "mixed": 1e-5}[precision]
    assert es <= tol and ee <= tol


# ------------------------------------------------------------------ 64^3
@pytest.fixture(scope="module")
def config2():
    print("64^3 elastic: oracle, first step of each load", flush=True)
    return {load: _oracle_elastic(64, 200, 20260002, load, 1.0, 1.0) for load in ("x", "xy")}


@pytest.mark.parametrize("precision", ["fp64", "refined", "mixed"])
@pytest.mark.parametrize("load", ["x", "xy"])
def test_config2_64_elastic_first_step_matches_oracle(config2, load, precision):
    from materialite_amd.models.small_strain_fft import LoadSchedule

    material, ref = config2[load]
    schedule = (LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x") if load == "x"
                else serving.stream_load_schedule("xy"))
    print(f"64^3 elastic {load} {precision}", flush=True)
    out, counts = _solve(material, schedule, serving.cubic_elastic(), precision, 1.0, 1.0,
                         global_tolerance=1e-8, linear_solver_tolerance=1e-10)
    es = _rel(out.extract("stress").components, ref["stress"])
    print(f"  stress {es:.3e}", flush=True)
    tol = {"fp64": 1e-9, "refined": 1e-6, "mixed": 1e-5}[precision]
    assert es <= tol


# ------------------------------------------------------------------ 128^3, 256^3 EVP
def _evp_compare(material, schedule, end_time, precision, ref, **kw):
    out, counts = _solve(material, schedule, kw.pop("model", None), precision, end_time, 0.2,
                         output_variables=["slip_resistances"], **kw)
    stress = out.extract("stress").components
    g = out.extract("slip_resistances").components
    ref_out, ref_counts = ref
    es = _rel(stress, ref_out.extract("stress").components)
    eg = _rel(g, ref_out.extract("slip_resistances").components)
    print(f"  {precision}: counts {counts}; stress {es:.3e}, slip resistances {eg:.3e}", flush=True)
    assert counts == ref_counts
    assert es <= STRESS_TOL[precision]
    assert eg <= SLIP_TOL


@pytest.fixture(scope="module")
def config3():
    from materialite_amd.models.small_strain_fft import LoadSchedule

    print("128^3 EVP: material and FP64 form", flush=True)
    material = serving.grain_material(128, 1000, 20260003)
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x")
    ref = _solve(material, ls, serving.config3_evp(), "fp64", 1.6, 0.2,
                 output_variables=["slip_resistances"])
    return material, ls, ref


@pytest.mark.parametrize("precision", ["refined", "mixed"])
def test_config3_128_evp_matches_fp64_form(config3, precision):
    material, ls, ref = config3
    print(f"128^3 EVP {precision}", flush=True)
    _evp_compare(material, ls, 1.6, precision, ref, model=serving.config3_evp())


@pytest.fixture(scope="module")
def config5():
    from materialite_amd.models.small_strain_fft import LoadSchedule

    print("256^3 two-phase EVP: material and FP64 form", flush=True)
    material, _ = serving.two_phase_material()
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "z")
    ref = _solve(material, ls, None, "fp64", 1.0, 0.2, phase_label="phase",
                 output_variables=["slip_resistances"])
    return material, ls, ref


@pytest.mark.parametrize("precision", ["refined", "mixed"])
def test_config5_256_two_phase_matches_fp64_form(config5, precision):
    material, ls, ref = config5
    print(f"256^3 two-phase {precision}", flush=True)
    _evp_compare(material, ls, 1.0, precision, ref, phase_label="phase")


# ------------------------------------------------------------------ serving
def test_config4_stream_requests_same_alone_and_in_mix():
    """Requests of every grid size of the stream (32^3-96^3), elastic and EVP,
    normal and shear loads, run alone and then four at once on one GPU."""
    specs = [serving.StreamSpec(seed=21, n=32, grains=50, model="evp", direction="xx", steps=4),
             serving.StreamSpec(seed=22, n=48, grains=120, model="elastic", direction="xy", steps=4),
             serving.StreamSpec(seed=23, n=64, grains=300, model="evp", direction="yz", steps=4),
             serving.StreamSpec(seed=24, n=96, grains=500, model="elastic", direction="zz", steps=4)]
    alone = []
    for sp in specs:
        print(f"alone: {sp}", flush=True)
        alone.append(serving.build_stream_request(sp, keep_material=True).run())
    server = serving.RveServer(hbm_budget=torch.cuda.mem_get_info()[0] // 2, max_slots=4)
    reqs = [serving.build_stream_request(sp, keep_material=True) for sp in specs]
    print("mix: submitting", flush=True)
    futs = [server.submit(r) for r in reqs]
    mixed = [f.result(timeout=300) for f in futs]
    server.close()
    assert server.peak_in_flight >= 2
    for sp, a, m in zip(specs, alone, mixed):
        assert a["krylov"] == m["krylov"], sp
        scale = np.abs(a["stress"]).max()
        assert np.abs(a["stress"] - m["stress"]).max() <= 1e-12 * scale, sp
