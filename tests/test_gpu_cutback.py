"""Cutback paths of the time loop on the device, against the oracle.

small_strain_fft.py:242-246 halves the increment and the fluctuation strain
when the constitutive update fails (at the Newton loop's first call, :137-139,
or after a MINRES correction, :162-164) or when the Newton loop reaches its
100-iteration cap (:209-211), and raises "min time increment reached" once
the increment falls below the minimum.  Each attempted increment is recorded
as (time, dt, converged, Newton iterations, Krylov iterations, constitutive
failure) by both the product and the oracle.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def _oracle_evp():
    from oracle import constitutive as C
    from oracle import tensors as T

    return C.ElasticViscoplastic(T.cubic_stiffness(250e3, 150e3, 120e3),
                                 T.schmid_mandel(*T.octahedral_slip_systems()), 1e-3, 20.0, 50.0,
                                 "voce", {"tau_1": 100.0, "theta_0": 500.0, "theta_1": 0.0})


def _material():
    """The reference's 8^3 Voronoi material (tests/test_elastic_model.py:100-123)
    and its rotations as the oracle builds them."""
    from tests.fixtures import voronoi_8
    from tests.test_gpu_solver import voronoi_material

    return voronoi_material(), voronoi_8()[0]


@pytest.mark.parametrize("precision", ["fp64", "refined"])
def test_constitutive_failure_cutbacks_match_oracle(precision):
    """128^3-workload EVP parameters on the 8^3 Voronoi material from an
    increment of 0.6 s: the first attempt fails in the update after its first
    MINRES correction, later ones at the Newton loop's first update; the
    accepted increments grow again after each success."""
    from materialite_amd import serving
    from materialite_amd.models.small_strain_fft import LoadSchedule, SmallStrainFFT
    from oracle.solver import LoadSchedule as OLS
    from oracle.solver import SmallStrainFFT as OSolver

    print(f"cutback {precision}", flush=True)
    material, rot = _material()
    osol = OSolver(OLS.from_constant_uniaxial_strain_rate(1e-3, "x"), end_time=1.6,
                   initial_time_increment=0.6)
    ref = osol.run([8, 8, 8], rot, _oracle_evp())
    solver = SmallStrainFFT(LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x"),
                            end_time=1.6, initial_time_increment=0.6,
                            constitutive_model=serving.config3_evp(), precision=precision)
    out = solver.run(material, output_variables=["slip_resistances"])
    hist, ref_hist = solver.stats["history"], osol.stats["history"]
    print("  device:", hist, "\n  oracle:", ref_hist, flush=True)
    failures = [h for h in ref_hist if h[5]]
    assert any(h[4] > 0 for h in failures) and any(h[4] == 0 for h in failures)
    assert len(hist) == len(ref_hist)
    for a, b in zip(hist, ref_hist):
        assert a[0] == pytest.approx(b[0], abs=1e-12) and a[1] == pytest.approx(b[1], abs=1e-12)
        assert (a[2], a[3], a[5]) == (b[2], b[3], b[5]), (a, b)
        # the refined form's MINRES stopping test may fall one iteration either side
        assert abs(a[4] - b[4]) <= (0 if precision == "fp64" else 1), (a, b)
    stress = out.extract("stress").components
    tol = 1e-9 if precision == "fp64" else 1e-6
    assert np.abs(stress - ref["stress"]).max() <= tol * np.abs(ref["stress"]).max()
    g = out.extract("slip_resistances").components
    g_ref = ref["slip_resistances"]
    assert np.abs(g - g_ref).max() <= tol * np.abs(g_ref).max()


def test_newton_cap_and_minimum_increment():
    """With a zero equilibrium tolerance no increment converges: each attempt
    stops at the 100-iteration cap, the increment halves, and the run ends
    with the reference's error once it falls below the minimum."""
    from materialite_amd.models.small_strain_fft import Elastic, LoadSchedule, SmallStrainFFT
    from materialite_amd import Order4SymmetricTensor

    print("newton cap", flush=True)
    material, _ = _material()
    solver = SmallStrainFFT(LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x"),
                            end_time=1.0, initial_time_increment=1.0, min_time_increment=0.3,
                            constitutive_model=Elastic(
                                Order4SymmetricTensor.from_cubic_constants(250e3, 150e3, 120e3)))
    with pytest.raises(ValueError, match="min time increment reached"):
        solver.run(material, global_tolerance=0.0)
    hist = solver.stats["history"]
    print("  device:", [h[:4] for h in hist], flush=True)
    # attempts at dt 1 and 0.5, each stopped at 100 Newton iterations; the
    # next halving (0.25) is below the minimum and raises before an attempt
    assert [h[1] for h in hist] == [1.0, 0.5]
    assert all(h[3] == 100 and not h[2] and not h[5] for h in hist)
