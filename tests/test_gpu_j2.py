"""IsotropicElasticPlastic (J2) on the device: the reference's goldens
(tests/test_isotropic_plasticity.py, tests/test_multiphase.py) through the
product API, and full fields and state against the oracle."""

import numpy as np
import pandas as pd
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from materialite_amd import Box, Material, Orientation  # noqa: E402
from materialite_amd.models.small_strain_fft import (  # noqa: E402
    IsotropicElasticPlastic, LoadSchedule, SmallStrainFFT, linear, perfect_plasticity, voce,
)
from oracle import constitutive as OC  # noqa: E402
from oracle import solver as OS  # noqa: E402
from tests import fixtures as F  # noqa: E402

PRECISIONS = ["fp64", "mixed", "refined"]
LAWS = {"perfect_plasticity": perfect_plasticity, "linear": linear, "voce": voce}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def material():
    return (Material(dimensions=[8, 8, 8], sizes=[3, 3, 3])
            .create_uniform_field("orientation", Orientation(np.eye(3)))
            .create_uniform_field("phase", 1))


def j2(name):
    hard, props, steps, golden = F.J2_CASES[name]
    return (IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, F.J2_YIELD, LAWS[name], props),
            steps, golden)


def rtol_for(precision):
    return 1e-7 if precision == "fp64" else 1e-5


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", list(LAWS))
def test_j2_golden(name, precision):
    # tests/test_isotropic_plasticity.py:30-140, with the same output-time stacking
    print(f"j2 {name} {precision}", flush=True)
    model_c, steps, golden = j2(name)
    dt = 1.0e-4
    model = SmallStrainFFT(load_schedule=LoadSchedule.from_constant_uniaxial_strain_rate(direction="z"),
                           end_time=steps * dt, initial_time_increment=dt,
                           constitutive_model=model_c, precision=precision)
    out_times = np.round((np.arange(steps) + 1) * dt, 10)
    out = model(material(), output_times=out_times, output_variables=["eq_plastic_strains"])
    stress = out.extract("stress")
    assert stress.components.shape == (512, steps, 6)
    final = stress.components[:, -1]
    assert_allclose(np.linalg.norm(final.mean(axis=0)), golden, rtol=rtol_for(precision))
    axial = stress.components.mean(axis=0)[:, 2]
    assert_allclose(final[:, 2], axial[-1], rtol=rtol_for(precision))
    eq = out.extract("eq_plastic_strains").components.mean(axis=0)
    rtol = rtol_for(precision)
    if name == "perfect_plasticity":
        assert_allclose(axial[9:], 150.0, rtol=rtol)
    elif name == "linear":
        assert_allclose(axial[12:], 1000.0 * eq[11:-1] + 150.0, rtol=rtol)
    else:
        tau1, theta0, theta1 = 30.0, 10000.0, 100.0
        v = 150.0 + (tau1 + theta1 * eq[10:-1]) * (1 - np.exp(-eq[10:-1] * abs(theta0 / tau1)))
        assert_allclose(axial[11:], v, rtol=rtol)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_j2_multiphase_golden(precision):
    # tests/test_multiphase.py:13-74
    print(f"j2 multiphase {precision}", flush=True)
    m1, m2 = j2("linear")[0], j2("linear")[0]
    mat = (Material(dimensions=[8, 8, 8], sizes=[3, 3, 3])
           .create_uniform_field("orientation", Orientation(np.eye(3)))
           .create_uniform_field("phase", 0)
           .insert_feature(Box(max_corner=[2, 2, 2]), fields={"phase": 1})
           .create_regional_fields(region_label="phase", regional_fields={
               "phase": [0, 1], "constitutive_model": [m1, m2]}))
    dt = 1.0e-4
    model = SmallStrainFFT(load_schedule=LoadSchedule.from_constant_uniaxial_strain_rate(direction="z"),
                           end_time=15 * dt, initial_time_increment=dt, precision=precision)
    out = model(mat, output_times=np.round((np.arange(15) + 1) * dt, 10),
                output_variables=["eq_plastic_strains"], phase_label="phase")
    stress = out.extract("stress").components
    assert_allclose(np.linalg.norm(stress[:, -1].mean(axis=0)), 150.3980086, rtol=rtol_for(precision))
    eq = out.extract("eq_plastic_strains").components
    assert_allclose(eq, np.tile(eq.mean(axis=0), (len(eq), 1)), rtol=1e-9 if precision == "fp64" else 1e-5)


def test_j2_two_phase_fields_match_oracle():
    """Heterogeneous J2 (two phases with different yield stresses and laws):
    stress, strain and state per point vs the oracle, FP64 form."""
    print("j2 two-phase vs oracle", flush=True)
    box = F.box_phase_8()
    soft = ("linear", {"hardening_rate": 1000.0}, 120.0)
    hard = ("voce", {"tau_1": 30.0, "theta_0": 10000.0, "theta_1": 100.0}, 180.0)
    laws = {"linear": linear, "voce": voce}
    mats = [IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, y, laws[h], p) for h, p, y in (soft, hard)]
    mat = (Material(dimensions=[8, 8, 8], sizes=[3, 3, 3])
           .create_uniform_field("orientation", Orientation(np.eye(3)))
           .create_fields({"phase": box.astype(np.int64)})
           .create_regional_fields(region_label="phase", regional_fields={
               "phase": [0, 1], "constitutive_model": mats}))
    dt, steps = 1.0e-4, 15
    model = SmallStrainFFT(load_schedule=LoadSchedule.from_constant_uniaxial_strain_rate(direction="z"),
                           end_time=steps * dt, initial_time_increment=dt, precision="fp64")
    out = model(mat, phase_label="phase", global_tolerance=1e-8, linear_solver_tolerance=1e-10,
                output_variables=["eq_plastic_strains", "yield_stresses", "plastic_strains"])

    omodels = [OC.IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, y, h, p) for h, p, y in (soft, hard)]
    omodel = OC.Multiphase(omodels, [np.nonzero(~box)[0], np.nonzero(box)[0]], 512)
    ls = OS.LoadSchedule.from_constant_uniaxial_strain_rate(1.0, "z")
    ref = OS.SmallStrainFFT(ls, end_time=steps * dt, initial_time_increment=dt).run(
        [8, 8, 8], np.broadcast_to(np.eye(3), (512, 3, 3)), omodel,
        global_tolerance=1e-8, linear_solver_tolerance=1e-10)
    s = out.extract("stress").components
    assert_allclose(s, ref["stress"], rtol=1e-9, atol=1e-9 * np.abs(ref["stress"]).max())
    eq = np.zeros(512)
    y = np.zeros(512)
    for m, idx in zip(omodels, omodel.indices):
        eq[idx] = m.eq
        y[idx] = m.y
    assert np.abs(eq).max() > 0  # both phases yielded
    assert_allclose(out.extract("eq_plastic_strains").components, eq, rtol=1e-8, atol=1e-14)
    assert_allclose(out.extract("yield_stresses").components, y, rtol=1e-12)
