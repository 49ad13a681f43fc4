"""The oracle reproduces the reference's end-to-end goldens (rtol 1e-7, the
``assert_allclose`` default the reference uses)."""

import numpy as np
import pytest
from numpy.testing import assert_allclose

from oracle import constitutive as C
from oracle import tensors as T
from oracle.solver import LoadSchedule, SmallStrainFFT
from tests import fixtures as F


def mean_norm(stress):
    return np.linalg.norm(stress.mean(axis=0))


def test_bicrystal_vs_golden():
    # tests/test_elastic_model.py:130-145
    ls = LoadSchedule.from_constant_rates([0, 0, 1.0, 0, 0, 0], np.zeros(6), [1, 1, 0, 0, 0, 0])
    out = SmallStrainFFT(ls, end_time=5e-5).run(
        [4, 4, 4], F.bicrystal_4(), C.Elastic(F.TI_STIFFNESS), linear_solver_tolerance=1e-7
    )
    assert_allclose(mean_norm(out["stress"]), 6.5890252)


def _pore_multiphase(model, pore):
    idx = [np.nonzero(~pore)[0], np.nonzero(pore)[0]]
    return C.Multiphase([model, C.Elastic(F.TI_STIFFNESS * 0.0)], idx, pore.size)


def test_stress_bc_closed_form():
    # tests/test_elastic_model.py:148-165
    rot, pore, _ = F.voronoi_8()
    ls = LoadSchedule.from_constant_rates(np.zeros(6), np.array([0, 0, 6, 0, 0, 0]) / 1e-4, np.ones(6))
    out = SmallStrainFFT(ls, end_time=1e-4).run(
        [8, 8, 8], rot, _pore_multiphase(C.Elastic(F.TI_STIFFNESS), pore),
        linear_solver_tolerance=1e-7,
    )
    assert_allclose(mean_norm(out["stress"]), 6.0)


def test_elastic_defect_vs_golden():
    # tests/test_elastic_model.py:168-180
    rot, pore, _ = F.voronoi_8()
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    out = SmallStrainFFT(ls, end_time=5e-5, initial_time_increment=5e-5).run(
        [8, 8, 8], rot, _pore_multiphase(C.Elastic(F.TI_STIFFNESS), pore),
        linear_solver_tolerance=1e-7,
    )
    assert_allclose(mean_norm(out["stress"]), 4.547568053)


@pytest.mark.parametrize("strain", [[-0.35, -0.35, 1.0, 0, 0, 0], [0, 0, 0, 0, 0, np.sqrt(2)]])
def test_single_crystal_closed_form(strain):
    # tests/test_elastic_model.py:183-222 (3x4x5 grid, identity orientation)
    rot = np.tile(np.eye(3), (60, 1, 1))
    strain = np.asarray(strain, float)
    ls = LoadSchedule.from_constant_rates(strain, np.zeros(6), np.zeros(6))
    out = SmallStrainFFT(ls, end_time=0.001).run(
        [3, 4, 5], rot, C.Elastic(F.TI_STIFFNESS), linear_solver_tolerance=1e-7
    )
    expected = F.TI_STIFFNESS @ strain * 0.001
    assert_allclose(out["stress"], np.tile(expected, (60, 1)), atol=1e-14)


@pytest.mark.parametrize("name", list(F.EVP_CASES))
def test_evp_vs_golden(name):
    # tests/test_elastic_viscoplastic_model.py:128-182
    props, end, expected, defect = F.EVP_CASES[name]
    rot, pore, _ = F.voronoi_8()
    model = F.evp_model(name, props)
    if defect:
        model = _pore_multiphase(model, pore)
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    out = SmallStrainFFT(ls, end_time=end, initial_time_increment=1e-3).run(
        [8, 8, 8], rot, model, global_tolerance=1e-7
    )
    assert_allclose(mean_norm(out["stress"]), expected)
    assert_allclose(out["strain"].mean(axis=0)[2], end)


def test_evp_strain_bcs_vs_golden():
    # tests/test_elastic_viscoplastic_model.py:185-224
    rot, _, _ = F.voronoi_8()
    rate = T.to_mandel(np.diag([-0.35, -0.35, 1.0]))
    ls = LoadSchedule.from_constant_rates(rate, np.zeros(6), np.zeros(6))
    out = SmallStrainFFT(ls, end_time=4e-3, initial_time_increment=1e-3).run(
        [8, 8, 8], rot, F.evp_model("perfect_plasticity", None), global_tolerance=1e-7
    )
    assert_allclose(mean_norm(out["stress"]), 428.328875)
    assert_allclose(out["stress"].mean(axis=0)[2], 426.923816)
    assert_allclose(out["strain"].mean(axis=0), rate * 0.004, atol=1e-14)


def _j2(name):
    hard, props, steps, golden = F.J2_CASES[name]
    return C.IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, F.J2_YIELD, hard, props), steps, golden


def _j2_run(name, steps):
    model = _j2(name)[0]
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(1.0, "z")
    out = SmallStrainFFT(ls, end_time=steps * 1e-4, initial_time_increment=1e-4).run(
        [8, 8, 8], np.broadcast_to(np.eye(3), (512, 3, 3)), model)
    return out


@pytest.mark.parametrize("name", ["perfect_plasticity", "linear", "voce"])
def test_j2_vs_golden(name):
    # tests/test_isotropic_plasticity.py:30-140
    _, steps, golden = _j2(name)
    out = _j2_run(name, steps)
    assert_allclose(mean_norm(out["stress"]), golden)
    axial = out["stress"][:, 2]
    assert_allclose(axial, axial.mean())
    # plastic flow is deviatoric
    assert_allclose(out["plastic_strains"][:, :3].sum(axis=-1), 0.0, atol=1e-15)
    # the reference's step relations: the axial stress of step k + 1 is the
    # hardened yield stress of the equivalent plastic strain of step k
    prev = _j2_run(name, steps - 1)["eq_plastic_strains"].mean()
    if name == "perfect_plasticity":
        assert_allclose(axial.mean(), 150.0)
    elif name == "linear":
        assert_allclose(axial.mean(), 1000.0 * prev + F.J2_YIELD)
    else:
        tau1, theta0, theta1 = 30.0, 10000.0, 100.0
        voce = F.J2_YIELD + (tau1 + theta1 * prev) * (1 - np.exp(-prev * abs(theta0 / tau1)))
        assert_allclose(axial.mean(), voce)


def test_j2_multiphase_vs_golden():
    # tests/test_multiphase.py:42-74: two identical linear-hardening phases
    box = F.box_phase_8()
    models = [_j2("linear")[0], _j2("linear")[0]]
    model = C.Multiphase(models, [np.nonzero(~box)[0], np.nonzero(box)[0]], 512)
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(1.0, "z")
    out = SmallStrainFFT(ls, end_time=15e-4, initial_time_increment=1e-4).run(
        [8, 8, 8], np.broadcast_to(np.eye(3), (512, 3, 3)), model)
    assert_allclose(mean_norm(out["stress"]), 150.3980086)
    eq = np.concatenate([m.eq for m in models])
    assert_allclose(eq, eq.mean())
