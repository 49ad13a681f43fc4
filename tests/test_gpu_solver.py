"""End-to-end SmallStrainFFT on the device: the reference's goldens and the oracle.

The materials are built with materialite_amd's own Material exactly as the
reference's tests build theirs (tests/test_elastic_model.py,
tests/test_elastic_viscoplastic_model.py).
"""

import numpy as np
import pandas as pd
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from materialite_amd import (  # noqa: E402
    Box, Material, Order2SymmetricTensor, Order4SymmetricTensor, Orientation, SlipSystem, Sphere,
)
from materialite_amd.models.small_strain_fft import (  # noqa: E402
    Elastic, ElasticViscoplastic, LoadSchedule, SmallStrainFFT, armstrong_frederick, linear,
    perfect_plasticity, voce,
)

TI = np.array([
    [160e3, 90e3, 66e3, 0, 0, 0],
    [90e3, 160e3, 66e3, 0, 0, 0],
    [66e3, 66e3, 181.7e3, 0, 0, 0],
    [0, 0, 0, 46.5e3, 0, 0],
    [0, 0, 0, 0, 46.5e3, 0],
    [0, 0, 0, 0, 0, 35e3],
])


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def stiffness():
    return Order4SymmetricTensor.from_voigt(TI)


def mean_norm(material):
    return material.extract("stress").mean().norm.components


def voronoi_material():
    rng = np.random.default_rng(0)
    return (Material(dimensions=[8, 8, 8], sizes=[3, 3, 3])
            .create_voronoi(num_regions=10, rng=rng)
            .assign_random_orientations(rng=rng)
            .create_uniform_field("phase", 1))


def with_pore(material, model):
    mid = material.sizes[2] / 2
    m = material.insert_feature(Sphere(radius=1, centroid=[mid, mid, mid]), fields={"phase": 2})
    table = pd.DataFrame({"phase": [1, 2], "constitutive_model": [model, Elastic(model.stiffness * 0.0)]})
    return m.create_regional_fields("phase", table)


PRECISIONS = ["fp64", "mixed", "refined"]


def rtol_for(precision):
    # FP64 reproduces the reference's goldens at assert_allclose's default;
    # the mixed forms are checked against them at the north-star field tolerance
    return 1e-7 if precision == "fp64" else 1e-5


@pytest.mark.parametrize("precision", PRECISIONS)
def test_bicrystal_golden(precision):
    print(f"bicrystal {precision}", flush=True)
    material = Material(dimensions=[4, 4, 4])
    material = material.create_uniform_field(
        "orientation", Orientation.from_euler_angles([0, np.pi / 4, 0])
    ).insert_feature(Box(max_corner=[None, None, material.sizes[2] / 2]),
                     fields={"orientation": Orientation.identity()})
    ls = LoadSchedule.from_constant_rates(Order2SymmetricTensor([0, 0, 1.0, 0, 0, 0]),
                                          Order2SymmetricTensor.zero(), np.array([1, 1, 0, 0, 0, 0]))
    model = SmallStrainFFT(load_schedule=ls, end_time=5.0e-5, constitutive_model=Elastic(stiffness()),
                           precision=precision)
    out = model(material, linear_solver_tolerance=1.0e-7)
    assert_allclose(mean_norm(out), 6.5890252, rtol=rtol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_elastic_defect_golden(precision):
    print(f"elastic defect {precision}", flush=True)
    material = with_pore(voronoi_material(), Elastic(stiffness()))
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    model = SmallStrainFFT(load_schedule=ls, end_time=5.0e-5, initial_time_increment=5.0e-5,
                           precision=precision)
    out = model(material, phase_label="phase", linear_solver_tolerance=1.0e-7)
    assert_allclose(mean_norm(out), 4.547568053, rtol=rtol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_stress_bc_closed_form(precision):
    print(f"stress bc {precision}", flush=True)
    material = with_pore(voronoi_material(), Elastic(stiffness()))
    end_time = 1.0e-4
    ls = LoadSchedule.from_constant_rates(Order2SymmetricTensor.zero(),
                                          Order2SymmetricTensor(np.array([0, 0, 6, 0, 0, 0])) / end_time,
                                          np.ones(6))
    model = SmallStrainFFT(load_schedule=ls, end_time=end_time, precision=precision)
    out = model(material, phase_label="phase", linear_solver_tolerance=1.0e-7)
    assert_allclose(mean_norm(out), 6.0, rtol=rtol_for(precision))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_single_crystal_closed_form(precision):
    print(f"single crystal {precision}", flush=True)
    material = Material(dimensions=[3, 4, 5]).create_uniform_field("orientation", Orientation.identity())
    rate = Order2SymmetricTensor([-0.35, -0.35, 1.0, 0, 0, 0])
    ls = LoadSchedule.from_constant_rates(rate, Order2SymmetricTensor.zero(), np.zeros(6))
    model = SmallStrainFFT(load_schedule=ls, end_time=0.001, constitutive_model=Elastic(stiffness()),
                           precision=precision)
    out = model(material, linear_solver_tolerance=1.0e-7)
    expected = (stiffness() @ rate * 0.001).components
    atol = 1e-14 if precision == "fp64" else 1e-6
    for s in out.extract("stress").components:
        assert_allclose(s, expected, atol=atol)


EVP = {
    "linear": (linear, {"hardening_rate": 10}, 4e-3, 361.276725, True),
    "voce": (voce, {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}, 5e-3, 541.043871, False),
    "armstrong_frederick": (armstrong_frederick, {"direct_hardening": 10000.0, "dynamic_recovery": 15.0},
                            5e-3, 539.345343, False),
    "perfect_plasticity": (perfect_plasticity, None, 4e-3, 442.608112, False),
}


def evp(fn, props):
    return ElasticViscoplastic(stiffness=stiffness(), slip_systems=SlipSystem.octahedral(),
                               reference_slip_rate=1.0, rate_exponent=10.0, slip_resistance=300.0,
                               hardening_function=fn, hardening_properties=props)


@pytest.mark.parametrize("precision", PRECISIONS)
@pytest.mark.parametrize("name", list(EVP))
def test_evp_golden(name, precision):
    print(f"evp {name} {precision}", flush=True)
    fn, props, end, expected, defect = EVP[name]
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    if defect:
        material = with_pore(voronoi_material(), evp(fn, props))
        model = SmallStrainFFT(load_schedule=ls, end_time=end, initial_time_increment=1e-3,
                               precision=precision)
        out = model(material, phase_label="phase", global_tolerance=1.0e-7)
    else:
        model = SmallStrainFFT(load_schedule=ls, end_time=end, initial_time_increment=1e-3,
                               constitutive_model=evp(fn, props), precision=precision)
        out = model(voronoi_material(), global_tolerance=1.0e-7)
    assert_allclose(mean_norm(out), expected, rtol=rtol_for(precision))
    assert_allclose(out.extract("strain").mean().components[2], end, rtol=1e-9)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_evp_strain_bcs_golden(precision):
    print(f"evp strain bcs {precision}", flush=True)
    rate = Order2SymmetricTensor.from_strain_voigt(np.array([-0.35, -0.35, 1.0, 0, 0, 0]))
    ls = LoadSchedule.from_constant_rates(rate, Order2SymmetricTensor.zero(), np.zeros(6))
    model = SmallStrainFFT(load_schedule=ls, end_time=4.0e-3, initial_time_increment=1.0e-3,
                           constitutive_model=evp(perfect_plasticity, None), precision=precision)
    out = model(voronoi_material(), global_tolerance=1.0e-7)
    assert_allclose(mean_norm(out), 428.328875, rtol=rtol_for(precision))
    assert_allclose(out.extract("stress").mean().components[2], 426.923816, rtol=rtol_for(precision))
    assert_allclose(out.extract("strain").mean().components, rate.components * 0.004, atol=1e-14)


def test_fields_match_oracle_fp64():
    """Full stress/strain fields and EVP state against the oracle on the same input."""
    from oracle import constitutive as C
    from oracle import tensors as T
    from oracle.solver import LoadSchedule as OLS
    from oracle.solver import SmallStrainFFT as OSolver

    print("fields vs oracle", flush=True)
    material = voronoi_material()
    rot = material.extract("orientation").rotation_matrix
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    model = SmallStrainFFT(load_schedule=ls, end_time=4e-3, initial_time_increment=1e-3,
                           constitutive_model=evp(voce, {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}))
    out = model(material, global_tolerance=1e-7, output_variables=["slip_resistances", "plastic_strains"])
    schmid = T.schmid_mandel(*T.octahedral_slip_systems())
    omodel = C.ElasticViscoplastic(T.stiffness_from_voigt(TI), schmid, 1.0, 10.0, 300.0, "voce",
                                   {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0})
    oracle = OSolver(OLS.from_constant_uniaxial_strain_rate(direction="z"), end_time=4e-3,
                     initial_time_increment=1e-3).run([8, 8, 8], rot, omodel, global_tolerance=1e-7)
    s, so = out.extract("stress").components, oracle["stress"]
    assert np.abs(s - so).max() <= 1e-9 * np.abs(so).max()
    assert np.abs(out.extract("strain").components - oracle["strain"]).max() <= 1e-9 * np.abs(oracle["strain"]).max()
    g = out.extract("slip_resistances").components
    assert np.abs(g - oracle["slip_resistances"]).max() <= 1e-9 * np.abs(oracle["slip_resistances"]).max()
    ep = out.extract("plastic_strains").components
    assert np.abs(ep - oracle["plastic_strains"]).max() <= 1e-9 * np.abs(oracle["plastic_strains"]).max()
