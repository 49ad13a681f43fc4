"""CPU tests of the product's host layer and the C-ABI library (no device calls)."""

import re
import os

import numpy as np
import pandas as pd
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from materialite_amd import Material, Order2SymmetricTensor, Order4SymmetricTensor, Orientation, SlipSystem, Sphere
from materialite_amd import _native
from materialite_amd.models.small_strain_fft import (
    Elastic, ElasticViscoplastic, LoadSchedule, SmallStrainFFT, armstrong_frederick, linear, voce,
)
from oracle import constitutive as OC
from oracle import tensors as T
from tests import fixtures as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "materialite_amd.h")).read()
    declared = set(re.findall(r"\b(mtl_\w+)\s*\(", header))
    lib = _native.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_native.EXPORTS), declared ^ set(_native.EXPORTS)


def test_step_struct_layout_matches_header():
    import ctypes
    assert ctypes.sizeof(_native.StepParams) == 4 * 6 * 8 + 5 * 8 + 2 * 4
    assert ctypes.sizeof(_native.StepReport) == 3 * 4 + 4 + 8 + 4 + 4 + 5 * 8


def test_material_fixture_matches_oracle_voronoi():
    rot, pore, regions = F.voronoi_8()
    rng = np.random.default_rng(0)
    m = (Material(dimensions=[8, 8, 8], sizes=[3, 3, 3]).create_voronoi(num_regions=10, rng=rng)
         .assign_random_orientations(rng=rng).create_uniform_field("phase", 1)
         .insert_feature(Sphere(radius=1, centroid=[1.5, 1.5, 1.5]), fields={"phase": 2}))
    assert_array_equal(m.extract("region"), regions)
    assert_allclose(m.extract("orientation").rotation_matrix, rot, atol=0)
    assert_array_equal(m.extract("phase") == 2, pore)


def test_material_voronoi_known_answer():
    # tests/test_material.py:801-822 of the reference
    m = Material(dimensions=[2, 3, 4], origin=[1, 1, 1]).create_voronoi(
        num_regions=3, rng=np.random.default_rng(12345))
    assert_array_equal(m.fields.region.to_numpy(),
                       [1, 1, 2, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 1, 1, 2, 0, 1, 1, 1, 0])


def test_tensor_constructors_match_oracle():
    assert_allclose(Order4SymmetricTensor.from_voigt(F.TI_VOIGT).components, F.TI_STIFFNESS)
    assert_allclose(Order4SymmetricTensor.from_cubic_constants(250, 150, 120).components,
                    T.cubic_stiffness(250, 150, 120))
    assert_allclose(Order4SymmetricTensor.from_isotropic_constants(200.0, 80.0).components,
                    T.isotropic_stiffness(200.0, 80.0))
    o = Orientation.random(5, rng=np.random.default_rng(3))
    assert_allclose(o.rotation_matrix, T.rotation_from_euler(T.random_euler(5, np.random.default_rng(3))))
    assert_allclose(o.rotation_matrix_mandel, T.rotation_matrix_mandel(o.rotation_matrix))
    p = SlipSystem.octahedral().schmid_tensor.components
    assert_allclose(p, T.schmid_mandel(*T.octahedral_slip_systems()), atol=1e-15)
    c = Order4SymmetricTensor.from_voigt(F.TI_VOIGT)
    assert_allclose(c.directional_modulus([1, 0, 0]).components, T.directional_modulus(F.TI_STIFFNESS, [1, 0, 0]))


def test_load_schedule_increments():
    # tests/test_load_schedule.py of the reference: masked rates and increments
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x")
    assert_allclose(ls.strain_increment(0.0, 0.5).components, [5e-4, 0, 0, 0, 0, 0], atol=1e-14)
    assert_allclose(ls.stress_mask, [0, 1, 1, 1, 1, 1])
    rate = Order2SymmetricTensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    ls = LoadSchedule.from_constant_rates(rate, rate * 10, np.array([1, 0, 1, 0, 1, 0]))
    assert_allclose(ls.strain_increment(1.0, 2.0).components, [0, 4, 0, 8, 0, 12], atol=1e-14)
    assert_allclose(ls.stress_increment(1.0, 2.0).components, [20, 0, 60, 0, 100, 0], atol=1e-14)
    assert_allclose(ls.stress(1.0, 2.0).components, [30, 0, 90, 0, 150, 0], atol=1e-14)
    with pytest.raises(ValueError):
        LoadSchedule(None, None, None, np.array([0, 2, 0, 0, 0, 0]))


def test_hardening_functions_match_oracle():
    rates = np.array([[0.002, -0.001, 0.0005]])
    g = np.full((1, 3), 50.0)
    acc = np.array([0.01])
    props = {"tau_1": 100.0, "theta_0": 500.0, "theta_1": 7.0}
    assert_allclose(voce(props, acc, g, rates, 0.5), OC.voce(props, acc, g, rates, 0.5))
    assert_allclose(linear({"hardening_rate": 3.0}, acc, g, rates, 0.5),
                    OC.linear({"hardening_rate": 3.0}, acc, g, rates, 0.5))
    p = {"direct_hardening": 100.0, "dynamic_recovery": 2.0}
    assert_allclose(armstrong_frederick(p, acc, g, rates, 0.5), OC.armstrong_frederick(p, acc, g, rates, 0.5))


def test_solver_argument_errors():
    m = Material(dimensions=[2, 2, 2]).create_uniform_field("orientation", Orientation.identity())
    ls = LoadSchedule.from_constant_uniaxial_strain_rate()
    with pytest.raises(ValueError, match="must specify"):
        SmallStrainFFT(ls).run(m)
    with pytest.raises(ValueError, match="cannot specify"):
        SmallStrainFFT(ls, constitutive_model=Elastic(Order4SymmetricTensor.identity())).run(m, phase_label="p")
    with pytest.raises(TypeError):
        ElasticViscoplastic(Order4SymmetricTensor.identity(), SlipSystem.octahedral(), 1.0, 10.0, 1.0,
                            lambda *a: None, {})


def test_periodic_spheres_matches_brute_force():
    from materialite_amd.microstructure import periodic_spheres

    dims = np.array([20, 24, 18])
    mask, centres = periodic_spheres(dims, 3.5, 0.1, np.random.default_rng(3))
    g = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    ref = np.zeros(len(g), bool)
    for c in centres:
        d = np.abs(g - c)
        d = np.minimum(d, dims - d)
        ref |= np.linalg.norm(d, axis=1) <= 3.5
    assert np.array_equal(mask, ref)
    assert mask.mean() >= 0.1
    # non-overlapping: periodic centre distances >= 2 r
    for i in range(len(centres)):
        d = np.abs(centres[i + 1:] - centres[i])
        d = np.minimum(d, dims - d)
        assert np.all(np.linalg.norm(d, axis=1) >= 7.0)


def test_hcp_slip_systems_match_oracle():
    for fam, ca in [("basal", 1.633), ("prismatic", 1.633), ("pyramidal_a", 1.5), ("pyramidal_ca", 1.5)]:
        ss = getattr(SlipSystem, fam)() if fam in ("basal", "prismatic") else getattr(SlipSystem, fam)(c_over_a=ca)
        n, d = T.hcp_slip_systems([fam], ca)
        assert_allclose(ss.normal.components, n, atol=1e-15)
        assert_allclose(ss.direction.components, d, atol=1e-15)
    both = SlipSystem.basal().concatenate(SlipSystem.prismatic())
    assert len(both) == 6
    from materialite_amd.tensor import Vector
    assert_allclose(SlipSystem.pyramidal_ca(c_over_a=1.5).max_schmid_factor(Vector([1.0, 1.0, 1.0])).components,
                    (7 + 5 * np.sqrt(3)) / (12 * np.sqrt(13)))
    assert_allclose(SlipSystem.octahedral().max_schmid_factor(Vector([1.0, 0.0, 0.0])).components,
                    1 / np.sqrt(6))


def test_unknown_precision_is_rejected():
    from materialite_amd._native import NativeSolver

    with pytest.raises(ValueError, match="precision must be one of"):
        NativeSolver((4, 4, 4), "fp16")
