"""Initial slip resistances given per slip system (elastic_viscoplastic.py:
66-69 adds the model's slip_resistance to a (points, systems) zero Scalar, so
one value per system broadcasts onto every point).

The oracle's per-system path is pinned by what the mathematics fixes: a slip
system whose resistance is effectively infinite never slips, so the model with
systems "switched off" that way gives the same update as the model without
those systems; and equal per-system values give the scalar model's update.
"""

import numpy as np
import pytest

from oracle import constitutive as OC
from oracle import tensors as OT
from tests import fixtures as F

LAWS = [("linear", {"hardening_rate": 50.0}),
        ("voce", {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}),
        ("armstrong_frederick", {"direct_hardening": 600.0, "dynamic_recovery": 2.0}),
        ("perfect_plasticity", None)]


def _run(model, rot, steps=2):
    stiff = model.initialize(rot)
    rng = np.random.default_rng(3)
    strain = np.tile([-0.0012, -0.0012, 0.004, 0, 0, 0], (rot.shape[0], 1))
    strain = strain + 2e-4 * rng.normal(size=strain.shape)
    out = []
    guess = 0.5 * np.einsum("pmn,pn->pm", stiff, strain)
    for k in range(steps):
        s, d, it = model.calculate_stress_and_tangent(strain * (1 + 0.2 * k), guess, 1e-3)
        assert it, "the update should converge"
        model.update_state_variables()
        out.append((s, d, it, model.g.copy(), model.gamma.copy()))
        guess = s
    return out


@pytest.mark.parametrize("law,props", LAWS)
def test_switched_off_systems_equal_fewer_systems(law, props):
    rng = np.random.default_rng(11)
    rot = OT.rotation_from_euler(OT.random_euler(40, rng))
    schmid = OT.schmid_mandel(*OT.octahedral_slip_systems())
    keep = np.array([0, 2, 3, 5, 7, 8, 10])
    g0 = np.full(12, 1e30)
    g0[keep] = 280.0 + 20.0 * np.arange(keep.size)
    full = OC.ElasticViscoplastic(F.TI_STIFFNESS, schmid, 1.0, 10.0, g0, law, props)
    part = OC.ElasticViscoplastic(F.TI_STIFFNESS, schmid[keep], 1.0, 10.0, g0[keep], law, props)
    for (s1, d1, i1, g1, y1), (s2, d2, i2, g2, y2) in zip(_run(full, rot), _run(part, rot)):
        assert i1 == i2
        np.testing.assert_allclose(s1, s2, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(d1, d2, rtol=1e-12, atol=1e-6)
        np.testing.assert_allclose(g1[:, keep], g2, rtol=1e-13)
        assert np.abs(y1[:, np.setdiff1d(np.arange(12), keep)]).max() < 1e-200
        np.testing.assert_allclose(y1[:, keep], y2, rtol=1e-12, atol=1e-18)


def test_equal_values_equal_scalar():
    rng = np.random.default_rng(12)
    rot = OT.rotation_from_euler(OT.random_euler(30, rng))
    schmid = OT.schmid_mandel(*OT.octahedral_slip_systems())
    law, props = LAWS[1]
    a = _run(OC.ElasticViscoplastic(F.TI_STIFFNESS, schmid, 1.0, 10.0, np.full(12, 300.0), law, props), rot)
    b = _run(OC.ElasticViscoplastic(F.TI_STIFFNESS, schmid, 1.0, 10.0, 300.0, law, props), rot)
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            np.testing.assert_array_equal(u, v)


def test_model_class_checks_shape():
    from materialite_amd import Order4SymmetricTensor, Scalar, SlipSystem
    from materialite_amd.models.small_strain_fft import ElasticViscoplastic, voce

    def make(g0):
        return ElasticViscoplastic(Order4SymmetricTensor.from_voigt(F.TI_VOIGT),
                                   SlipSystem.octahedral(), 1.0, 10.0, g0, voce,
                                   {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0})

    assert make(300.0)._initial_slip_resistances() == 300.0
    assert make(Scalar(300.0))._initial_slip_resistances() == 300.0
    per = make(Scalar(np.arange(12.0) + 200, "s"))._initial_slip_resistances()
    np.testing.assert_array_equal(per, np.arange(12.0) + 200)
    with pytest.raises(ValueError, match="expected"):
        make(np.ones(5))._initial_slip_resistances()
    with pytest.raises(NotImplementedError, match="dims"):
        make(Scalar(np.ones(4), "p"))._initial_slip_resistances()
