"""Unit pins of the oracle: known answers, closed forms and invariants."""

import numpy as np
import pytest
import scipy.sparse.linalg as spla
from numpy.testing import assert_allclose, assert_array_equal

from oracle import constitutive as C
from oracle import microstructure as M
from oracle import projection as P
from oracle import tensors as T
from oracle.minres import minres


# ------------------------------------------------------------------ tensors
def test_directional_modulus_known_answer():
    # tests/test_tensor.py:214-224, 1438-1445
    mtex = np.array(
        [
            [320.5, 68.2, 71.6, 0, 0, 0],
            [68.2, 196.5, 76.8, 0, 0, 0],
            [71.6, 76.8, 233.5, 0, 0, 0],
            [0, 0, 0, 64, 0, 0],
            [0, 0, 0, 0, 77, 0],
            [0, 0, 0, 0, 0, 78.7],
        ]
    )
    assert_allclose(T.directional_modulus(T.stiffness_from_voigt(mtex), [1, 1, 1]), 183.0467131980609)


def test_voigt_to_mandel_scaling():
    c = T.cubic_stiffness(250.0, 150.0, 120.0)
    expected = np.diag([250.0] * 3 + [240.0] * 3)
    expected[:3, :3] += 150.0 * (1 - np.eye(3))
    assert_allclose(c, expected, atol=1e-12)


def test_random_orientation_known_answer():
    # tests/test_orientation.py:108-111 and :133-139 (default_rng(12345))
    e = T.random_euler(2, np.random.default_rng(12345))
    assert_allclose(e[:, :2], [[1.42839436, 0.93386543], [1.99025135, 1.2105451]], rtol=1e-7)
    assert_allclose(e[:, 2], [2.45741378, 2.09113158], rtol=1e-7)


def test_mandel_rotation_matches_cartesian_fourth_order():
    rng = np.random.default_rng(3)
    r = T.rotation_from_euler(T.random_euler(5, rng))
    c = T.stiffness_from_voigt(np.array(
        [[5.0, 1, 2, 0.1, 0.2, 0.3], [1, 6, 1.5, 0.4, 0.1, 0.2], [2, 1.5, 7, 0.3, 0.1, 0.5],
         [0.1, 0.4, 0.3, 2, 0.1, 0.2], [0.2, 0.1, 0.1, 0.1, 3, 0.3], [0.3, 0.2, 0.5, 0.2, 0.3, 4]]))
    c4 = T.mandel4_to_cartesian(c)
    # C_specimen_abcd = R_ia R_jb R_kc R_ld C_crystal_ijkl
    spec4 = np.einsum("pia,pjb,pkc,pld,ijkl->pabcd", r, r, r, r, c4)
    got = T.order4_to_specimen(np.broadcast_to(c, (5, 6, 6)), T.rotation_matrix_mandel(r))
    assert_allclose(got, T.mandel4_from_cartesian(spec4), atol=1e-12)
    # 2nd order: crystal -> specimen round trip
    v = rng.normal(size=(5, 6))
    assert_allclose(T.order2_to_specimen(T.order2_to_crystal(v, r), r), v, atol=1e-13)


def test_reference_rotation_matrix_mandel_table():
    # tensor.py:1668-1700 spot entries
    r = T.rotation_from_euler([0.3, 1.1, -0.7])
    rm = T.rotation_matrix_mandel(r)
    r2 = np.sqrt(2)
    assert_allclose(rm[0, 3], r2 * r[0, 1] * r[0, 2])
    assert_allclose(rm[3, 0], r2 * r[1, 0] * r[2, 0])
    assert_allclose(rm[3, 4], r[1, 0] * r[2, 2] + r[1, 2] * r[2, 0])
    assert_allclose(rm[5, 5], r[0, 0] * r[1, 1] + r[0, 1] * r[1, 0])


def test_octahedral_systems():
    # tests/test_slip_system.py: every direction lies in its plane, unit vectors
    n, d = T.octahedral_slip_systems()
    assert_allclose(np.einsum("si,si->s", n, d), 0, atol=1e-15)
    p = T.schmid_mandel(n, d)
    assert_allclose(p[:, :3].sum(axis=1), 0, atol=1e-15)  # traceless
    assert_allclose(np.linalg.norm(p, axis=1), np.sqrt(0.5))


# ------------------------------------------------------------------ microstructure
def test_reference_voronoi_known_answers():
    # tests/test_material.py:801-822
    coords, sizes, _ = M.grid_coordinates([2, 3, 4], origin=[1, 1, 1])
    got = M.reference_voronoi(coords, sizes, np.ones(3), 3, np.random.default_rng(12345))
    assert_array_equal(got, [1, 1, 2, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 1, 1, 2, 0, 1, 1, 1, 0])
    got = M.reference_voronoi(coords, sizes, np.ones(3), 3, np.random.default_rng(12345), periodic=True)
    assert_array_equal(got, [0, 1, 2, 0, 0, 1, 0, 0, 0, 1, 2, 0, 0, 1, 2, 0, 0, 1, 0, 0, 0, 1, 2, 0])


# ------------------------------------------------------------------ projection
@pytest.mark.parametrize("dims", [(3, 4, 5), (4, 4, 4), (2, 3, 6), (5, 5, 5), (1, 4, 3), (6, 2, 1)])
@pytest.mark.parametrize("mask", [np.zeros(6), np.array([0, 1, 1, 1, 1, 1.0])])
def test_closed_form_equals_literal_build(dims, mask):
    assert_allclose(P.projection_closed_form(dims, mask), P.projection_literal(dims, mask), atol=2e-15)


@pytest.mark.parametrize("dims", [(4, 6, 8), (3, 5, 7)])
def test_projection_invariants(dims):
    g = P.projection_closed_form(dims, np.zeros(6)).reshape(-1, 6, 6)
    assert_allclose(g, np.swapaxes(g, -1, -2), atol=1e-15)  # symmetric
    assert_allclose(g @ g, g, atol=1e-14)  # idempotent
    tr = np.trace(g, axis1=-2, axis2=-1).reshape(dims)
    nyq = np.zeros(dims, bool)
    for ax, n in enumerate(dims):
        if n % 2 == 0:
            sl = [slice(None)] * 3
            sl[ax] = n // 2
            nyq[tuple(sl)] = True
    nyq[0, 0, 0] = True
    assert_allclose(tr[~nyq], 3.0, atol=1e-13)
    assert_allclose(tr[nyq], 0.0, atol=1e-15)
    # even in k
    g3 = g.reshape(*dims, 6, 6)
    flipped = g3[(-np.arange(dims[0])) % dims[0]][:, (-np.arange(dims[1])) % dims[1]][:, :, (-np.arange(dims[2])) % dims[2]]
    assert_allclose(flipped, g3, atol=1e-15)


def test_apply_projection_brute_force_dft_and_mask():
    dims = (3, 4, 5)
    rng = np.random.default_rng(1)
    f = rng.normal(size=(60, 6))
    mask = np.array([1, 1, 0, 0, 0, 0.0])
    ghat = P.projection_literal(dims, mask)
    # brute-force 3-D DFT matrices
    dft = [np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n) for n in dims]
    grid = f.reshape(*dims, 6)
    spec = np.einsum("ax,by,cz,xyzi->abci", *dft, grid)
    prod = np.einsum("xyzij,xyzj->xyzi", ghat, spec)
    back = np.einsum("ax,by,cz,abci->xyzi", *[d.conj() for d in dft], prod).real / 60
    assert_allclose(P.apply_projection(f, dims, ghat=ghat), back.reshape(-1, 6), atol=1e-13)
    assert_allclose(P.apply_projection(f, dims, stress_mask=mask), back.reshape(-1, 6), atol=1e-13)
    # a homogeneous field projects to its masked mean only
    h = np.tile(rng.normal(size=6), (60, 1))
    assert_allclose(P.apply_projection(h, dims, stress_mask=mask), h * mask, atol=1e-14)


# ------------------------------------------------------------------ minres
@pytest.mark.parametrize("rtol", [1e-5, 1e-10])
def test_minres_matches_scipy_iterates(rtol):
    rng = np.random.default_rng(7)
    a = rng.normal(size=(40, 40))
    a = a + a.T + 5 * np.eye(40)
    a[:, -3:] = 0
    a[-3:, :] = 0  # singular but consistent, like zero-stiffness pores
    b = a @ rng.normal(size=40)
    mine, ref = [], []
    x, itn, _ = minres(lambda v: a @ v, b, rtol=rtol, callback=lambda xk: mine.append(xk.copy()))
    xs, info = spla.minres(spla.aslinearoperator(a), b, rtol=rtol, callback=lambda xk: ref.append(xk.copy()))
    assert itn == len(ref)
    assert_allclose(np.array(mine), np.array(ref), rtol=0, atol=1e-13)
    assert_allclose(x, xs, atol=1e-13)


# ------------------------------------------------------------------ hardening closed forms
def test_voce_integrates_the_voce_law():
    # tests/test_hardening_laws.py: g(G) = g0 + (tau1 + theta1 G)(1 - exp(-G theta0/tau1))
    props = {"tau_1": 100.0, "theta_0": 500.0, "theta_1": 7.0}
    g0, acc, g = 50.0, np.zeros(1), np.full((1, 3), 50.0)
    rates = np.array([[0.002, -0.001, 0.0005]])
    for _ in range(20):
        g = C.voce(props, acc, g, rates, 0.5)
        acc = acc + 0.5 * np.abs(rates).sum()
    gam = acc[0]
    expected = g0 + (100.0 + 7.0 * gam) * (1 - np.exp(-gam * 500.0 / 100.0))
    assert_allclose(g, expected, rtol=1e-12)


def test_linear_and_af_increments():
    rates = np.array([[0.002, -0.001]])
    assert_allclose(C.linear({"hardening_rate": 10}, np.zeros(1), np.full((1, 2), 5.0), rates, 2.0), 5.0 + 10 * 0.006)
    d = 0.006
    assert_allclose(
        C.armstrong_frederick({"direct_hardening": 100.0, "dynamic_recovery": 2.0}, np.zeros(1), np.full((1, 2), 5.0), rates, 2.0),
        5.0 + (100.0 - 2.0 * 5.0) * d / (1 + 2.0 * d),
    )


def test_hcp_slip_systems_known_answers():
    # tests/test_slip_system.py:62-200 (c/a = 1.5 for the pyramidal families)
    r3 = np.sqrt(3)
    r3b2 = r3 / 2
    n, d = T.hcp_slip_systems(["basal"])
    np.testing.assert_allclose(n, [[0, 0, 1]] * 3, atol=1e-15)
    np.testing.assert_allclose(d, [[1, 0, 0], [-0.5, r3b2, 0], [-0.5, -r3b2, 0]], atol=1e-15)
    n, d = T.hcp_slip_systems(["prismatic"])
    np.testing.assert_allclose(n, [[r3b2, 0.5, 0], [0, -1, 0], [-r3b2, 0.5, 0]], atol=1e-15)
    np.testing.assert_allclose(d, [[-0.5, r3b2, 0], [1, 0, 0], [-0.5, -r3b2, 0]], atol=1e-15)
    a_over_c = 2 / 3
    rows = [[1, 1 / r3, a_over_c], [0, -2 / r3, a_over_c], [-1, 1 / r3, a_over_c],
            [-1, -1 / r3, a_over_c], [0, 2 / r3, a_over_c], [1, -1 / r3, a_over_c]]
    n, d = T.hcp_slip_systems(["pyramidal_a"], c_over_a=1.5)
    np.testing.assert_allclose(n, np.array(rows) * 0.75, atol=1e-15)
    np.testing.assert_allclose(d, [[-0.5, r3b2, 0], [1, 0, 0], [-0.5, -r3b2, 0],
                                   [-0.5, r3b2, 0], [1, 0, 0], [0.5, r3b2, 0]], atol=1e-15)
    n, d = T.hcp_slip_systems(["pyramidal_ca"], c_over_a=1.5)
    np.testing.assert_allclose(n, np.repeat(np.array(rows), 2, axis=0) * 0.75, atol=1e-15)
    ca = 1.5
    dirs = np.array([[-1.5, -3 * r3b2, 3 * ca], [-3, 0, 3 * ca], [1.5, 3 * r3b2, 3 * ca],
                     [-1.5, 3 * r3b2, 3 * ca], [3, 0, 3 * ca], [1.5, -3 * r3b2, 3 * ca],
                     [3, 0, 3 * ca], [1.5, 3 * r3b2, 3 * ca], [-1.5, -3 * r3b2, 3 * ca],
                     [1.5, -3 * r3b2, 3 * ca], [-3, 0, 3 * ca], [-1.5, 3 * r3b2, 3 * ca]])
    np.testing.assert_allclose(d, dirs * 2 / (3 * np.sqrt(13)), atol=1e-15)
    # maximum Schmid factors along [111] (tests/test_slip_system.py:148-190)
    l111 = [1, 1, 1]
    assert np.isclose(T.max_schmid_factor(*T.octahedral_slip_systems(), l111), 2 / (3 * np.sqrt(6)))
    assert np.isclose(T.max_schmid_factor(*T.hcp_slip_systems(["basal"]), l111), (1 + r3) / 6)
    assert np.isclose(T.max_schmid_factor(*T.hcp_slip_systems(["prismatic"]), l111), 1 / 3)
    assert np.isclose(T.max_schmid_factor(*T.hcp_slip_systems(["pyramidal_a"], 1.5), l111),
                      (1 + r3) / 6)
    assert np.isclose(T.max_schmid_factor(*T.hcp_slip_systems(["pyramidal_ca"], 1.5), l111),
                      (7 + 5 * r3) / (12 * np.sqrt(13)))


def test_oracle_periodic_shift_gives_shifted_stress():
    """tests/test_elastic_model.py:225-270: a microstructure shifted
    periodically gives the same stress field, shifted.  The oracle's solver on
    a 6^3 grid of 8 grains (random orientations), one uniaxial-z increment,
    shifted by (1, 3, 2) voxels."""
    from oracle import constitutive as C
    from oracle import tensors as T
    from oracle.solver import LoadSchedule as OLS
    from oracle.solver import SmallStrainFFT as OSolver
    from tests import fixtures as F

    n = 6
    rng = np.random.default_rng(21)
    grains = rng.integers(0, 8, size=(n, n, n))
    rot_g = T.rotation_from_euler(T.random_euler(8, rng))

    def solve(g):
        rot = rot_g[g.ravel()]
        out = OSolver(OLS.from_constant_uniaxial_strain_rate(direction="z"), end_time=1.0,
                      initial_time_increment=1.0).run([n] * 3, rot, C.Elastic(F.TI_STIFFNESS),
                                                      global_tolerance=1e-8,
                                                      linear_solver_tolerance=1e-10)
        return out["stress"].reshape(n, n, n, 6)

    shift = (1, 3, 2)
    base = solve(grains)
    moved = solve(np.roll(grains, shift, axis=(0, 1, 2)))
    want = np.roll(base, shift, axis=(0, 1, 2))
    assert np.abs(moved - want).max() <= 1e-9 * np.abs(base).max()
    assert np.abs(base - base.mean(axis=(0, 1, 2))).max() > 1e-3 * np.abs(base).max()
