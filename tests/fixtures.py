"""Seeded inputs of the reference's own SmallStrainFFT tests, rebuilt with the oracle.

Cases follow ``tests/test_elastic_model.py`` and
``tests/test_elastic_viscoplastic_model.py`` of the reference.
"""

import numpy as np

from oracle import constitutive as C
from oracle import microstructure as M
from oracle import tensors as T

# tests/test_elastic_model.py:66-79 (transversely isotropic stiffness, MPa)
TI_VOIGT = np.array(
    [
        [160e3, 90e3, 66e3, 0, 0, 0],
        [90e3, 160e3, 66e3, 0, 0, 0],
        [66e3, 66e3, 181.7e3, 0, 0, 0],
        [0, 0, 0, 46.5e3, 0, 0],
        [0, 0, 0, 0, 46.5e3, 0],
        [0, 0, 0, 0, 0, 35e3],
    ]
)
TI_STIFFNESS = T.stiffness_from_voigt(TI_VOIGT)

# tests/test_elastic_viscoplastic_model.py:62-112
EVP_CASES = {
    "linear": ({"hardening_rate": 10}, 4e-3, 361.276725, True),
    "voce": ({"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}, 5e-3, 541.043871, False),
    "armstrong_frederick": (
        {"direct_hardening": 10000.0, "dynamic_recovery": 15.0}, 5e-3, 539.345343, False,
    ),
    "perfect_plasticity": (None, 4e-3, 442.608112, False),
}


def bicrystal_4():
    """tests/test_elastic_model.py:82-91: 4^3, Euler (0, pi/4, 0) above z = 1.5."""
    coords, sizes, _ = M.grid_coordinates([4, 4, 4])
    rot = np.tile(T.rotation_from_euler([0, np.pi / 4, 0]), (64, 1, 1))
    rot[M.box_inside(coords, max_corner=[None, None, sizes[2] / 2])] = np.eye(3)
    return rot


def voronoi_8():
    """tests/test_elastic_model.py:100-123: 8^3 over sizes 3, 10 Voronoi regions,
    default_rng(0), unit-radius spherical pore at the centre (phase 2)."""
    rng = np.random.default_rng(0)
    coords, sizes, _ = M.grid_coordinates([8, 8, 8], sizes=[3, 3, 3])
    regions = M.reference_voronoi(coords, sizes, np.zeros(3), 10, rng)
    rot = M.random_region_rotations(regions, rng)
    pore = M.sphere_inside(coords, 1, [1.5, 1.5, 1.5])
    return rot, pore, regions


def evp_model(hardening, props, rate_exponent=10.0):
    schmid = T.schmid_mandel(*T.octahedral_slip_systems())
    return C.ElasticViscoplastic(TI_STIFFNESS, schmid, 1.0, rate_exponent, 300.0, hardening, props)


# J2 plasticity cases of tests/test_isotropic_plasticity.py:19-140 (8^3, identity
# orientation, sizes 3, uniaxial z, dt 1e-4): (hardening, properties, steps, golden)
J2_CASES = {
    "perfect_plasticity": ("perfect_plasticity", None, 15, 150.0),
    "linear": ("linear", {"hardening_rate": 1000.0}, 15, 150.3980086),
    "voce": ("voce", {"tau_1": 30.0, "theta_0": 10000.0, "theta_1": 100.0}, 50, 171.694205),
}
J2_MODULUS, J2_SHEAR_MODULUS, J2_YIELD = 150000.0, 60000.0, 150.0


def box_phase_8():
    """Phase of each point of the 8^3 grid with sizes 3 when a Box(max_corner=[2,2,2])
    is inserted (tests/test_multiphase.py:13-38): coordinates i * 3/7 <= 2."""
    i = np.arange(8) * 3.0 / 7.0
    inside = i <= 2.0
    return (inside[:, None, None] & inside[None, :, None] & inside[None, None, :]).ravel()
