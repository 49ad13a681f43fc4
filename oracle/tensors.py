"""Mandel-basis tensor algebra, rotations and crystallography (oracle, FP64).

Restates the subset of ``materialite/tensor.py``, ``materialite/basis_operations.py``
and ``materialite/slip_system.py`` that the SmallStrainFFT path uses, on plain
NumPy arrays:

* symmetric 2nd-order tensors are Mandel 6-vectors ``[xx, yy, zz, r2*yz, r2*xz, r2*xy]``
  (``basis_operations.py:37-48``);
* symmetric 4th-order tensors are Mandel 6x6 matrices (``tensor.py:1384-1640``);
* orientations are 3x3 matrices R with ``v_crystal = R v_specimen``
  (``tensor.py:1855-1896``).
"""

import numpy as np

R2 = np.sqrt(2.0)

# basis_operations.py:37-48 (Mandel basis, orthonormal under A:B)
MANDEL_BASIS = np.array(
    [
        [[1.0, 0, 0], [0, 0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 0, 0], [0, 0, 1.0]],
        [[0, 0, 0], [0, 0, 1 / R2], [0, 1 / R2, 0]],
        [[0, 0, 1 / R2], [0, 0, 0], [1 / R2, 0, 0]],
        [[0, 1 / R2, 0], [1 / R2, 0, 0], [0, 0, 0]],
    ]
)

# basis_operations.py:88-98 (stress-Voigt basis: ones on both off-diagonal slots)
STRESS_VOIGT_BASIS = np.array(
    [
        [[1.0, 0, 0], [0, 0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 0, 0], [0, 0, 1.0]],
        [[0, 0, 0], [0, 0, 1.0], [0, 1.0, 0]],
        [[0, 0, 1.0], [0, 0, 0], [1.0, 0, 0]],
        [[0, 1.0, 0], [1.0, 0, 0], [0, 0, 0]],
    ]
)

# basis_operations.py:75-85 (strain-Voigt basis: halves on off-diagonal slots)
STRAIN_VOIGT_BASIS = np.array(
    [
        [[1.0, 0, 0], [0, 0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 0, 0], [0, 0, 1.0]],
        [[0, 0, 0], [0, 0, 0.5], [0, 0.5, 0]],
        [[0, 0, 0.5], [0, 0, 0], [0.5, 0, 0]],
        [[0, 0.5, 0], [0.5, 0, 0], [0, 0, 0]],
    ]
)


def to_mandel(cart):
    """Cartesian (...,3,3) symmetric -> Mandel (...,6)  (tensor.py:1159-1165)."""
    return np.einsum("nij,...ij->...n", MANDEL_BASIS, cart)


def from_mandel(v):
    """Mandel (...,6) -> cartesian (...,3,3)  (tensor.py:1192-1194)."""
    return np.einsum("nij,...n->...ij", MANDEL_BASIS, v)


def mandel4_from_cartesian(c4):
    """(...,3,3,3,3) -> Mandel (...,6,6)  (tensor.py:142-146, 1414-1416)."""
    return np.einsum("mij,nkl,...ijkl->...mn", MANDEL_BASIS, MANDEL_BASIS, c4)


def mandel4_to_cartesian(m):
    return np.einsum("mij,nkl,...mn->...ijkl", MANDEL_BASIS, MANDEL_BASIS, m)


def stiffness_from_voigt(voigt):
    """Voigt 6x6 stiffness -> Mandel 6x6  (tensor.py:1407-1412, 158-164)."""
    c4 = np.einsum(
        "mn,mij,nkl->ijkl", np.asarray(voigt, float), STRESS_VOIGT_BASIS, STRESS_VOIGT_BASIS
    )
    return mandel4_from_cartesian(c4)


def cubic_stiffness(c11, c12, c44):
    """tensor.py:1421-1434."""
    v = np.array(
        [
            [c11, c12, c12, 0, 0, 0],
            [c12, c11, c12, 0, 0, 0],
            [c12, c12, c11, 0, 0, 0],
            [0, 0, 0, c44, 0, 0],
            [0, 0, 0, 0, c44, 0],
            [0, 0, 0, 0, 0, c44],
        ],
        float,
    )
    return stiffness_from_voigt(v)


def isotropic_stiffness(modulus, shear_modulus):
    """tensor.py:1452-1467 (Lame lambda from E and G)."""
    e, g = modulus, shear_modulus
    lam = g * (e - 2 * g) / (3 * g - e)
    v = np.array(
        [
            [lam + 2 * g, lam, lam, 0, 0, 0],
            [lam, lam + 2 * g, lam, 0, 0, 0],
            [lam, lam, lam + 2 * g, 0, 0, 0],
            [0, 0, 0, g, 0, 0],
            [0, 0, 0, 0, g, 0],
            [0, 0, 0, 0, 0, g],
        ],
        float,
    )
    return stiffness_from_voigt(v)


def directional_modulus(c_mandel, direction):
    """1 / (d d : C^-1 : d d)  (tensor.py:1571-1574)."""
    d = np.asarray(direction, float)
    d = d / np.linalg.norm(d)
    dd = to_mandel(np.outer(d, d))
    return 1.0 / (dd @ np.linalg.inv(c_mandel) @ dd)


def rotation_from_euler(euler):
    """Bunge (z1, x2, z3) -> R, v_crystal = R v_specimen  (tensor.py:1855-1896)."""
    e = np.asarray(euler, float)
    z1, x2, z3 = e[..., 0], e[..., 1], e[..., 2]
    c1, c2, c3 = np.cos(z1), np.cos(x2), np.cos(z3)
    s1, s2, s3 = np.sin(z1), np.sin(x2), np.sin(z3)
    r = np.zeros(e.shape[:-1] + (3, 3))
    r[..., 0, 0] = c1 * c3 - c2 * s1 * s3
    r[..., 0, 1] = c3 * s1 + c1 * c2 * s3
    r[..., 0, 2] = s2 * s3
    r[..., 1, 0] = -c1 * s3 - c2 * c3 * s1
    r[..., 1, 1] = c1 * c2 * c3 - s1 * s3
    r[..., 1, 2] = c3 * s2
    r[..., 2, 0] = s1 * s2
    r[..., 2, 1] = -c1 * s2
    r[..., 2, 2] = c2
    return r


def random_euler(n, rng):
    """Haar-uniform Euler angles; draw order z1, cos(x2), z3  (tensor.py:1899-1912)."""
    z1 = rng.random((n,)) * 2.0 * np.pi
    cos_x2 = rng.random((n,)) * 2.0 - 1.0
    z3 = rng.random((n,)) * 2.0 * np.pi
    return np.stack([z1, np.arccos(cos_x2), z3], axis=-1)


def rotation_matrix_mandel(r):
    """Mandel form Rm of R: (R.A.R^T) in Mandel = Rm @ A_mandel  (tensor.py:1668-1700).

    Built here from the basis definition, Rm_mn = B^m_ij R_ik R_jl B^n_kl, which
    equals the reference's explicit table entry by entry.
    """
    return np.einsum("mij,...ik,...jl,nkl->...mn", MANDEL_BASIS, r, r, MANDEL_BASIS)


def order2_to_crystal(v, r):
    """sigma'_mn = R_mi R_nj sigma_ij  (tensor.py:1347-1360)."""
    return to_mandel(np.einsum("...mi,...nj,...ij->...mn", r, r, from_mandel(v)))


def order2_to_specimen(v, r):
    """sigma_mn = R_im R_jn sigma'_ij  (tensor.py:1362-1375)."""
    return to_mandel(np.einsum("...im,...jn,...ij->...mn", r, r, from_mandel(v)))


def order4_to_specimen(c, rm):
    """D_ab = Rm_ia Rm_jb C_ij  (tensor.py:1620-1634)."""
    return np.einsum("...ia,...jb,...ij->...ab", rm, rm, c)


def order4_to_crystal(c, rm):
    """D'_ab = Rm_ai Rm_bj C_ij  (tensor.py:1604-1618)."""
    return np.einsum("...ai,...bj,...ij->...ab", rm, rm, c)


def mandel_dev(v):
    """Deviatoric part of a Mandel vector  (tensor.py:1241-1244)."""
    out = np.array(v, float, copy=True)
    tr = out[..., 0] + out[..., 1] + out[..., 2]
    out[..., :3] -= tr[..., None] / 3.0
    return out


def octahedral_slip_systems():
    """FCC {111}<110> normals and directions, unit length  (slip_system.py:35-65)."""
    normals = np.array(
        [[1, 1, -1]] * 3 + [[1, -1, -1]] * 3 + [[1, -1, 1]] * 3 + [[1, 1, 1]] * 3, float
    )
    directions = np.array(
        [
            [0, 1, 1], [1, 0, 1], [1, -1, 0],
            [0, 1, -1], [1, 0, 1], [1, 1, 0],
            [0, 1, 1], [1, 0, -1], [1, 1, 0],
            [0, 1, -1], [1, 0, -1], [1, -1, 0],
        ],
        float,
    )
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    return normals, directions


def hexagonal_direction(uvtw, c_over_a=1.633):
    """Miller-Bravais direction -> unit Cartesian (slip_system.py:158-164)."""
    d = np.atleast_2d(np.asarray(uvtw, float))
    x = d[:, 0] * 1.5
    y = np.sqrt(3.0) * (0.5 * d[:, 0] + d[:, 1])
    z = c_over_a * d[:, 3]
    v = np.stack([x, y, z], axis=1)
    return v / np.sqrt((v * v).sum(axis=1))[:, None]


def hexagonal_plane(hkil, c_over_a=1.633):
    """Miller-Bravais plane -> unit Cartesian normal (slip_system.py:167-173)."""
    p = np.atleast_2d(np.asarray(hkil, float))
    v = np.stack([p[:, 0], (p[:, 0] + 2.0 * p[:, 1]) / np.sqrt(3.0), p[:, 3] / c_over_a], axis=1)
    return v / np.sqrt((v * v).sum(axis=1))[:, None]


HCP_FAMILIES = {
    # family: (plane indices per system, direction indices per system) (slip_system.py:67-137)
    "basal": ([[0, 0, 0, 1]] * 3, [[2, -1, -1, 0], [-1, 2, -1, 0], [-1, -1, 2, 0]]),
    "prismatic": ([[1, 0, -1, 0], [0, -1, 1, 0], [-1, 1, 0, 0]],
                  [[-1, 2, -1, 0], [2, -1, -1, 0], [-1, -1, 2, 0]]),
    "pyramidal_a": ([[1, 0, -1, 1], [0, -1, 1, 1], [-1, 1, 0, 1], [-1, 0, 1, 1],
                     [0, 1, -1, 1], [1, -1, 0, 1]],
                    [[-1, 2, -1, 0], [2, -1, -1, 0], [-1, -1, 2, 0], [-1, 2, -1, 0],
                     [2, -1, -1, 0], [1, 1, -2, 0]]),
    "pyramidal_ca": ([h for h in [[1, 0, -1, 1], [0, -1, 1, 1], [-1, 1, 0, 1], [-1, 0, 1, 1],
                                  [0, 1, -1, 1], [1, -1, 0, 1]] for _ in (0, 1)],
                     [[-1, -1, 2, 3], [-2, 1, 1, 3], [1, 1, -2, 3], [-1, 2, -1, 3],
                      [2, -1, -1, 3], [1, -2, 1, 3], [2, -1, -1, 3], [1, 1, -2, 3],
                      [-1, -1, 2, 3], [1, -2, 1, 3], [-2, 1, 1, 3], [-1, 2, -1, 3]]),
}


def hcp_slip_systems(families, c_over_a=1.633):
    """Unit normals and directions of the concatenated HCP families."""
    normals = np.concatenate([hexagonal_plane(HCP_FAMILIES[f][0], c_over_a) for f in families])
    directions = np.concatenate(
        [hexagonal_direction(HCP_FAMILIES[f][1], c_over_a) for f in families])
    return normals, directions


def max_schmid_factor(normals, directions, load):
    """max_s |(d_s (x) n_s) : l (x) l| for a unit load direction l."""
    l = np.asarray(load, float) / np.linalg.norm(load)
    return np.max(np.abs((directions @ l) * (normals @ l)))


def schmid_mandel(normals, directions):
    """sym(d (x) n) per system as Mandel 6-vectors  (slip_system.py:139-141)."""
    p = np.einsum("si,sj->sij", directions, normals)
    return to_mandel(0.5 * (p + np.swapaxes(p, -1, -2)))
