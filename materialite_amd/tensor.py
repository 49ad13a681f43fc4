"""Mandel-basis tensors and orientations used on the SmallStrainFFT path.

API-compatible subset of ``materialite/tensor.py``: the classes keep the
reference's names, constructors, ``components`` layout and the operations the
solver's users call (construction from Voigt/cubic/isotropic constants,
``@`` and ``*`` products, ``norm``, ``dev``, ``mean``, frame rotations,
``repeat``, indexing).  Leading (point/slip/time) dimensions are tracked as a
``dims`` string as in the reference; broadcasting here is positional (numpy)
over leading dimensions.
"""

from numbers import Number

import numpy as np

R2 = np.sqrt(2.0)

# Mandel basis (basis_operations.py:37-48)
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
# Voigt bases (basis_operations.py:75-98)
_STRESS_VOIGT = np.array(
    [
        [[1.0, 0, 0], [0, 0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 0, 0], [0, 0, 1.0]],
        [[0, 0, 0], [0, 0, 1.0], [0, 1.0, 0]],
        [[0, 0, 1.0], [0, 0, 0], [1.0, 0, 0]],
        [[0, 1.0, 0], [1.0, 0, 0], [0, 0, 0]],
    ]
)
_STRAIN_VOIGT = np.array(
    [
        [[1.0, 0, 0], [0, 0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 1.0, 0], [0, 0, 0]],
        [[0, 0, 0], [0, 0, 0], [0, 0, 1.0]],
        [[0, 0, 0], [0, 0, 0.5], [0, 0.5, 0]],
        [[0, 0, 0.5], [0, 0, 0], [0.5, 0, 0]],
        [[0, 0.5, 0], [0.5, 0, 0], [0, 0, 0]],
    ]
)


def _default_dims(n):
    return "pst"[:n] if n <= 3 else "".join(chr(ord("a") + i) for i in range(n))


class Tensor:
    _rank = 0  # number of trailing component axes

    def __init__(self, components, dims=None):
        if isinstance(components, Tensor):
            dims = components.dims_str if dims is None else dims
            components = components.components
        c = np.asarray(components, dtype=np.float64)
        nlead = c.ndim - self._rank
        if nlead < 0:
            raise ValueError(f"{type(self).__name__} needs at least {self._rank} component axes")
        self.components = c
        self.dims_str = _default_dims(nlead) if dims is None else dims
        if len(self.dims_str) != nlead:
            raise ValueError(
                f"tried to create {type(self).__name__} with dimensions {self.dims_str} "
                f"but components have shape {c.shape}"
            )

    def _new(self, components, dims=None):
        return type(self)(components, self.dims_str if dims is None else dims)

    @property
    def shape(self):
        return self.components.shape[: len(self.dims_str)]

    def __len__(self):
        if not self.dims_str:
            raise TypeError("tensor without leading dimensions has no len()")
        return self.components.shape[0]

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def __getitem__(self, key):
        c = self.components[key]
        nlead = c.ndim - self._rank
        dims = self.dims_str[len(self.dims_str) - nlead:] if nlead > 0 else ""
        return type(self)(c, dims)

    def __setitem__(self, key, item):
        self.components[key] = item.components if isinstance(item, Tensor) else item

    def copy(self):
        return type(self)(self.components.copy(), self.dims_str)

    def __neg__(self):
        return self._new(-self.components)

    def _dims_for(self, other):
        a, b = self.dims_str, other.dims_str
        return a if len(a) >= len(b) else b

    def __add__(self, other):
        if isinstance(other, Number):
            return self._new(self.components + other)
        if type(other) is not type(self):
            return NotImplemented
        return type(self)(self.components + other.components, self._dims_for(other))

    __radd__ = __add__

    def __sub__(self, other):
        return self + (-other)

    def __truediv__(self, other):
        if isinstance(other, Number):
            return self._new(self.components / other)
        if isinstance(other, Scalar):
            return self._new(self.components / other.components.reshape(
                other.components.shape + (1,) * self._rank))
        return NotImplemented

    def mean(self, dim=None):
        if not self.dims_str:
            return self.copy()
        if dim is None:
            return type(self)(self.components.mean(axis=tuple(range(len(self.dims_str)))), "")
        ax = self.dims_str.index(dim)
        return type(self)(self.components.mean(axis=ax), self.dims_str.replace(dim, ""))

    def sum(self, dim=None):
        if not self.dims_str:
            return self.copy()
        if dim is None:
            return type(self)(self.components.sum(axis=tuple(range(len(self.dims_str)))), "")
        ax = self.dims_str.index(dim)
        return type(self)(self.components.sum(axis=ax), self.dims_str.replace(dim, ""))

    def repeat(self, shape, dims=None):
        shape = tuple(shape) if hasattr(shape, "__iter__") else (shape,)
        comp = np.broadcast_to(self.components, shape + self.components.shape).copy()
        return type(self)(comp, dims if dims is not None else _default_dims(len(shape)))

    @classmethod
    def from_list(cls, tensors):
        return cls(np.array([t.components for t in tensors]))

    @classmethod
    def from_stack(cls, tensors, new_dim, axis=0):
        comp = np.stack([t.components for t in tensors], axis=axis)
        dims = tensors[0].dims_str
        return cls(comp, dims[:axis] + new_dim + dims[axis:])

    def __repr__(self):
        return f"{type(self).__name__}({self.components!r}, dims='{self.dims_str}')"


class Scalar(Tensor):
    _rank = 0

    @classmethod
    def zero(cls):
        return cls(0.0)

    @property
    def abs(self):
        return self._new(np.abs(self.components))

    def max(self, dim=None):
        if dim is None:
            return Scalar(self.components.max(), "")
        ax = self.dims_str.index(dim)
        return Scalar(self.components.max(axis=ax), self.dims_str.replace(dim, ""))

    def __mul__(self, other):
        if isinstance(other, Number):
            return self._new(self.components * other)
        if isinstance(other, Scalar):
            return Scalar(self.components * other.components, self._dims_for(other))
        return NotImplemented

    __rmul__ = __mul__

    def __float__(self):
        return float(self.components)


class Vector(Tensor):
    _rank = 1

    @property
    def norm(self):
        return Scalar(np.linalg.norm(self.components, axis=-1), self.dims_str)

    @property
    def unit(self):
        return self._new(self.components / np.linalg.norm(self.components, axis=-1, keepdims=True))

    def outer(self, other):
        c = np.einsum("...i,...j->...ij", self.components, other.components)
        return Order2SymmetricTensor.from_cartesian(0.5 * (c + np.swapaxes(c, -1, -2)),
                                                    self._dims_for(other))


class Order2SymmetricTensor(Tensor):
    """Symmetric 2nd-order tensor as a Mandel 6-vector (tensor.py:1128-1381)."""

    _rank = 1

    @classmethod
    def zero(cls):
        return cls(np.zeros(6))

    @classmethod
    def identity(cls):
        return cls.from_cartesian(np.eye(3))

    @classmethod
    def from_cartesian(cls, matrices, dims=None):
        return cls(np.einsum("nij,...ij->...n", MANDEL_BASIS, np.asarray(matrices, float)), dims)

    @classmethod
    def from_mandel(cls, components, dims=None):
        return cls(components, dims)

    @classmethod
    def from_strain_voigt(cls, vectors, dims=None):
        cart = np.einsum("nij,...n->...ij", _STRAIN_VOIGT, np.asarray(vectors, float))
        return cls.from_cartesian(cart, dims)

    @classmethod
    def from_stress_voigt(cls, vectors, dims=None):
        cart = np.einsum("nij,...n->...ij", _STRESS_VOIGT, np.asarray(vectors, float))
        return cls.from_cartesian(cart, dims)

    @property
    def cartesian(self):
        return np.einsum("nij,...n->...ij", MANDEL_BASIS, self.components)

    @property
    def mandel(self):
        return self.components

    @property
    def stress_voigt(self):
        c = self.cartesian
        return np.stack([c[..., 0, 0], c[..., 1, 1], c[..., 2, 2], c[..., 1, 2], c[..., 0, 2],
                         c[..., 0, 1]], axis=-1)

    @property
    def strain_voigt(self):
        c = self.cartesian
        return np.stack([c[..., 0, 0], c[..., 1, 1], c[..., 2, 2], 2 * c[..., 1, 2],
                         2 * c[..., 0, 2], 2 * c[..., 0, 1]], axis=-1)

    @property
    def norm(self):
        return Scalar(np.sqrt(np.einsum("...n,...n", self.components, self.components)),
                      self.dims_str)

    @property
    def trace(self):
        return Scalar(self.components[..., :3].sum(axis=-1), self.dims_str)

    @property
    def dev(self):
        c = self.components.copy()
        c[..., :3] -= c[..., :3].sum(axis=-1, keepdims=True) / 3.0
        return self._new(c)

    def __mul__(self, other):
        if isinstance(other, Number):
            return self._new(other * self.components)
        if isinstance(other, Scalar):
            return self._new(self.components * other.components[..., None],
                             self._dims_for(other))
        if isinstance(other, Order2SymmetricTensor):
            return Scalar(np.einsum("...n,...n->...", self.components, other.components),
                          self._dims_for(other))
        return NotImplemented

    __rmul__ = __mul__

    def outer(self, other):
        return Order4SymmetricTensor(
            np.einsum("...m,...n->...mn", self.components, other.components),
            self._dims_for(other))

    def to_crystal_frame(self, orientations):
        rm = orientations.rotation_matrix_mandel
        return Order2SymmetricTensor(np.einsum("...ij,...j->...i", rm, self.components),
                                     self._dims_for(orientations))

    def to_specimen_frame(self, orientations):
        rm = orientations.rotation_matrix_mandel
        return Order2SymmetricTensor(np.einsum("...ji,...j->...i", rm, self.components),
                                     self._dims_for(orientations))


class Order4SymmetricTensor(Tensor):
    """Minor/major-symmetric 4th-order tensor as a Mandel 6x6 (tensor.py:1384-1640)."""

    _rank = 2

    @classmethod
    def zero(cls):
        return cls(np.zeros((6, 6)))

    @classmethod
    def identity(cls):
        return cls(np.eye(6))

    @classmethod
    def from_mandel(cls, components, dims=None):
        return cls(components, dims)

    @classmethod
    def from_cartesian(cls, array_4d, dims=None):
        return cls(np.einsum("mij,nkl,...ijkl->...mn", MANDEL_BASIS, MANDEL_BASIS,
                             np.asarray(array_4d, float)), dims)

    @classmethod
    def from_voigt(cls, matrix, dims=None):
        c4 = np.einsum("...mn,mij,nkl->...ijkl", np.asarray(matrix, float), _STRESS_VOIGT,
                       _STRESS_VOIGT)
        return cls.from_cartesian(c4, dims)

    @classmethod
    def from_cubic_constants(cls, C11, C12, C44):
        v = np.array([[C11, C12, C12, 0, 0, 0], [C12, C11, C12, 0, 0, 0],
                      [C12, C12, C11, 0, 0, 0], [0, 0, 0, C44, 0, 0], [0, 0, 0, 0, C44, 0],
                      [0, 0, 0, 0, 0, C44]], float)
        return cls.from_voigt(v)

    @classmethod
    def from_transverse_isotropic_constants(cls, C11, C12, C13, C33, C44):
        C66 = (C11 - C12) / 2.0
        v = np.array([[C11, C12, C13, 0, 0, 0], [C12, C11, C13, 0, 0, 0],
                      [C13, C13, C33, 0, 0, 0], [0, 0, 0, C44, 0, 0], [0, 0, 0, 0, C44, 0],
                      [0, 0, 0, 0, 0, C66]], float)
        return cls.from_voigt(v)

    @classmethod
    def from_isotropic_constants(cls, modulus, shear_modulus):
        E, G = modulus, shear_modulus
        L = G * (E - 2 * G) / (3 * G - E)
        v = np.array([[L + 2 * G, L, L, 0, 0, 0], [L, L + 2 * G, L, 0, 0, 0],
                      [L, L, L + 2 * G, 0, 0, 0], [0, 0, 0, G, 0, 0], [0, 0, 0, 0, G, 0],
                      [0, 0, 0, 0, 0, G]], float)
        return cls.from_voigt(v)

    @property
    def mandel(self):
        return self.components

    @property
    def cartesian(self):
        return np.einsum("mij,nkl,...mn->...ijkl", MANDEL_BASIS, MANDEL_BASIS, self.components)

    @property
    def voigt(self):
        return np.einsum("...ijkl,mij,nkl->...mn", self.cartesian, _STRAIN_VOIGT, _STRAIN_VOIGT)

    @property
    def inverse(self):
        return self._new(np.linalg.inv(self.components))

    @property
    def inv(self):
        return self.inverse

    @property
    def norm(self):
        return Scalar(np.sqrt(np.einsum("...mn,...mn", self.components, self.components)),
                      self.dims_str)

    def __mul__(self, other):
        if isinstance(other, Number):
            return self._new(other * self.components)
        if isinstance(other, Scalar):
            return self._new(self.components * other.components[..., None, None],
                             self._dims_for(other))
        return NotImplemented

    __rmul__ = __mul__

    def __matmul__(self, other):
        if isinstance(other, Order2SymmetricTensor):
            return Order2SymmetricTensor(
                np.einsum("...mn,...n->...m", self.components, other.components),
                self._dims_for(other))
        if isinstance(other, Order4SymmetricTensor):
            return Order4SymmetricTensor(
                np.einsum("...mn,...no->...mo", self.components, other.components),
                self._dims_for(other))
        return NotImplemented

    def directional_modulus(self, direction):
        d = Vector(direction).unit
        dd = d.outer(d)
        return Scalar(1.0 / np.einsum("...m,...mn,...n->...", dd.components,
                                      np.linalg.inv(self.components), dd.components))

    def to_specimen_frame(self, orientations):
        rm = orientations.rotation_matrix_mandel
        return Order4SymmetricTensor(np.einsum("...ia,...jb,...ij->...ab", rm, rm,
                                               self.components),
                                     self._dims_for(orientations))

    def to_crystal_frame(self, orientations):
        rm = orientations.rotation_matrix_mandel
        return Order4SymmetricTensor(np.einsum("...ai,...bj,...ij->...ab", rm, rm,
                                               self.components),
                                     self._dims_for(orientations))


class Orientation:
    """Rotation R with v_crystal = R v_specimen (tensor.py:1643-2032)."""

    def __init__(self, rotation_matrix, dims=None):
        if isinstance(rotation_matrix, Orientation):
            self.rotation_matrix = rotation_matrix.rotation_matrix.copy()
            self.dims_str = rotation_matrix.dims_str
            return
        r = np.asarray(rotation_matrix, dtype=np.float64)
        self.rotation_matrix = r
        self.dims_str = _default_dims(r.ndim - 2) if dims is None else dims

    @property
    def shape(self):
        return self.rotation_matrix.shape[:-2]

    def __len__(self):
        return self.rotation_matrix.shape[0]

    def __getitem__(self, key):
        r = self.rotation_matrix[key]
        return Orientation(r, self.dims_str[len(self.dims_str) - (r.ndim - 2):])

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def copy(self):
        return Orientation(self.rotation_matrix.copy(), self.dims_str)

    @classmethod
    def identity(cls):
        return cls(np.eye(3))

    @classmethod
    def from_rotation_matrix(cls, rotation_matrix, dims=None):
        return cls(rotation_matrix, dims)

    @classmethod
    def from_euler_angles(cls, euler_angles, in_degrees=False, dims=None):
        """Bunge (z1, x2, z3) (tensor.py:1855-1896)."""
        e = np.asarray(euler_angles, dtype=np.float64)
        if in_degrees:
            e = e * np.pi / 180.0
        z1, x2, z3 = e[..., 0], e[..., 1], e[..., 2]
        c1, c2, c3 = np.cos(z1), np.cos(x2), np.cos(z3)
        s1, s2, s3 = np.sin(z1), np.sin(x2), np.sin(z3)
        r = np.empty(e.shape[:-1] + (3, 3))
        r[..., 0, 0] = c1 * c3 - c2 * s1 * s3
        r[..., 0, 1] = c3 * s1 + c1 * c2 * s3
        r[..., 0, 2] = s2 * s3
        r[..., 1, 0] = -c1 * s3 - c2 * c3 * s1
        r[..., 1, 1] = c1 * c2 * c3 - s1 * s3
        r[..., 1, 2] = c3 * s2
        r[..., 2, 0] = s1 * s2
        r[..., 2, 1] = -c1 * s2
        r[..., 2, 2] = c2
        return cls(np.squeeze(r), dims)

    @classmethod
    def random(cls, shape=1, rng=None, dims=None):
        """Haar-uniform; draw order z1, cos(x2), z3 (tensor.py:1899-1912)."""
        rng = np.random.default_rng() if rng is None else rng
        shape = tuple(shape) if hasattr(shape, "__iter__") else (shape,)
        z1 = rng.random(shape) * 2.0 * np.pi
        x2 = np.arccos(rng.random(shape) * 2.0 - 1.0)
        z3 = rng.random(shape) * 2.0 * np.pi
        e = np.stack([z1, x2, z3], axis=-1)
        return cls.from_euler_angles(e, dims=dims)

    @classmethod
    def from_list(cls, orientations):
        return cls(np.array([o.rotation_matrix for o in orientations]))

    @property
    def euler_angles(self):
        """Bunge (z1, x2, z3) in radians (tensor.py:1919-1956, D. Eberly's
        "Euler Angle Formulas"): the general case for -1 < R22 < 1, and
        x2 = 0 or pi with z3 = 0 on the two degenerate sides."""
        r = self.rotation_matrix
        r22 = r[..., 2, 2]
        general = np.logical_and(r22 < 1.0, r22 > -1.0)
        with np.errstate(invalid="ignore"):
            out = np.stack([np.arctan2(r[..., 2, 0], -r[..., 2, 1]), np.arccos(r22),
                            np.arctan2(r[..., 0, 2], r[..., 1, 2])], axis=-1)
        if np.all(general):
            return np.squeeze(out)
        zeros = np.zeros(r22.shape)
        lower = np.stack([np.arctan2(r[..., 1, 0], r[..., 0, 0]), np.pi * np.ones(r22.shape),
                          zeros], axis=-1)
        upper = np.stack([np.arctan2(-r[..., 1, 0], r[..., 0, 0]), zeros, zeros], axis=-1)
        out = np.where(general[..., None], out, np.where((r22 >= 1.0)[..., None], upper, lower))
        return np.squeeze(out)

    @property
    def euler_angles_in_degrees(self):
        return self.euler_angles * 180.0 / np.pi

    @property
    def rotation_matrix_mandel(self):
        """(R.A.R^T) in Mandel = Rm @ A (tensor.py:1668-1700)."""
        return np.einsum("mij,...ik,...jl,nkl->...mn", MANDEL_BASIS, self.rotation_matrix,
                         self.rotation_matrix, MANDEL_BASIS)

    def repeat(self, shape, dims=None):
        shape = tuple(shape) if hasattr(shape, "__iter__") else (shape,)
        r = np.broadcast_to(self.rotation_matrix, shape + self.rotation_matrix.shape).copy()
        return Orientation(r, dims if dims is not None else _default_dims(len(shape)))

    def __repr__(self):
        return f"Orientation({self.rotation_matrix!r})"
