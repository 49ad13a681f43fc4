"""Slip systems (``materialite/slip_system.py``): FCC octahedral and the HCP
basal, prismatic, pyramidal <a> and pyramidal <c+a> families, Miller-Bravais
to Cartesian conversion, Schmid tensors, concatenation and maximum Schmid
factors."""

import numpy as np

from .tensor import Order2SymmetricTensor, Scalar, Vector


def _v(rows):
    return Vector(np.array(rows, float), "s")


class SlipSystem:
    def __init__(self, normals, directions):
        self.normal = normals.unit
        self.direction = directions.unit

    def __len__(self):
        return len(self.normal)

    def __repr__(self):
        return ("SlipSystem(normals: " + str(np.round(self.normal.components, 3))
                + ",\n directions: " + str(np.round(self.direction.components, 3)) + ")")

    @classmethod
    def octahedral(cls):
        """FCC {111}<110> (slip_system.py:35-65)."""
        normals = [[1, 1, -1]] * 3 + [[1, -1, -1]] * 3 + [[1, -1, 1]] * 3 + [[1, 1, 1]] * 3
        directions = [
            [0, 1, 1], [1, 0, 1], [1, -1, 0],
            [0, 1, -1], [1, 0, 1], [1, 1, 0],
            [0, 1, 1], [1, 0, -1], [1, 1, 0],
            [0, 1, -1], [1, 0, -1], [1, -1, 0],
        ]
        return cls(_v(normals), _v(directions))

    @classmethod
    def basal(cls):
        """HCP (0001)<11-20>, 3 systems (slip_system.py:67-72)."""
        normals = hexagonal_planes_to_cartesian([[0, 0, 0, 1]] * 3)
        directions = hexagonal_directions_to_cartesian([[2, -1, -1, 0], [-1, 2, -1, 0],
                                                        [-1, -1, 2, 0]])
        return cls(_v(normals), _v(directions))

    @classmethod
    def prismatic(cls):
        """HCP {10-10}<11-20>, 3 systems (slip_system.py:74-80)."""
        normals = hexagonal_planes_to_cartesian([[1, 0, -1, 0], [0, -1, 1, 0], [-1, 1, 0, 0]])
        directions = hexagonal_directions_to_cartesian([[-1, 2, -1, 0], [2, -1, -1, 0],
                                                        [-1, -1, 2, 0]])
        return cls(_v(normals), _v(directions))

    @classmethod
    def pyramidal_a(cls, c_over_a=1.633):
        """HCP {10-11}<11-20>, 6 systems (slip_system.py:82-103)."""
        normals = hexagonal_planes_to_cartesian(
            [[1, 0, -1, 1], [0, -1, 1, 1], [-1, 1, 0, 1], [-1, 0, 1, 1], [0, 1, -1, 1],
             [1, -1, 0, 1]], c_over_a)
        directions = hexagonal_directions_to_cartesian(
            [[-1, 2, -1, 0], [2, -1, -1, 0], [-1, -1, 2, 0], [-1, 2, -1, 0], [2, -1, -1, 0],
             [1, 1, -2, 0]], c_over_a)
        return cls(_v(normals), _v(directions))

    @classmethod
    def pyramidal_ca(cls, c_over_a=1.633):
        """HCP {10-11}<11-23>, 12 systems (slip_system.py:105-137)."""
        planes = [[1, 0, -1, 1], [0, -1, 1, 1], [-1, 1, 0, 1], [-1, 0, 1, 1], [0, 1, -1, 1],
                  [1, -1, 0, 1]]
        normals = hexagonal_planes_to_cartesian([p for p in planes for _ in range(2)], c_over_a)
        directions = hexagonal_directions_to_cartesian(
            [[-1, -1, 2, 3], [-2, 1, 1, 3], [1, 1, -2, 3], [-1, 2, -1, 3], [2, -1, -1, 3],
             [1, -2, 1, 3], [2, -1, -1, 3], [1, 1, -2, 3], [-1, -1, 2, 3], [1, -2, 1, 3],
             [-2, 1, 1, 3], [-1, 2, -1, 3]], c_over_a)
        return cls(_v(normals), _v(directions))

    @property
    def schmid_tensor(self):
        """sym(d (x) n) per system (slip_system.py:139-141, used as .sym)."""
        c = np.einsum("si,sj->sij", self.direction.components, self.normal.components)
        return Order2SymmetricTensor.from_cartesian(0.5 * (c + np.swapaxes(c, -1, -2)), "s")

    def max_schmid_factor(self, load_direction=Vector([0, 0, 1.0])):
        """max_s |P_s : l (x) l| (slip_system.py:143-146)."""
        l = np.asarray(load_direction.unit.components, float)
        p = np.einsum("si,sj->sij", self.direction.components, self.normal.components)
        f = np.abs(np.einsum("sij,...i,...j->...s", p, l, l))
        return Scalar(f.max(axis=-1))

    def concatenate(self, other):
        """slip_system.py:148-156."""
        normals = np.concatenate([self.normal.components, other.normal.components], axis=0)
        directions = np.concatenate([self.direction.components, other.direction.components],
                                    axis=0)
        return SlipSystem(Vector(normals, "s"), Vector(directions, "s"))


def hexagonal_directions_to_cartesian(directions, c_over_a=1.633):
    """Miller-Bravais [uvtw] -> unit Cartesian (slip_system.py:158-164):
    x = 3u/2, y = sqrt(3)(u/2 + v), z = w c/a."""
    d = np.atleast_2d(np.asarray(directions, float))
    xyz = np.c_[d[:, 0] * 3 / 2, (d[:, 0] / 2 + d[:, 1]) * np.sqrt(3), d[:, 3] * c_over_a]
    return xyz / np.linalg.norm(xyz, axis=1, keepdims=True)


def hexagonal_planes_to_cartesian(planes, c_over_a=1.633):
    """Miller-Bravais (hkil) -> unit Cartesian normal (slip_system.py:167-173):
    x = h, y = (h + 2k)/sqrt(3), z = l a/c."""
    p = np.atleast_2d(np.asarray(planes, float))
    xyz = np.c_[p[:, 0], (p[:, 0] + 2 * p[:, 1]) / np.sqrt(3), p[:, 3] / c_over_a]
    return xyz / np.linalg.norm(xyz, axis=1, keepdims=True)
