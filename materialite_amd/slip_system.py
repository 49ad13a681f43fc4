"""Slip systems (subset of ``materialite/slip_system.py``)."""

import numpy as np

from .tensor import Order2SymmetricTensor, Vector


class SlipSystem:
    def __init__(self, normals, directions):
        self.normal = normals.unit
        self.direction = directions.unit

    def __len__(self):
        return len(self.normal)

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
        return cls(Vector(np.array(normals, float), "s"), Vector(np.array(directions, float), "s"))

    @property
    def schmid_tensor(self):
        """sym(d (x) n) per system (slip_system.py:139-141, used as .sym)."""
        c = np.einsum("si,sj->sij", self.direction.components, self.normal.components)
        return Order2SymmetricTensor.from_cartesian(0.5 * (c + np.swapaxes(c, -1, -2)), "s")
