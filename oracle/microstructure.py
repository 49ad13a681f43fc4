"""Seeded microstructure fixtures as the reference's tests build them (oracle).

Restates ``Material`` grid coordinates (``material.py:103-131``, ``util.py:40-42``),
``create_voronoi`` (``material.py:337-383``), ``assign_random_orientations``
(``material.py:385-419``) and the ``Sphere`` / ``Box`` features
(``material.py:1115-1135, 1187-1232``), so the end-to-end goldens of the
reference's tests can be regenerated without importing the reference.
"""

import numpy as np
from scipy import spatial

from .tensors import random_euler, rotation_from_euler


def grid_coordinates(dims, sizes=None, spacing=(1, 1, 1), origin=(0, 0, 0)):
    """Point coordinates, x slowest and z fastest (material.py:103-131)."""
    dims = np.asarray(dims)
    if sizes is not None:
        sizes = np.asarray(sizes, float)
        spacing = np.divide(sizes, dims - 1, out=np.zeros_like(sizes), where=dims != 1)
    spacing = np.asarray(spacing, float)
    sizes = spacing * (dims - 1)
    x, y, z = np.mgrid[0 : dims[0], 0 : dims[1], 0 : dims[2]]
    ids = np.c_[x.ravel(), y.ravel(), z.ravel()]
    return ids * spacing + np.asarray(origin, float), sizes, spacing


def _repeat27(data, sx, sy, sz):
    """util.py:57-88 (the 27 periodic images, in the reference's order)."""
    shifts = [(0, 0, 0)]
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dx, dy, dz) != (0, 0, 0):
                    shifts.append((dx, dy, dz))
    return np.concatenate([data + np.array(s) * [sx, sy, sz] for s in shifts], axis=0)


def reference_voronoi(coords, sizes, origin, num_regions, rng, periodic=False):
    """Nearest-seed region ids (material.py:337-383)."""
    seeds = rng.random((num_regions, 3)) * sizes + origin
    if periodic:
        ref = np.tile(np.arange(num_regions), 27)
        tree = spatial.cKDTree(_repeat27(seeds, *sizes))
        _, idx = tree.query(_repeat27(coords, *sizes), k=1)
        return ref[idx[: coords.shape[0]]]
    _, idx = spatial.cKDTree(seeds).query(coords, k=1)
    return np.asarray(idx)


def random_region_rotations(regions, rng):
    """Per-point R from Haar-random per-region orientations (material.py:385-419).

    Regions are numbered in order of first appearance for the draw, as pandas'
    ``unique`` does.
    """
    _, first = np.unique(regions, return_index=True)
    order = np.asarray(regions)[np.sort(first)]
    rot = rotation_from_euler(random_euler(len(order), rng))
    lookup = {r: i for i, r in enumerate(order)}
    return rot[[lookup[r] for r in regions]]


def sphere_inside(coords, radius, centroid):
    """material.py:1134-1135."""
    return np.linalg.norm(coords - np.asarray(centroid, float), axis=-1) <= radius


def box_inside(coords, min_corner=(None,) * 3, max_corner=(None,) * 3):
    """material.py:1227-1232 (None is an open side)."""
    lo = np.array([-np.inf if v is None else v for v in min_corner], float)
    hi = np.array([np.inf if v is None else v for v in max_corner], float)
    return np.all(coords - lo >= 0, axis=-1) & np.all(-(coords - hi) >= 0, axis=-1)
