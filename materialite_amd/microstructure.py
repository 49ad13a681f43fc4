"""Seeded synthetic microstructures for the benchmark workloads.

Periodic Voronoi grains on an n^3 voxel torus (voxel i at coordinate i, period
n, minimum-image distances) with Haar-uniform orientations drawn per grain in
the order z1, cos(x2), z3 from the same generator (tensor.py:1899-1912 of the
reference).  The reference's own ``create_voronoi(periodic=True)`` uses the
period n-1 (material.py:366-377), which is not a torus of the voxel grid; this
generator is the truly periodic variant the FFT solver's periodic cell needs.
"""

import numpy as np
from scipy import spatial

from .tensor import Orientation


def periodic_voronoi(dims, num_grains, rng):
    """Grain id per point (x slowest, z fastest) and the seed points."""
    dims = np.asarray(dims, dtype=float)
    seeds = rng.random((num_grains, 3)) * dims
    tree = spatial.cKDTree(seeds, boxsize=dims)
    ix, iy, iz = (np.arange(int(d), dtype=float) for d in dims)
    grains = np.empty(int(np.prod(dims)), dtype=np.int32)
    plane = int(dims[1] * dims[2])
    yy, zz = np.meshgrid(iy, iz, indexing="ij")
    yz = np.c_[yy.ravel(), zz.ravel()]
    for x in range(int(dims[0])):  # one x-plane at a time keeps memory flat
        pts = np.c_[np.full(plane, ix[x]), yz]
        _, idx = tree.query(pts, k=1)
        grains[x * plane:(x + 1) * plane] = idx
    return grains, seeds


def grain_rotations(num_grains, rng):
    """Haar-random rotation matrices, one per grain (num_grains, 3, 3)."""
    return Orientation.random(num_grains, rng=rng).rotation_matrix.reshape(num_grains, 3, 3)


def periodic_spheres(dims, radius, volume_fraction, rng, max_tries=200000):
    """Random sequential addition of non-overlapping periodic spheres (centre
    distance >= 2 radius on the torus) until the voxel volume fraction is
    reached; voxel i (coordinate i) belongs to a sphere when its periodic
    distance to the centre is <= radius.  Returns (inclusion mask over the
    points, x slowest, z fastest; sphere centres).  Each sphere marks only the
    voxels of its bounding box, so the cost is independent of the grid size."""
    dims = np.asarray(dims, dtype=int)
    n = int(np.prod(dims))
    mask = np.zeros(n, dtype=bool)
    centres = np.empty((0, 3))
    target = volume_fraction * n
    count = 0
    tries = 0
    r_box = int(np.ceil(radius)) + 1
    while count < target and tries < max_tries:
        tries += 1
        c = rng.random(3) * dims
        if len(centres):
            d = np.abs(centres - c)
            d = np.minimum(d, dims - d)
            if np.any(np.linalg.norm(d, axis=1) < 2 * radius):
                continue
        centres = np.vstack([centres, c])
        axes = []
        for k in range(3):
            lo = int(np.floor(c[k])) - r_box
            ii = np.arange(lo, lo + 2 * r_box + 2)
            di = np.abs(ii - c[k])
            di = np.minimum(di, dims[k] - di)
            axes.append((np.mod(ii, dims[k]), di))
        (ix, dx), (iy, dy), (iz, dz) = axes
        inside = (dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2
                  <= radius ** 2)
        flat = ((ix[:, None, None] * dims[1] + iy[None, :, None]) * dims[2]
                + iz[None, None, :])[inside]
        flat = np.unique(flat)
        count += int(np.count_nonzero(~mask[flat]))
        mask[flat] = True
    return mask, centres
