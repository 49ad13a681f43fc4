"""Point-grid material container (API-compatible subset of materialite's Material).

Follows ``materialite/material.py``: points in ``np.mgrid`` order (x slowest, z
fastest, ``util.py:40-42``), coordinates ``origin + id * spacing``
(``material.py:103-131``), per-point fields, per-region fields merged on a
region label (``material.py:200-243, 510-573``), seeded Voronoi regions and
Haar-random orientations (``material.py:337-419``), geometric features
(``material.py:421-470, 1094-1232``).

Tensor-valued fields are kept as whole tensors (one array per field) instead of
one Python object per point, so a 128^3 grid costs one array copy, not two
million objects; ``extract`` returns the same tensor types as the reference.
"""

from copy import deepcopy

import warnings

import numpy as np
import pandas as pd
from scipy import spatial

from .tensor import Order2SymmetricTensor, Order4SymmetricTensor, Orientation, Scalar, Tensor, Vector


def cartesian_grid(dimensions):
    x, y, z = np.mgrid[0 : dimensions[0], 0 : dimensions[1], 0 : dimensions[2]]
    return np.c_[np.ravel(x), np.ravel(y), np.ravel(z)]


def _is_tensorlike(v):
    return isinstance(v, (Tensor, Orientation))


def _take(v, idx):
    if isinstance(v, Orientation):
        return Orientation(v.rotation_matrix[idx], "p")
    if isinstance(v, Tensor):
        return type(v)(v.components[idx], "p" + v.dims_str[1:])
    return np.asarray(v, dtype=object)[idx]


class Material:
    def __init__(self, dimensions=(16, 16, 16), origin=(0, 0, 0), spacing=(1, 1, 1), sizes=None,
                 fields=None):
        self._dimensions = np.array(dimensions)
        self._origin = np.array(origin, dtype=float)
        spacing = np.array(spacing, dtype=float)
        if sizes is not None:
            sizes = np.array(sizes, dtype=float)
            spacing = np.divide(sizes, self._dimensions - 1, out=np.zeros_like(sizes),
                                where=self._dimensions != 1)
            sizes[np.where(self._dimensions == 1)[0]] = 0
        else:
            sizes = spacing * (self._dimensions - 1)
        self._spacing, self._sizes = spacing, sizes
        grid = cartesian_grid(self._dimensions)
        base = pd.DataFrame({
            "x": grid[:, 0] * spacing[0] + self._origin[0],
            "y": grid[:, 1] * spacing[1] + self._origin[1],
            "z": grid[:, 2] * spacing[2] + self._origin[2],
        })
        self._tensors = {}
        if fields is not None:
            fields = pd.DataFrame(fields) if not isinstance(fields, pd.DataFrame) else fields
            if len(fields) != self.num_points:
                raise ValueError("length of fields does not match number of points from dimensions")
            if any(k in fields.columns for k in ("x", "y", "z")):
                fields = fields.sort_values(by=["x", "y", "z"], ignore_index=True)
                base = fields
            else:
                base = pd.concat([base, fields.reset_index(drop=True)], axis=1)
        self.fields = base.assign(x_id=grid[:, 0], y_id=grid[:, 1], z_id=grid[:, 2])
        self.state = dict()
        self._regional_fields = dict()

    # ------------------------------------------------------------ geometry
    @property
    def dimensions(self):
        return self._dimensions.copy()

    @property
    def origin(self):
        return self._origin.copy()

    @property
    def spacing(self):
        return self._spacing.copy()

    @property
    def sizes(self):
        return self._sizes.copy()

    @property
    def num_points(self):
        return int(np.prod(self._dimensions))

    def copy(self):
        return deepcopy(self)

    def run(self, model, **kwargs):
        return model(self, **kwargs)

    # ------------------------------------------------------------ fields
    def create_fields(self, fields):
        material = self.copy()
        scalars = {}
        for k, v in fields.items():
            if _is_tensorlike(v):
                material._tensors[k] = v.copy()
                material.fields = material.fields.drop(columns=[k], errors="ignore")
            else:
                scalars[k] = v
                material._tensors.pop(k, None)
        if scalars:
            material.fields = material.fields.assign(**scalars)
        return material

    def create_uniform_field(self, label, value):
        if isinstance(value, Orientation):
            return self.create_fields({label: value.repeat(self.num_points)})
        if isinstance(value, Tensor):
            return self.create_fields({label: value.repeat(self.num_points)})
        return self.create_fields({label: [value] * self.num_points})

    def create_regional_fields(self, region_label, regional_fields):
        if isinstance(regional_fields, pd.DataFrame):
            regional_fields = {k: regional_fields[k].to_list() for k in regional_fields.columns}
        regional_fields = dict(regional_fields)
        regions = np.asarray(regional_fields[region_label])
        if len(np.unique(regions)) != len(regions):
            raise ValueError(
                f"Could not create regional field: {region_label} has non-unique values in regional field")
        material = self.copy()
        table = material._regional_fields.get(region_label, {region_label: regions})
        order = np.argsort(regions)
        for k, v in regional_fields.items():
            if k == region_label:
                continue
            if isinstance(v, list) and v and isinstance(v[0], Orientation):
                v = Orientation.from_list(v)
            elif isinstance(v, list) and v and isinstance(v[0], Tensor):
                v = type(v[0]).from_list(v)
            table[k] = _take(v, order) if _is_tensorlike(v) else np.asarray(v, dtype=object)[order]
        table[region_label] = regions[order]
        material._regional_fields[region_label] = table
        return material

    def extract_regional_field(self, region_label, field_label=None):
        table = self._regional_fields[region_label]
        if field_label is None:
            return pd.DataFrame({k: (list(v) if _is_tensorlike(v) else list(v))
                                 for k, v in table.items()})
        return table[field_label]

    def _regional_lookup(self, label):
        for region_label, table in self._regional_fields.items():
            if label in table and label != region_label:
                point_regions = self.fields[region_label].to_numpy()
                idx = np.searchsorted(table[region_label], point_regions)
                return _take(table[label], idx)
        return None

    def extract(self, labels):
        if isinstance(labels, str):
            if labels in self._tensors:
                return self._tensors[labels].copy()
            if labels in self.fields.columns:
                data = self.fields[labels].to_list()
                if data and hasattr(type(data[0]), "from_list"):
                    return type(data[0]).from_list(data)
                return np.array(data)
            regional = self._regional_lookup(labels)
            if regional is not None:
                if isinstance(regional, np.ndarray) and regional.dtype == object:
                    try:
                        return regional.astype(float)
                    except (TypeError, ValueError):
                        return regional
                return regional
            raise KeyError(labels)
        return self.fields[list(labels)].to_numpy()

    def get_region_indices(self, region_label="region"):
        regions = self.fields[region_label]
        return {r: idx.tolist() for r, idx in regions.groupby(regions).groups.items()}

    # ------------------------------------------------------------ microstructure
    def create_voronoi(self, num_regions=10, label="region", rng=None, periodic=False):
        """Nearest-seed regions (material.py:337-383)."""
        rng = np.random.default_rng() if rng is None else rng
        seeds = rng.random((num_regions, 3)) * self.sizes + self.origin
        pts = self.fields[["x", "y", "z"]].to_numpy()
        if periodic:
            shifts = [(0, 0, 0)] + [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                                    for dx in (-1, 0, 1) if (dx, dy, dz) != (0, 0, 0)]
            rep = lambda d: np.concatenate([d + np.array(s) * self.sizes for s in shifts])
            ref = np.tile(np.arange(num_regions), 27)
            _, idx = spatial.cKDTree(rep(seeds)).query(rep(pts), k=1)
            regions = ref[idx[: self.num_points]]
        else:
            _, regions = spatial.cKDTree(seeds).query(pts, k=1)
        return self.create_fields({label: np.asarray(regions)})

    def assign_random_orientations(self, region_label="region", orientation_label="orientation",
                                   rng=None):
        """Haar-random orientation per region, drawn in order of first appearance
        (material.py:385-419)."""
        rng = np.random.default_rng() if rng is None else rng
        unique_regions = self.fields[region_label].unique()
        orientations = Orientation.random(len(unique_regions), rng=rng)
        if len(unique_regions) == 1:
            orientations = Orientation(orientations.rotation_matrix.reshape(1, 3, 3), "p")
        return self.create_regional_fields(
            region_label, {region_label: unique_regions, orientation_label: orientations})

    def export_to_evpfft(self, orientation_label="orientation", grain_label="grain",
                         phase_label="phase", output="fields.txt", euler_angles_to_degrees=True):
        """EVP-FFT microstructure file (material.py:1035-1091): one line per point,
        "e1 e2 e3 x y z grain phase" with Bunge angles (%.5f), 1-based voxel ids,
        sorted z slowest, x fastest, space separated."""
        dims = self.dimensions
        if not np.array_equal([power_of_two_below(d) for d in dims], dims):
            warnings.warn("Outputting to EVP-FFT with dimensions that are not powers of 2. Use "
                          "Material.crop_by_range() or Material.crop_by_id_range() if the "
                          "dimensions must be powers of 2.")
        o = self.extract(orientation_label)
        e = np.atleast_2d(o.euler_angles_in_degrees if euler_angles_to_degrees else o.euler_angles)
        table = pd.DataFrame({
            "euler_angle_1": e[:, 0], "euler_angle_2": e[:, 1], "euler_angle_3": e[:, 2],
            "x": self.extract("x_id") + 1, "y": self.extract("y_id") + 1,
            "z": self.extract("z_id") + 1,
            "grain": self.extract(grain_label), "phase": self.extract(phase_label),
        }).sort_values(by=["z", "y", "x"])
        table.to_csv(output, header=False, sep=" ", index=False, float_format="%.5f")

    def insert_feature(self, feature, fields):
        """Set field values at points inside a feature (material.py:421-470)."""
        pts = self.fields[["x", "y", "z"]].to_numpy()
        inside = feature.check_inside(pts)
        new = {}
        for label, value in fields.items():
            if label in self._tensors or (label not in self.fields.columns and _is_tensorlike(value)):
                if label in self._tensors:
                    cur = self._tensors[label].copy()
                elif isinstance(value, Orientation):
                    cur = Orientation.identity().repeat(self.num_points)
                else:
                    cur = type(value).zero().repeat(self.num_points)
                if isinstance(cur, Orientation):
                    cur.rotation_matrix[inside] = value.rotation_matrix
                else:
                    cur.components[inside] = value.components
                new[label] = cur
            else:
                cur = (self.fields[label].to_numpy().copy() if label in self.fields.columns
                       else np.zeros(self.num_points, dtype=type(value)))
                cur = cur.astype(object) if cur.dtype == object else cur
                cur[inside] = value
                new[label] = cur
        return self.create_fields(new)


class Sphere:
    """material.py:1115-1135."""

    def __init__(self, radius, centroid):
        self.radius = float(radius)
        self.centroid = np.asarray(centroid, dtype=float)

    def check_inside(self, points):
        return np.linalg.norm(np.asarray(points) - self.centroid, axis=-1) <= self.radius


class Box:
    """material.py:1187-1232 (None is an open side)."""

    def __init__(self, min_corner=(None, None, None), max_corner=(None, None, None)):
        self.min_corner = np.array([-np.inf if v is None else v for v in min_corner], float)
        self.max_corner = np.array([np.inf if v is None else v for v in max_corner], float)

    def check_inside(self, points):
        p = np.asarray(points)
        return np.all(p - self.min_corner >= 0, axis=-1) & np.all(-(p - self.max_corner) >= 0, axis=-1)


__all__ = ["Material", "Sphere", "Box", "Vector", "Scalar", "Order2SymmetricTensor",
           "Order4SymmetricTensor", "Orientation"]


# ------------------------------------------------------------------ EVP-FFT text files
def power_of_two_below(n):
    """Largest power of two <= n (material.py, util)."""
    return 1 << (int(n).bit_length() - 1)


def import_evpfft(file, spacing=(1, 1, 1), origin=(0, 0, 0), grain_label="grain",
                  phase_label="phase", orientation_label="orientation",
                  euler_angles_to_radians=True):
    """Material from an EVP-FFT microstructure file (importers.py:153-200):
    dimensions from the largest 1-based voxel ids, Bunge angles in degrees
    (radians when euler_angles_to_radians is False)."""
    table = pd.read_csv(file, header=None, delimiter=" ",
                        names=["e1", "e2", "e3", "x_id", "y_id", "z_id", grain_label,
                               phase_label]).sort_values(by=["x_id", "y_id", "z_id"])
    dims = table[["x_id", "y_id", "z_id"]].max().to_numpy()
    angles = table[["e1", "e2", "e3"]].to_numpy()
    orientations = Orientation.from_euler_angles(angles, in_degrees=euler_angles_to_radians)
    material = Material(dimensions=dims, spacing=spacing, origin=origin)
    if len(table) != material.num_points:
        raise ValueError("EVP-FFT file does not cover the full voxel grid")
    return material.create_fields({
        grain_label: table[grain_label].to_numpy(),
        phase_label: table[phase_label].to_numpy(),
        orientation_label: Orientation(np.reshape(orientations.rotation_matrix, (-1, 3, 3)), "p"),
    })

