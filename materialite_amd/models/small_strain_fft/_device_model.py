"""The constitutive-model protocol of the reference, run on a device point handle.

``initialize``, ``calculate_stress_and_tangent``, ``update_state_variables``,
``generate_outputs`` and ``postprocess`` (``small_strain_fft/elastic.py:23-46``,
``elastic_viscoplastic.py:52-294``, ``isotropic_elastic_plastic.py:55-192``):
``initialize`` creates a handle over the given orientations
(``mtl_points_create``) and registers the model on it; the update runs in the
same HIP kernels as inside ``SmallStrainFFT``.  State variables the device
holds are read from it when they are looked up in ``state_variables``.
There is no CPU fallback: without the library or a HIP device these raise.

``SmallStrainFFT`` does not go through this protocol; it registers the model
on its own grid handle, where the state stays on the device for the run.
"""

import numpy as np

from ..._native import NativePoints
from ...tensor import Order2SymmetricTensor, Order4SymmetricTensor, Scalar


class DeviceStateVariables(dict):
    """A model's ``state_variables``: host entries as in a dict, plus names
    whose values are read from the device when looked up."""

    def __init__(self, fetch, device_names):
        super().__init__()
        self._fetch = fetch
        self._device = tuple(device_names)

    def __getitem__(self, key):
        if key in self._device:
            return self._fetch(key)
        return super().__getitem__(key)

    def get(self, key, default=None):
        if key in self._device:
            return self._fetch(key)
        return super().get(key, default)

    def __contains__(self, key):
        return key in self._device or super().__contains__(key)

    def keys(self):
        return list(super().keys()) + [k for k in self._device if not super().__contains__(k)]


def _rotations(orientations):
    r = np.asarray(orientations.rotation_matrix, dtype=np.float64)
    return r.reshape(-1, 3, 3)


class DeviceModel:
    """Protocol mixin.  A model class provides ``_add_to(native)`` (registers
    itself on a handle, returns the model id) and ``_device_fields`` (state
    variable name -> tensor type and dims of the device field)."""

    _device_fields = {}

    def initialize(self, orientations):
        rot = _rotations(orientations)
        n = rot.shape[0]
        self._close_points()
        points = NativePoints(n, "fp64")
        try:
            self._add_to(points)
            points.set_microstructure(rot, np.zeros(n, np.int32))
            points.initialize()
        except BaseException:
            points.close()
            raise
        self._points = points
        self.state_variables = DeviceStateVariables(self._fetch, self._device_fields)
        self.state_variables["orientations"] = orientations
        return self._tangent()

    def calculate_stress_and_tangent(self, strain, guess_stress, time_increment):
        points = self._require_points()
        stress, iterations = points.constitutive_host(
            np.asarray(strain.components).reshape(-1, 6),
            np.asarray(guess_stress.components).reshape(-1, 6), float(time_increment))
        if not iterations:
            return None, None, False
        return Order2SymmetricTensor(stress, "p"), self._tangent(), iterations

    def update_state_variables(self):
        self._require_points().commit(1.0)

    def generate_outputs(self, output_variables=None):
        if output_variables is None:
            return dict()
        return {k: self.state_variables.get(k, None) for k in output_variables}

    def postprocess(self, output_variables=None):
        outputs = self.generate_outputs(output_variables)
        self.state_variables = dict()
        self._close_points()
        return outputs

    # ------------------------------------------------------------ helpers
    def _require_points(self):
        points = getattr(self, "_points", None)
        if points is None:
            raise RuntimeError(f"{type(self).__name__}: call initialize(orientations) first")
        return points

    def _close_points(self):
        points = getattr(self, "_points", None)
        if points is not None:
            points.close()
            self._points = None

    def _tangent(self):
        points = self._require_points()
        return Order4SymmetricTensor(points.get_field("tangent").reshape(-1, 6, 6), "p")

    def _fetch(self, name):
        tensor_type, dims = self._device_fields[name]
        data = self._require_points().get_field(name)
        if tensor_type is Scalar and dims == "p":
            data = data[:, 0]
        return tensor_type(data, dims)

    def __del__(self):
        try:
            self._close_points()
        except Exception:
            pass
