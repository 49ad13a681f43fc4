"""Linear elastic constitutive model (``small_strain_fft/elastic.py:16-49``).

Evaluated on the device: sigma = Rm^T C Rm : eps per point, inside
``SmallStrainFFT`` on its grid handle, or through the reference's protocol
(``initialize``, ``calculate_stress_and_tangent``, ...) on a point handle.
"""

import numpy as np

from ._device_model import DeviceModel


class Elastic(DeviceModel):
    def __init__(self, stiffness):
        self.stiffness = stiffness
        self.state_variables = dict()
        self.available_state_variables = None
        self._state_variable_info = dict()

    def _add_to(self, native):
        return native.add_elastic(self.stiffness.components)

    def initialize(self, orientations):
        """elastic.py:23-27: the stiffness in the specimen frame of every point."""
        tangent = super().initialize(orientations)
        self.state_variables["stiffnesses"] = tangent
        return tangent

    def __repr__(self):
        return f"Elastic(stiffness={np.round(self.stiffness.voigt, 3)})"
