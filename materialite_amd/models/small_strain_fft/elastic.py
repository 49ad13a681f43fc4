"""Linear elastic constitutive model (``small_strain_fft/elastic.py:16-49``).

Evaluated on the device by the solver: sigma = Rm^T C Rm : eps per point.
"""

import numpy as np


class Elastic:
    def __init__(self, stiffness):
        self.stiffness = stiffness
        self.state_variables = dict()
        self.available_state_variables = None
        self._state_variable_info = dict()

    def __repr__(self):
        return f"Elastic(stiffness={np.round(self.stiffness.voigt, 3)})"
