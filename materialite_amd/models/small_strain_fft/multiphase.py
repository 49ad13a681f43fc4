"""Per-phase model assignment (``small_strain_fft/multiphase.py:18-95``).

Inside ``SmallStrainFFT`` every point gets a model id and the device evaluates
each model on its own point subset, failing the whole update if any model
fails.  Called directly, this class runs the reference's protocol over its
member models (each on its own device point handle) and scatters their
results to the points of their phases.
"""

import numpy as np

from ...tensor import Order2SymmetricTensor, Order4SymmetricTensor


class Multiphase:
    def __init__(self, phases, models, indices, num_points):
        self.phases = phases
        self.models = models
        self.indices = indices
        self.num_points = num_points
        self.available_state_variables = []
        for m in self.models:
            if m.available_state_variables is not None:
                self.available_state_variables += m.available_state_variables
        self._output_variable_info = None

    def initialize(self, orientations):
        """multiphase.py:31-36."""
        tangent = np.zeros((self.num_points, 6, 6))
        for m, i in zip(self.models, self.indices):
            tangent[np.asarray(i, int)] = m.initialize(orientations[np.asarray(i, int)]).components
        return Order4SymmetricTensor(tangent, "p")

    def calculate_stress_and_tangent(self, strain, guess_stress, time_increment):
        """multiphase.py:38-51: (None, None, False) as soon as a model fails."""
        stress = np.zeros((self.num_points, 6))
        tangent = np.zeros((self.num_points, 6, 6))
        iterations = []
        for m, i in zip(self.models, self.indices):
            i = np.asarray(i, int)
            s_i, t_i, converged = m.calculate_stress_and_tangent(strain[i], guess_stress[i],
                                                                 time_increment)
            if not converged:
                return None, None, converged
            stress[i] = s_i.components
            tangent[i] = t_i.components
            iterations.append(converged)
        return (Order2SymmetricTensor(stress, "p"), Order4SymmetricTensor(tangent, "p"),
                np.max(iterations))

    def update_state_variables(self):
        for m in self.models:
            m.update_state_variables()

    def postprocess(self, output_variables):
        outputs = self.generate_outputs(output_variables)
        for m in self.models:
            m.state_variables = dict()
            close = getattr(m, "_close_points", None)
            if close is not None:
                close()
        return outputs

    def generate_outputs(self, output_variables=None):
        """multiphase.py:64-95: each variable at the points of the models that
        have it, zero elsewhere; a variable no model has is an error."""
        outputs = dict()
        if output_variables is None:
            return outputs
        if self._output_variable_info is None:
            self._output_variable_info = dict()
            for v in output_variables:
                if v not in self.available_state_variables:
                    raise ValueError(
                        f"output variable {v} is not available in any constitutive models in this "
                        "simulation")
                width, type_v = 0, None
                for m in self.models:
                    type_vm, width_vm = m._state_variable_info.get(v, (None, None))
                    if type_vm is None:
                        continue
                    width = max(width, width_vm)
                    type_v = type_vm
                shape = [self.num_points, width] if width > 0 else self.num_points
                self._output_variable_info[v] = (type_v, shape)
        for v in output_variables:
            type_v, shape = self._output_variable_info[v]
            data = type_v.zero().repeat(shape)
            for m, i in zip(self.models, self.indices):
                value = m.state_variables.get(v, None)
                if value is None:
                    continue
                i = np.asarray(i, int)
                if data.components[i].shape == value.components.shape:
                    data.components[i] = value.components
                else:
                    data.components[i, : value.components.shape[-1]] = value.components
            outputs[v] = data
        return outputs
