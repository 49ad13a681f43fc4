"""Per-phase model assignment (``small_strain_fft/multiphase.py:18-95``).

The solver gives every point a model id; the device evaluates each model on
its own point subset and fails the whole update if any model fails, as the
reference's dispatcher does.
"""


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
