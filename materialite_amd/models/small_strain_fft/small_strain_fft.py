"""SmallStrainFFT: Galerkin Newton-Krylov FFT solver on MI355X.

Drop-in for ``materialite/models/small_strain_fft/small_strain_fft.py:29-398``:
same constructor and ``run`` keywords, same time stepping (adaptive increment,
cutback, output times, post-processor), same logger name and messages.  Each
load increment (predictor, Newton loop, MINRES corrections, convergence tests)
runs on the device through ``mtl_solver_step``; this module keeps the
reference's host-side control flow and the ``Material`` round trip.

Extra keyword (constructor): ``precision`` = ``"fp64"`` (default; every field,
FFT and Krylov vector in FP64, reproducing the reference), ``"refined"`` (FP32
tangent and MINRES w / x vectors; matvec spectrum and Lanczos vectors stored as
an FP32 value plus a bf16 remainder, FFT arithmetic in FP64; FP64 strain/stress,
residual FFT and norms: follows the FP64 form's iterates to ~1e-8) or
``"mixed"`` (FP32 Krylov vectors, tangents, matvec FFT storage and arithmetic;
the fastest form, a few 1e-6 to 1e-5 from the FP64 form).
"""

import logging
from collections import defaultdict

import numpy as np

from ..._native import NativeSolver, StepParams
from ...tensor import Order2SymmetricTensor, Scalar
from .elastic import Elastic
from .elastic_viscoplastic import ElasticViscoplastic
from .isotropic_elastic_plastic import IsotropicElasticPlastic
from .multiphase import Multiphase


class SmallStrainFFT:
    def __init__(self, load_schedule, end_time=1, initial_time_increment=None,
                 max_time_increment=None, min_time_increment=1.0e-9, start_time=0.0,
                 constitutive_model=None, precision="fp64"):
        self.load_schedule = load_schedule
        self.end_time = end_time
        self.initial_time_increment = (
            self.end_time if initial_time_increment is None else initial_time_increment)
        self.max_time_increment = (
            self.initial_time_increment if max_time_increment is None else max_time_increment)
        self.min_time_increment = min_time_increment
        self.start_time = start_time
        self._constitutive_model = constitutive_model
        self.precision = precision
        self._logger = logging.getLogger("SmallStrainFFT")
        self._sizes = None
        self._dimensions = None
        self._num_points = None
        self.stats = {}

    def __call__(self, material, **kwargs):
        return self.run(material, **kwargs)

    # ------------------------------------------------------------ setup
    def _get_constitutive_model(self, material, phase_label, constitutive_model):
        """small_strain_fft.py:384-398 -> (models, model id per point)."""
        if phase_label is not None and constitutive_model is not None:
            raise ValueError("cannot specify phase and constitutive model")
        if phase_label is None and constitutive_model is None:
            raise ValueError("must specify phase label or provide a constitutive model")
        n = self._num_points
        if phase_label is None:
            if isinstance(constitutive_model, Multiphase):
                # a Multiphase built by the caller (multiphase.py:18-29): its
                # models on the points of their index lists
                model_of_point = np.full(n, -1, np.int32)
                for k, idx in enumerate(constitutive_model.indices):
                    model_of_point[np.asarray(idx, int)] = k
                if np.any(model_of_point < 0):
                    raise ValueError("every point needs a constitutive model")
                return list(constitutive_model.models), model_of_point
            return [constitutive_model], np.zeros(n, np.int32)
        table = material.extract_regional_field(phase_label)
        phases = table[phase_label].to_list()
        models = table["constitutive_model"].to_list()
        indices = material.get_region_indices(region_label=phase_label)
        model_of_point = np.full(n, -1, np.int32)
        for k, p in enumerate(phases):
            model_of_point[np.asarray(indices.get(p, []), int)] = k
        if np.any(model_of_point < 0):
            raise ValueError("every point's phase needs a constitutive model")
        self._multiphase = Multiphase(phases, models, [indices.get(p, []) for p in phases], n)
        return models, model_of_point

    @staticmethod
    def _register(native, model):
        if isinstance(model, (Elastic, ElasticViscoplastic, IsotropicElasticPlastic)):
            return model._add_to(native)
        raise TypeError(
            f"constitutive model {type(model).__name__} has no device implementation "
            "(supported: Elastic, ElasticViscoplastic, IsotropicElasticPlastic, Multiphase of them)")

    def _new_time_increment(self, time, time_increment, next_output_time, max_constit_iters):
        """small_strain_fft.py:266-279."""
        fixed = np.min([next_output_time - time, self.end_time - time])
        eligible = [self.max_time_increment, fixed]
        eligible.append(time_increment * 2.0 if max_constit_iters < 11 else time_increment)
        return np.min(eligible)

    # ------------------------------------------------------------ outputs
    def _outputs(self, native, models, output_variables):
        out = {}
        if output_variables is None:
            return out
        available = []
        for m in models:
            available += m.available_state_variables or []
        for v in output_variables:
            if v not in available:
                raise ValueError(
                    f"output variable {v} is not available in any constitutive models in this simulation")
            data = native.get_field(v)
            if v == "plastic_strains":
                out[v] = Order2SymmetricTensor(data, "p")
            elif v in ("eq_plastic_strains", "yield_stresses"):
                out[v] = Scalar(data[:, 0], "p")
            else:
                out[v] = Scalar(data, "ps")
        return out

    # ------------------------------------------------------------ run
    def run(self, material, orientation_label="orientation", phase_label=None,
            output_variables=None, output_times=None, global_tolerance=1.0e-3,
            strain_correction_tolerance=1.0e-2, linear_solver_tolerance=1.0e-5,
            postprocessor=None, increment_callback=None):
        """small_strain_fft.py:60-264.

        ``increment_callback(info)`` (an extension) is called after every
        converged increment with a dict of the increment's counters; a true
        return value ends the run there.
        """
        self._sizes = material.sizes
        self._dimensions = material.dimensions
        self._num_points = material.num_points
        time = self.start_time
        time_tolerance = self.min_time_increment / 2.0
        if output_times is not None:
            time_idx = 0
            output_times = np.append(output_times, np.inf)
            next_output_time = output_times[time_idx]
        else:
            next_output_time = np.inf
        new_state = defaultdict(list)

        orientations = material.extract(orientation_label)
        rotations = np.broadcast_to(orientations.rotation_matrix,
                                    (self._num_points, 3, 3))
        models, model_of_point = self._get_constitutive_model(
            material, phase_label, self._constitutive_model)

        native = NativeSolver(self._dimensions, self.precision)
        try:
            for m in models:
                self._register(native, m)
            native.set_microstructure(rotations, model_of_point)
            native.initialize()

            ls = self.load_schedule
            mask = np.asarray(ls.stress_mask, float)
            old_max_strain_increment = 0.0
            summed_von_mises_stress = 0.0
            time_step_id = 0
            time_increment = self.initial_time_increment
            krylov = newton = 0
            floored = False
            # every attempted increment: (time, dt, converged, Newton iterations,
            # Krylov iterations, constitutive failure); kept up to date so that
            # it survives a "min time increment reached" error
            history = []
            self.stats = {"history": history}
            while time < (self.end_time - time_tolerance):
                self._logger.info(f"New increment: time {time} + {time_increment}")
                d_strain = ls.strain_increment(time, time_increment).components
                d_stress = ls.stress_increment(time, time_increment).components
                applied = ls.stress(time, time_increment).components
                max_strain_increment = d_strain[np.argmax(np.abs(d_strain))]
                p = StepParams()
                p.strain_increment[:] = [float(v) for v in d_strain]
                p.stress_increment[:] = [float(v) for v in d_stress]
                p.applied_stress[:] = [float(v) for v in applied]
                p.stress_mask[:] = [float(v) for v in mask]
                p.time_increment = float(time_increment)
                p.global_tolerance = float(global_tolerance)
                p.strain_correction_tolerance = float(strain_correction_tolerance)
                p.linear_solver_tolerance = float(linear_solver_tolerance)
                p.summed_von_mises_stress = float(summed_von_mises_stress)
                p.time_step_id = int(time_step_id)
                p.reset_fluctuation = int(max_strain_increment * old_max_strain_increment < 0.0)
                rep = native.step(p)
                if rep.linear_solver_tolerance_used > linear_solver_tolerance and not floored:
                    floored = True
                    self._logger.warning(
                        f"precision {self.precision!r}: linear_solver_tolerance "
                        f"{linear_solver_tolerance:g} raised to {rep.linear_solver_tolerance_used:g}")
                krylov += rep.krylov_iterations
                newton += rep.newton_iterations
                history.append((float(time), float(time_increment), bool(rep.converged),
                                int(rep.newton_iterations), int(rep.krylov_iterations),
                                bool(rep.constitutive_failed)))
                self._logger.debug(f"global iterations {rep.newton_iterations}, "
                                   f"linear solver iterations {rep.krylov_iterations}, "
                                   f"equilibrium error {rep.equilibrium_error}")
                if not rep.converged and not rep.constitutive_failed:
                    self._logger.info("Too many iterations in macroloading loop")
                if rep.converged:
                    summed_von_mises_stress += rep.mean_von_mises_stress
                    old_max_strain_increment = max_strain_increment
                    time_step_id += 1
                    time += time_increment
                    # the reference collects outputs, advances the output time and then
                    # sizes the next increment (small_strain_fft.py:221-236); the device
                    # commit (state, and fluctuation scaled by the new increment) comes
                    # first here, which leaves the collected fields unchanged
                    output_now = time >= (next_output_time - time_tolerance)
                    upcoming = output_times[time_idx + 1] if output_now else next_output_time
                    new_time_increment = self._new_time_increment(
                        time, time_increment, upcoming, rep.max_constitutive_iterations)
                    native.commit(new_time_increment / time_increment)
                    if output_now:
                        outputs = self._outputs(native, models, output_variables)
                        outputs.update({"stress": Order2SymmetricTensor(native.get_field("stress"), "p"),
                                        "strain": Order2SymmetricTensor(native.get_field("strain"), "p")})
                        if postprocessor is not None:
                            outputs = postprocessor(outputs)
                        for k, v in outputs.items():
                            new_state[k].append(v)
                        time_idx += 1
                        next_output_time = output_times[time_idx]
                    time_increment = new_time_increment
                    if increment_callback is not None and increment_callback({
                            "step": time_step_id, "time": time,
                            "krylov_iterations": int(rep.krylov_iterations),
                            "newton_iterations": int(rep.newton_iterations),
                            "constitutive_iterations": int(rep.max_constitutive_iterations),
                            "num_points": self._num_points}):
                        break
                else:
                    time_increment = time_increment / 2.0
                    native.rollback()
                    if time_increment < self.min_time_increment:
                        raise ValueError("min time increment reached")

            outputs = self._outputs(native, models, output_variables)
            outputs.update({"stress": Order2SymmetricTensor(native.get_field("stress"), "p"),
                            "strain": Order2SymmetricTensor(native.get_field("strain"), "p")})
            self.stats = {"krylov_iterations": int(krylov), "newton_iterations": int(newton),
                          "steps": int(time_step_id), "history": history,
                          "linear_solver_tolerance_used": max(
                              float(linear_solver_tolerance),
                              1.0e-5 if self.precision == "mixed" else 0.0)}
        finally:
            native.close()

        if output_times is None:
            return material.create_fields(outputs)
        if postprocessor is None:
            for k, v in new_state.items():
                new_state[k] = type(v[0]).from_stack(v, new_dim="t", axis=len(v[0].dims_str))
            return material.create_fields(dict(new_state))
        new_material = material.create_fields(outputs)
        new_material.state.update(new_state)
        return new_material
