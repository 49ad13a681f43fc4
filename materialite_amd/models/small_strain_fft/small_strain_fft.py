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
        run = self.start(material, orientation_label=orientation_label, phase_label=phase_label,
                         output_variables=output_variables, output_times=output_times,
                         global_tolerance=global_tolerance,
                         strain_correction_tolerance=strain_correction_tolerance,
                         linear_solver_tolerance=linear_solver_tolerance,
                         postprocessor=postprocessor, increment_callback=increment_callback)
        try:
            while not run.finished:
                run.apply(run.native.step(run.next_params()))
            return run.result()
        finally:
            run.close()

    def start(self, material, **kwargs):
        """The run of ``material`` (the keywords of ``run``) as an object that
        is advanced one load increment at a time: ``next_params()`` gives the
        increment's device parameters and ``apply(report)`` takes the device's
        report of it.  ``run`` drives one such object; ``run_batch`` drives
        several in lockstep so that their MINRES solves share launches."""
        return _Run(self, material, **kwargs)

    @staticmethod
    def run_batch(jobs):
        """Solves several (solver, material, keywords) jobs of one grid size and
        precision together: every round, the unfinished runs each take one load
        increment on their own schedule and time step, and the MINRES solves of
        those increments run as one batch on the device (mtl_solver_step_batch).
        Returns the materials in the order of the jobs; each is the result
        ``solver.run(material, **keywords)`` gives."""
        runs = [solver.start(material, **kw) for solver, material, kw in jobs]
        try:
            drive_batch(runs)
            return [r.result() for r in runs]
        finally:
            for r in runs:
                r.close()


def drive_batch(runs):
    """Advances the given runs (same grid and precision) until all have
    finished, one increment of every unfinished run per device call."""
    while True:
        live = [r for r in runs if not r.finished]
        if not live:
            return
        reps = NativeSolver.step_batch([r.native for r in live], [r.next_params() for r in live])
        for r, rep in zip(live, reps):
            r.apply(rep)


class _Run:
    """The time loop of one SmallStrainFFT.run (small_strain_fft.py:106-246):
    adaptive increment, cutback, output times and post-processor, with each
    increment's predictor, Newton loop and MINRES corrections on the device."""

    def __init__(self, solver, material, orientation_label="orientation", phase_label=None,
                 output_variables=None, output_times=None, global_tolerance=1.0e-3,
                 strain_correction_tolerance=1.0e-2, linear_solver_tolerance=1.0e-5,
                 postprocessor=None, increment_callback=None):
        sv = self.solver = solver
        sv._sizes = material.sizes
        sv._dimensions = material.dimensions
        sv._num_points = material.num_points
        self.material = material
        self.output_variables = output_variables
        self.postprocessor = postprocessor
        self.increment_callback = increment_callback
        self.global_tolerance = global_tolerance
        self.strain_correction_tolerance = strain_correction_tolerance
        self.linear_solver_tolerance = linear_solver_tolerance
        self.time = sv.start_time
        self.time_tolerance = sv.min_time_increment / 2.0
        self.output_times = None
        if output_times is not None:
            self.time_idx = 0
            self.output_times = np.append(output_times, np.inf)
            self.next_output_time = self.output_times[0]
        else:
            self.next_output_time = np.inf
        self.new_state = defaultdict(list)
        orientations = material.extract(orientation_label)
        rotations = np.broadcast_to(orientations.rotation_matrix, (sv._num_points, 3, 3))
        self.models, model_of_point = sv._get_constitutive_model(
            material, phase_label, sv._constitutive_model)
        self.native = NativeSolver(sv._dimensions, sv.precision)
        try:
            for m in self.models:
                sv._register(self.native, m)
            self.native.set_microstructure(rotations, model_of_point)
            self.native.initialize()
        except BaseException:
            self.native.close()
            raise
        self.mask = np.asarray(sv.load_schedule.stress_mask, float)
        self.old_max_strain_increment = 0.0
        self.summed_von_mises_stress = 0.0
        self.time_step_id = 0
        self.time_increment = sv.initial_time_increment
        self.krylov = self.newton = 0
        self.floored = False
        self.stopped = False
        self.final = None
        # every attempted increment: (time, dt, converged, Newton iterations,
        # Krylov iterations, constitutive failure); kept up to date so that it
        # survives a "min time increment reached" error
        self.history = []
        sv.stats = {"history": self.history}

    @property
    def finished(self):
        return self.stopped or self.time >= (self.solver.end_time - self.time_tolerance)

    def next_params(self):
        sv, ls = self.solver, self.solver.load_schedule
        t, dt = self.time, self.time_increment
        sv._logger.info(f"New increment: time {t} + {dt}")
        d_strain = ls.strain_increment(t, dt).components
        d_stress = ls.stress_increment(t, dt).components
        applied = ls.stress(t, dt).components
        self._max_strain_increment = d_strain[np.argmax(np.abs(d_strain))]
        p = StepParams()
        p.strain_increment[:] = [float(v) for v in d_strain]
        p.stress_increment[:] = [float(v) for v in d_stress]
        p.applied_stress[:] = [float(v) for v in applied]
        p.stress_mask[:] = [float(v) for v in self.mask]
        p.time_increment = float(dt)
        p.global_tolerance = float(self.global_tolerance)
        p.strain_correction_tolerance = float(self.strain_correction_tolerance)
        p.linear_solver_tolerance = float(self.linear_solver_tolerance)
        p.summed_von_mises_stress = float(self.summed_von_mises_stress)
        p.time_step_id = int(self.time_step_id)
        p.reset_fluctuation = int(self._max_strain_increment * self.old_max_strain_increment < 0.0)
        return p

    def apply(self, rep):
        sv, native = self.solver, self.native
        if rep.linear_solver_tolerance_used > self.linear_solver_tolerance and not self.floored:
            self.floored = True
            sv._logger.warning(
                f"precision {sv.precision!r}: linear_solver_tolerance "
                f"{self.linear_solver_tolerance:g} raised to {rep.linear_solver_tolerance_used:g}")
        self.krylov += rep.krylov_iterations
        self.newton += rep.newton_iterations
        self.history.append((float(self.time), float(self.time_increment), bool(rep.converged),
                             int(rep.newton_iterations), int(rep.krylov_iterations),
                             bool(rep.constitutive_failed)))
        sv._logger.debug(f"global iterations {rep.newton_iterations}, "
                         f"linear solver iterations {rep.krylov_iterations}, "
                         f"equilibrium error {rep.equilibrium_error}")
        if not rep.converged and not rep.constitutive_failed:
            sv._logger.info("Too many iterations in macroloading loop")
        if not rep.converged:
            self.time_increment = self.time_increment / 2.0
            native.rollback()
            if self.time_increment < sv.min_time_increment:
                raise ValueError("min time increment reached")
            return
        self.summed_von_mises_stress += rep.mean_von_mises_stress
        self.old_max_strain_increment = self._max_strain_increment
        self.time_step_id += 1
        self.time += self.time_increment
        # the reference collects outputs, advances the output time and then
        # sizes the next increment (small_strain_fft.py:221-236); the device
        # commit (state, and fluctuation scaled by the new increment) comes
        # first here, which leaves the collected fields unchanged
        output_now = self.time >= (self.next_output_time - self.time_tolerance)
        upcoming = self.output_times[self.time_idx + 1] if output_now else self.next_output_time
        new_time_increment = sv._new_time_increment(
            self.time, self.time_increment, upcoming, rep.max_constitutive_iterations)
        native.commit(new_time_increment / self.time_increment)
        if output_now:
            outputs = self._fields()
            if self.postprocessor is not None:
                outputs = self.postprocessor(outputs)
            for k, v in outputs.items():
                self.new_state[k].append(v)
            self.time_idx += 1
            self.next_output_time = self.output_times[self.time_idx]
        self.time_increment = new_time_increment
        if self.increment_callback is not None and self.increment_callback({
                "step": self.time_step_id, "time": self.time,
                "krylov_iterations": int(rep.krylov_iterations),
                "newton_iterations": int(rep.newton_iterations),
                "constitutive_iterations": int(rep.max_constitutive_iterations),
                "num_points": sv._num_points}):
            self.stopped = True
        if self.finished:
            self._finish()

    def _fields(self):
        native = self.native
        outputs = self.solver._outputs(native, self.models, self.output_variables)
        outputs.update({"stress": Order2SymmetricTensor(native.get_field("stress"), "p"),
                        "strain": Order2SymmetricTensor(native.get_field("strain"), "p")})
        return outputs

    def _finish(self):
        sv = self.solver
        self.final = self._fields()
        sv.stats = {"krylov_iterations": int(self.krylov), "newton_iterations": int(self.newton),
                    "steps": int(self.time_step_id), "history": self.history,
                    "linear_solver_tolerance_used": max(
                        float(self.linear_solver_tolerance),
                        1.0e-5 if sv.precision == "mixed" else 0.0)}
        self.close()

    def close(self):
        if self.native is not None:
            self.native.close()

    def result(self):
        """The new Material (small_strain_fft.py:248-264)."""
        if self.final is None:
            self._finish()
        material, outputs, new_state = self.material, self.final, self.new_state
        if self.output_times is None:
            return material.create_fields(outputs)
        if self.postprocessor is None:
            for k, v in new_state.items():
                new_state[k] = type(v[0]).from_stack(v, new_dim="t", axis=len(v[0].dims_str))
            return material.create_fields(dict(new_state))
        new_material = material.create_fields(outputs)
        new_material.state.update(new_state)
        return new_material
