"""SmallStrainFFT Newton-Krylov solver and load schedules (oracle, FP64).

Restates ``SmallStrainFFT.run`` / ``_get_new_time_increment`` /
``_get_fluctuation_strain`` / ``_left_hand_side``
(``materialite/models/small_strain_fft/small_strain_fft.py:60-299, 370-374``) and
``LoadSchedule`` (``small_strain_fft/load_schedule.py:17-119``, all five
factories) on plain arrays:
fields are (N, 6) Mandel arrays in the reference's point order (x slowest, z
fastest), tangents are (N, 6, 6).
"""

import logging

import numpy as np

from .minres import minres
from .projection import apply_projection, projection_closed_form
from .tensors import mandel_dev


# ---------------------------------------------------------------- load schedules
class LoadSchedule:
    """load_schedule.py:17-30; callables return Mandel 6-vectors."""

    def __init__(self, strain_increment, stress_increment, stress, stress_mask):
        self.strain_increment = strain_increment
        self.stress_increment = stress_increment
        self.stress = stress
        self.stress_mask = np.asarray(stress_mask, float)
        if np.any(np.logical_and(self.stress_mask != 1, self.stress_mask != 0)):
            raise ValueError("stress mask can only have zeros and ones")

    @classmethod
    def from_constant_rates(cls, strain_rate, stress_rate, stress_mask, start_time=0):
        """load_schedule.py:32-40."""
        mask = np.asarray(stress_mask, float)
        de = np.asarray(strain_rate, float) * (1 - mask)
        ds = np.asarray(stress_rate, float) * mask
        return cls(
            lambda t, dt: de * dt,
            lambda t, dt: ds * dt,
            lambda t, dt: ds * (t + dt - start_time),
            mask,
        )

    @classmethod
    def from_constant_uniaxial_strain_rate(cls, magnitude=1.0, direction="z", start_time=0):
        """load_schedule.py:42-61."""
        rates = {"x": [magnitude, 0, 0], "y": [0, magnitude, 0], "z": [0, 0, magnitude]}
        masks = {"x": [0, 1, 1], "y": [1, 0, 1], "z": [1, 1, 0]}
        mask = np.concatenate([masks[direction], np.ones(3)])
        rate = np.concatenate([rates[direction], np.zeros(3)])
        return cls.from_constant_rates(rate, np.zeros(6), mask, start_time)


def _ramp(times, strains, stresses, stress_mask):
    """load_schedule.py:63-75: linear in time through two (time, strain,
    stress) states, each masked."""
    mask = np.asarray(stress_mask, float)
    e = np.asarray(strains, float) * (1 - mask)
    s = np.asarray(stresses, float) * mask
    t0, span = times[0], times[1] - times[0]
    de, ds = (e[1] - e[0]) / span, (s[1] - s[0]) / span
    return LoadSchedule(lambda t, dt: de * dt, lambda t, dt: ds * dt,
                        lambda t, dt: ds * (t + dt - t0) + s[0], mask)


def _cyclic(amplitude, frequency, start_time, ramp_rate):
    a = np.asarray(amplitude, float)
    return lambda t: a * (1 + ramp_rate * (t - start_time)) * np.sin(
        2 * np.pi * frequency * (t - start_time))


def _cyclic_stress(stress_amplitude, frequency, stress_mask=np.ones(6), start_time=0, ramp_rate=0):
    """load_schedule.py:77-98: A(t) = a (1 + r (t - t0)) sin(2 pi f (t - t0)) on
    the stress-controlled components; no strain increments."""
    mask = np.asarray(stress_mask, float)
    amp = _cyclic(np.asarray(stress_amplitude, float) * mask, frequency, start_time, ramp_rate)
    return LoadSchedule(lambda t, dt: np.zeros(6), lambda t, dt: amp(t + dt) - amp(t),
                        lambda t, dt: amp(t + dt), mask)


def _cyclic_strain(strain_amplitude, frequency, stress_mask, start_time=0, ramp_rate=0):
    """load_schedule.py:100-119: the same amplitude law on the strain-controlled
    components; zero applied stress."""
    mask = np.asarray(stress_mask, float)
    amp = _cyclic(np.asarray(strain_amplitude, float) * (1 - mask), frequency, start_time, ramp_rate)
    return LoadSchedule(lambda t, dt: amp(t + dt) - amp(t), lambda t, dt: np.zeros(6),
                        lambda t, dt: np.zeros(6), mask)


LoadSchedule.from_ramp = staticmethod(_ramp)
LoadSchedule.from_cyclic_stress = staticmethod(_cyclic_stress)
LoadSchedule.from_cyclic_strain = staticmethod(_cyclic_strain)


def _von_mises_of_mean(stress):
    return float(np.linalg.norm(mandel_dev(stress.mean(axis=0)))) * np.sqrt(1.5)


class SmallStrainFFT:
    """Array-level restatement of small_strain_fft.py:29-299."""

    MAX_ITERATIONS = 100

    def __init__(self, load_schedule, end_time=1, initial_time_increment=None,
                 max_time_increment=None, min_time_increment=1.0e-9, start_time=0.0):
        self.load_schedule = load_schedule
        self.end_time = end_time
        self.initial_time_increment = (
            end_time if initial_time_increment is None else initial_time_increment
        )
        self.max_time_increment = (
            self.initial_time_increment if max_time_increment is None else max_time_increment
        )
        self.min_time_increment = min_time_increment
        self.start_time = start_time
        self.logger = logging.getLogger("SmallStrainFFT.oracle")
        self.stats = {}

    def _new_time_increment(self, time, dt, next_output_time, max_constit_iters):
        """small_strain_fft.py:266-279."""
        fixed = min(next_output_time - time, self.end_time - time)
        eligible = [self.max_time_increment, fixed]
        eligible.append(dt * 2.0 if max_constit_iters < 11 else dt)
        return min(eligible)

    def run(self, dims, rotations, model, global_tolerance=1.0e-3,
            strain_correction_tolerance=1.0e-2, linear_solver_tolerance=1.0e-5,
            max_steps=None):
        """small_strain_fft.py:60-264 -> dict(stress, strain, stats).

        ``max_steps`` stops after that many converged increments (used to time a
        bounded CPU sample); None runs to ``end_time``.
        """
        dims = tuple(int(d) for d in dims)
        n_pts = int(np.prod(dims))
        ls = self.load_schedule
        mask = ls.stress_mask
        time = self.start_time
        time_tol = self.min_time_increment / 2.0
        next_output_time = np.inf

        def project(field):
            return apply_projection(field, dims, stress_mask=mask).ravel()

        def lhs(x, tangent):
            return project(np.einsum("pmn,pn->pm", tangent, x.reshape(n_pts, 6)))

        tangent = model.initialize(rotations)
        stress = np.zeros((n_pts, 6))
        strain = np.zeros((n_pts, 6))
        old_tangent = tangent.copy()
        old_strain = np.zeros((n_pts, 6))
        old_stress = np.zeros((n_pts, 6))
        old_fluct = np.zeros((n_pts, 6))
        old_max_inc = 0.0
        summed_vm = 0.0
        step_id = 0
        dt = self.initial_time_increment
        krylov_total = 0
        newton_total = 0
        krylov_per_solve = []
        steps_done = 0
        # every attempted increment: (time, dt, converged, Newton iterations,
        # Krylov iterations, constitutive failure)
        history = []
        self.stats = {"history": history}

        while time < (self.end_time - time_tol):
            if max_steps is not None and steps_done >= max_steps:
                break
            d_strain = np.asarray(ls.strain_increment(time, dt), float)
            d_stress = np.asarray(ls.stress_increment(time, dt), float)
            applied = np.asarray(ls.stress(time, dt), float)
            max_inc = d_strain[np.argmax(np.abs(d_strain))]
            if max_inc * old_max_inc < 0.0:
                old_fluct = np.zeros((n_pts, 6))
            strain = old_strain + d_strain + old_fluct
            if not np.all(d_stress < 1.0e-14):
                strain = strain + np.linalg.inv(tangent.mean(axis=0)) @ d_stress
            if step_id == 0:
                guess = d_stress + np.einsum("pmn,n->pm", tangent, d_strain)
            else:
                guess = old_stress + d_stress

            converged = False
            iteration = 0
            all_iters = []
            attempt_krylov = 0
            failed = False
            while not converged:
                iteration += 1
                if iteration == 1:
                    stress, tangent, it = model.calculate_stress_and_tangent(strain, guess, dt)
                    if not it:
                        tangent = old_tangent.copy()
                        failed = True
                        break
                    b = project(applied - stress)
                    vm = _von_mises_of_mean(stress)
                    eq_err = np.max(np.abs(b)) / vm
                    guess = stress.copy()
                    all_iters.append(it)
                old_eq_err = max(eq_err, global_tolerance)
                x, itn, _ = minres(lambda v: lhs(v, tangent), b, rtol=linear_solver_tolerance,
                                   maxiter=6 * n_pts * 5)
                krylov_total += itn
                attempt_krylov += itn
                krylov_per_solve.append(itn)
                newton_total += 1
                fluct = x.reshape(n_pts, 6)
                strain = strain + fluct
                old_fluct = old_fluct + fluct

                stress, tangent, it = model.calculate_stress_and_tangent(strain, guess, dt)
                if not it:
                    tangent = old_tangent.copy()
                    failed = True
                    break
                all_iters.append(it)
                b = project(applied - stress)
                vm = _von_mises_of_mean(stress)
                time_avg_vm = (summed_vm + vm) / (step_id + 1)
                if vm / time_avg_vm < 1.0e-12:
                    vm = time_avg_vm
                eq_err = np.max(np.abs(b)) / vm
                max_dstress = np.max(
                    np.linalg.norm(mandel_dev(stress - old_stress), axis=-1)
                ) * np.sqrt(1.5)
                if max_dstress / vm < 1.0e-12:
                    max_dstress = vm
                strain_err = np.max(
                    np.linalg.norm(np.einsum("pmn,pn->pm", tangent, fluct), axis=-1)
                ) / max_dstress
                projected = strain_err * eq_err / old_eq_err
                if eq_err < global_tolerance and (
                    strain_err < strain_correction_tolerance
                    or projected < strain_correction_tolerance
                    or step_id == 0
                    or eq_err < 1e-8
                ):
                    converged = True
                guess = stress.copy()
                if iteration == self.MAX_ITERATIONS and not converged:
                    self.logger.info("Too many iterations in macroloading loop")
                    break

            history.append((time, dt, converged, iteration, attempt_krylov, failed))
            if converged:
                model.update_state_variables()
                old_stress = stress.copy()
                old_strain = strain.copy()
                old_tangent = tangent.copy()
                summed_vm += vm
                old_max_inc = max_inc
                step_id += 1
                steps_done += 1
                time += dt
                new_dt = self._new_time_increment(time, dt, next_output_time, max(all_iters))
                old_fluct = old_fluct * new_dt / dt
                dt = new_dt
            else:
                dt = dt / 2.0
                old_fluct = old_fluct / 2.0
                if dt < self.min_time_increment:
                    raise ValueError("min time increment reached")

        self.stats = {
            "krylov_iterations": krylov_total,
            "newton_iterations": newton_total,
            "krylov_per_solve": krylov_per_solve,
            "steps": step_id,
            "history": history,
        }
        out = {"stress": stress, "strain": strain}
        out.update(model.outputs() if hasattr(model, "outputs") else {})
        return out


def projection_table(dims, stress_mask):
    """The materialised (Nx,Ny,Nz,6,6) Ghat the reference keeps, for small grids."""
    return projection_closed_form(dims, stress_mask)
