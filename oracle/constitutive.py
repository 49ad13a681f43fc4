"""Constitutive models on the SmallStrainFFT path (oracle, FP64).

Restates, on plain arrays (points p, Mandel components m, slip systems s):

* ``Elastic``              ``small_strain_fft/elastic.py:16-49``
* ``ElasticViscoplastic``  ``small_strain_fft/elastic_viscoplastic.py:20-294``
  (crystal-frame implicit stress update, power-law slip, masked iteration once at
  least half of the points have converged, consistent tangent J^-1 C)
* hardening laws           ``small_strain_fft/hardening_laws.py:18-90``
* ``Multiphase``           ``small_strain_fft/multiphase.py:18-95``
* ``IsotropicElasticPlastic`` ``small_strain_fft/isotropic_elastic_plastic.py:17-192``
  (J2 radial return with isotropic hardening, fixed-point iteration on the
  equivalent plastic strain increment until every point has converged)
"""

import logging

import numpy as np

from .tensors import (
    directional_modulus,
    isotropic_stiffness,
    order2_to_crystal,
    order2_to_specimen,
    order4_to_specimen,
    rotation_matrix_mandel,
)


# ---------------------------------------------------------------- hardening laws
def _slip_increment(rates, dt):
    return dt * np.abs(rates).sum(axis=-1)


def perfect_plasticity(props, old_acc, old_g, rates, dt):
    """hardening_laws.py:18-25."""
    return old_g.copy()


def linear(props, old_acc, old_g, rates, dt):
    """hardening_laws.py:28-39."""
    return old_g + props["hardening_rate"] * _slip_increment(rates, dt)[:, None]


def voce(props, old_acc, old_g, rates, dt):
    """Closed-form integrated Voce increment (hardening_laws.py:42-72)."""
    d = _slip_increment(rates, dt)[:, None]
    theta0, theta1, tau1 = props["theta_0"], props["theta_1"], props["tau_1"]
    ratio = np.abs(theta0 / tau1)
    acc = old_acc[:, None]
    e0 = np.exp(-ratio * acc)
    ed = np.exp(-ratio * d)
    return (
        old_g
        + theta1 * d
        - (ratio * tau1 - theta1) / ratio * e0 * (ed - 1)
        - theta1 / ratio * e0 * (ed * (1 + ratio * (acc + d)) - 1 - ratio * acc)
    )


def armstrong_frederick(props, old_acc, old_g, rates, dt):
    """hardening_laws.py:75-90."""
    h = props["direct_hardening"]
    r = props["dynamic_recovery"]
    d = _slip_increment(rates, dt)[:, None]
    return old_g + (h - r * old_g) * d / (1 + r * d)


HARDENING = {
    "perfect_plasticity": perfect_plasticity,
    "linear": linear,
    "voce": voce,
    "armstrong_frederick": armstrong_frederick,
}


# ---------------------------------------------------------------- elastic
class Elastic:
    """sigma = C_specimen : eps, tangent C_specimen, 1 iteration (elastic.py:23-31)."""

    def __init__(self, stiffness):
        self.stiffness = np.asarray(stiffness, float)
        self.tangent = None

    def initialize(self, rotations):
        rm = rotation_matrix_mandel(rotations)
        self.tangent = order4_to_specimen(np.broadcast_to(self.stiffness, rm.shape), rm)
        return self.tangent.copy()

    def calculate_stress_and_tangent(self, strain, guess_stress, dt):
        return np.einsum("pmn,pn->pm", self.tangent, strain), self.tangent.copy(), 1

    def update_state_variables(self):
        pass

    def outputs(self):
        return {}


# ---------------------------------------------------------------- elasto-viscoplastic
class ElasticViscoplastic:
    """Crystal-plasticity stress update (elastic_viscoplastic.py:52-271)."""

    TOLERANCE = 2.0e-5
    MAX_ITERATIONS = 60

    def __init__(self, stiffness, schmid, reference_slip_rate, rate_exponent,
                 slip_resistance, hardening, hardening_properties):
        self.stiffness = np.asarray(stiffness, float)
        self.ref_modulus = float(directional_modulus(self.stiffness, [1, 0, 0]))
        self.schmid = np.asarray(schmid, float)  # (S, 6) Mandel sym(d n)
        self.gamma0 = float(reference_slip_rate)
        self.n = float(rate_exponent)
        # a number or one value per slip system, broadcast onto every point
        # (elastic_viscoplastic.py:66-69 adds it to a (points, systems) zero)
        self.g0 = np.asarray(slip_resistance, float)
        self.hardening = HARDENING[hardening] if isinstance(hardening, str) else hardening
        self.props = hardening_properties
        self._pp = np.einsum("sm,sn->smn", self.schmid, self.schmid)
        self.logger = logging.getLogger("ElasticViscoplastic")
        self.last_iterations = 0

    def initialize(self, rotations):
        n_pts = rotations.shape[0]
        n_sys = self.schmid.shape[0]
        self.R = np.asarray(rotations, float)
        self.Rm = rotation_matrix_mandel(self.R)
        g = np.zeros((n_pts, n_sys)) + self.g0
        self.g = g.copy()
        self.g_old = g.copy()
        self.ep = np.zeros((n_pts, 6))  # crystal frame, trial
        self.ep_old = np.zeros((n_pts, 6))
        self.ep_specimen = np.zeros((n_pts, 6))
        self.gamma = np.zeros((n_pts, n_sys))
        self.gamma_old = np.zeros((n_pts, n_sys))
        self.acc = np.zeros(n_pts)
        self.acc_old = np.zeros(n_pts)
        return order4_to_specimen(np.broadcast_to(self.stiffness, self.Rm.shape), self.Rm)

    def _run_iteration(self, g, stress, ep_old, dt, strain, acc_old, g_old):
        """elastic_viscoplastic.py:190-253."""
        prev = stress.copy()
        tau = stress @ self.schmid.T
        pref = self.gamma0 * np.abs(tau / g) ** (self.n - 1) / g
        rates = pref * tau
        ep = ep_old + (rates @ self.schmid) * dt
        m = np.einsum("ps,smn->pmn", self.n * pref, self._pp)
        jinv = np.linalg.inv(np.eye(6) + (dt * self.stiffness) @ m)
        resid = stress - (strain - ep) @ self.stiffness.T
        stress = prev - np.einsum("pmn,pn->pm", jinv, resid)
        g_new = self.hardening(self.props, acc_old, g_old, rates, dt)
        diff = np.linalg.norm(stress - prev, axis=-1)
        return stress, diff, g_new, ep, rates * dt, jinv

    def calculate_stress_and_tangent(self, strain, guess_stress, dt):
        """elastic_viscoplastic.py:91-188."""
        guess_norm = np.linalg.norm(guess_stress, axis=-1)
        stress = order2_to_crystal(guess_stress, self.R)
        strain_c = order2_to_crystal(strain, self.R)
        g = self.g
        n_pts = stress.shape[0]
        converged = False
        iteration = 0
        frac = 0.0
        conv = None
        while not converged:
            iteration += 1
            if frac >= 0.5:
                nc = ~conv
                (stress[nc], diff[nc], g[nc], ep[nc], dgam[nc], jinv[nc]) = self._run_iteration(
                    g[nc], stress[nc], self.ep_old[nc], dt, strain_c[nc],
                    self.acc_old[nc], self.g_old[nc],
                )
            else:
                stress, diff, g, ep, dgam, jinv = self._run_iteration(
                    g, stress, self.ep_old, dt, strain_c, self.acc_old, self.g_old
                )
            with np.errstate(divide="ignore", invalid="ignore"):
                conv = np.logical_or(
                    diff / guess_norm <= self.TOLERANCE, diff / self.ref_modulus < 1.0e-12
                )
            frac = np.count_nonzero(conv) / n_pts
            converged = bool(np.all(conv))
            if iteration == self.MAX_ITERATIONS and not converged:
                self.logger.info("Too many iterations in material loop")
                self.last_iterations = iteration
                return None, None, False
        tangent_c = jinv @ self.stiffness
        self.ep = ep
        self.g = g
        self.gamma = self.gamma_old + dgam
        self.acc = self.acc_old + np.abs(dgam).sum(axis=-1)
        self.last_iterations = iteration
        return (
            order2_to_specimen(stress, self.R),
            order4_to_specimen(tangent_c, self.Rm),
            iteration,
        )

    def update_state_variables(self):
        """elastic_viscoplastic.py:255-271."""
        self.ep_old = self.ep.copy()
        self.ep_specimen = order2_to_specimen(self.ep, self.R)
        self.g_old = self.g.copy()
        self.gamma_old = self.gamma.copy()
        self.acc_old = self.acc.copy()

    def outputs(self):
        return {
            "slip_resistances": self.g.copy(),
            "plastic_strains": self.ep_specimen.copy(),
            "slip_system_shear_strains": self.gamma.copy(),
        }


# ---------------------------------------------------------------- multiphase
class Multiphase:
    """Per-phase dispatch and scatter (multiphase.py:18-51)."""

    def __init__(self, models, indices, num_points):
        self.models = models
        self.indices = [np.asarray(i, int) for i in indices]
        self.num_points = num_points

    def initialize(self, rotations):
        tangent = np.zeros((self.num_points, 6, 6))
        for m, i in zip(self.models, self.indices):
            tangent[i] = m.initialize(rotations[i])
        return tangent

    def calculate_stress_and_tangent(self, strain, guess_stress, dt):
        stress = np.zeros((self.num_points, 6))
        tangent = np.zeros((self.num_points, 6, 6))
        iterations = []
        for m, i in zip(self.models, self.indices):
            s, t, it = m.calculate_stress_and_tangent(strain[i], guess_stress[i], dt)
            if not it:
                return None, None, it
            stress[i] = s
            tangent[i] = t
            iterations.append(it)
        return stress, tangent, max(iterations)

    def update_state_variables(self):
        for m in self.models:
            m.update_state_variables()


# ---------------------------------------------------------------- J2 plasticity
MANDEL_IDENTITY = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


class IsotropicElasticPlastic:
    """J2 plasticity with isotropic hardening (isotropic_elastic_plastic.py:17-192).

    The hardening law is one of the ``HARDENING`` functions applied per point
    with a single "slip system" (equivalent plastic strain rate), as the
    reference's laws fall back to for Scalars (hardening_laws.py:34-37).
    """

    TOLERANCE = 1.0e-7      # isotropic_elastic_plastic.py:79
    MAX_ITERATIONS = 100    # :80

    def __init__(self, modulus, shear_modulus, yield_stress, hardening, properties):
        self.modulus = modulus
        self.shear_modulus = shear_modulus
        # isotropic_elastic_plastic.py:29-31
        self.bulk_modulus = shear_modulus * modulus / (3 * (3 * shear_modulus - modulus))
        self.stiffness = isotropic_stiffness(modulus, shear_modulus)
        self.yield_stress = float(yield_stress)
        self.hardening = HARDENING[hardening] if isinstance(hardening, str) else hardening
        self.properties = properties

    def initialize(self, rotations):
        n = len(rotations)
        self.ep = np.zeros((n, 6))
        self.ep_old = np.zeros((n, 6))
        self.eq = np.zeros(n)
        self.eq_old = np.zeros(n)
        self.y = np.full(n, self.yield_stress)
        self.y_old = np.full(n, self.yield_stress)
        return np.broadcast_to(self.stiffness, (n, 6, 6)).copy()

    def _harden(self, rate, dt):
        return self.hardening(self.properties, self.eq_old, self.y_old[:, None],
                              rate[:, None], dt)[:, 0]

    def calculate_stress_and_tangent(self, strain, guess_stress, dt):
        """isotropic_elastic_plastic.py:78-168."""
        g, k = self.shear_modulus, self.bulk_modulus
        n = len(strain)
        ee = strain - self.ep_old
        tr = ee[:, 0] + ee[:, 1] + ee[:, 2]
        dev = ee - (1 / 3 * tr)[:, None] * MANDEL_IDENTITY
        s_pred = 2.0 * g * dev
        mean = k * tr
        mises = np.sqrt(1.5) * np.sqrt(np.einsum("pn,pn->p", s_pred, s_pred))
        with np.errstate(divide="ignore", invalid="ignore"):
            ratios = s_pred / mises[:, None]
        jinv = -1.0 / (3.0 * g)
        elastic = mises <= self.y_old
        tangent = np.broadcast_to(self.stiffness, (n, 6, 6)).copy()
        if np.all(elastic):
            return s_pred + mean[:, None] * MANDEL_IDENTITY, tangent, 1
        deps = np.zeros(n)
        y = self.y_old.copy()
        iteration = 0
        while True:
            iteration += 1
            r = mises - 3.0 * g * deps - y
            r[r <= 0.0] = 0.0
            prev = deps.copy()
            deps = deps - jinv * r
            y = self._harden(deps / dt, dt)
            converged = np.all(np.abs(deps - prev) < self.TOLERANCE)
            if converged:
                break
            if iteration == self.MAX_ITERATIONS:
                return None, None, False
        with np.errstate(invalid="ignore"):
            self.ep = self.ep_old + 1.5 * ratios * deps[:, None]
        self.eq = self.eq_old + deps
        self.y = y
        ratio = 3.0 * g / mises
        stress = (1.0 - ratio * deps)[:, None] * s_pred + mean[:, None] * MANDEL_IDENTITY
        dev_identity = np.eye(6) - np.outer(MANDEL_IDENTITY, MANDEL_IDENTITY) / 3.0
        term1 = 2.0 * g * ratio * deps
        term2 = ratio ** 2 * (deps / mises - 1 / (3 * g))
        d = (tangent - term1[:, None, None] * dev_identity
             + term2[:, None, None] * np.einsum("pi,pj->pij", s_pred, s_pred))
        d[elastic] = tangent[elastic]
        return stress, d, iteration

    def update_state_variables(self):
        self.ep_old = self.ep.copy()
        self.eq_old = self.eq.copy()
        self.y_old = self.y.copy()

    def outputs(self):
        return {"plastic_strains": self.ep.copy(), "eq_plastic_strains": self.eq.copy(),
                "yield_stresses": self.y.copy()}
