"""Slip-resistance hardening laws (``small_strain_fft/hardening_laws.py``).

The functions evaluate the closed forms on host arrays when called directly;
``ElasticViscoplastic`` recognises them by identity and runs the same law
inside the HIP constitutive kernel.
"""

import numpy as np

from ...tensor import Scalar


def _increment(rates, dt):
    c = rates.components if isinstance(rates, Scalar) else np.asarray(rates, float)
    return dt * (np.abs(c).sum(axis=-1) if c.ndim > 1 else np.abs(c))


def _g(old):
    return old.components if isinstance(old, Scalar) else np.asarray(old, float)


def _wrap(old, values):
    return Scalar(values, old.dims_str) if isinstance(old, Scalar) else values


def perfect_plasticity(properties, old_accumulated_slip, old_slip_resistances,
                       plastic_slip_rates, time_increment):
    """hardening_laws.py:18-25."""
    return old_slip_resistances


def linear(properties, old_accumulated_slip, old_slip_resistances, plastic_slip_rates,
           time_increment):
    """hardening_laws.py:28-39."""
    d = _increment(plastic_slip_rates, time_increment)
    g = _g(old_slip_resistances)
    d = d[..., None] if g.ndim > np.ndim(d) else d
    return _wrap(old_slip_resistances, g + properties["hardening_rate"] * d)


def voce(properties, old_accumulated_slip, old_slip_resistances, plastic_slip_rates,
         time_increment):
    """Integrated Voce law, closed-form increment (hardening_laws.py:42-72)."""
    g = _g(old_slip_resistances)
    d = _increment(plastic_slip_rates, time_increment)
    acc = _g(old_accumulated_slip)
    if g.ndim > np.ndim(d):
        d = d[..., None]
        acc = np.asarray(acc)[..., None]
    t0, t1, tau1 = properties["theta_0"], properties["theta_1"], properties["tau_1"]
    r = np.abs(t0 / tau1)
    e0 = np.exp(-r * acc)
    ed = np.exp(-r * d)
    out = (g + t1 * d - (r * tau1 - t1) / r * e0 * (ed - 1)
           - t1 / r * e0 * (ed * (1 + r * (acc + d)) - 1 - r * acc))
    return _wrap(old_slip_resistances, out)


def armstrong_frederick(properties, old_accumulated_slip, old_slip_resistances,
                        plastic_slip_rates, time_increment):
    """hardening_laws.py:75-90."""
    h, rd = properties["direct_hardening"], properties["dynamic_recovery"]
    g = _g(old_slip_resistances)
    d = _increment(plastic_slip_rates, time_increment)
    d = d[..., None] if g.ndim > np.ndim(d) else d
    return _wrap(old_slip_resistances, g + (h - rd * g) * d / (1 + rd * d))


# kernel identifiers (include/materialite_amd.h MTL_HARDENING_*) and the order of
# the property values the kernel reads
KERNEL_LAWS = {
    perfect_plasticity: (0, ()),
    linear: (1, ("hardening_rate",)),
    voce: (2, ("theta_0", "theta_1", "tau_1")),
    armstrong_frederick: (3, ("direct_hardening", "dynamic_recovery")),
}
