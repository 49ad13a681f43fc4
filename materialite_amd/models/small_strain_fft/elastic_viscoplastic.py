"""Elasto-viscoplastic crystal-plasticity model
(``small_strain_fft/elastic_viscoplastic.py:20-294``).

The per-point implicit update, its masked iteration order and the consistent
tangent run in the HIP constitutive kernels, inside ``SmallStrainFFT`` on its
grid handle or through the reference's protocol (``initialize``,
``calculate_stress_and_tangent``, ``update_state_variables``, ...) on a point
handle; this class carries the parameters and names the state variables.
"""

import numpy as np

from ...tensor import Order2SymmetricTensor, Scalar, Vector
from ._device_model import DeviceModel
from .hardening_laws import KERNEL_LAWS


class ElasticViscoplastic(DeviceModel):
    # state held on the device: trial slip resistances and slip shears, the
    # committed plastic strain in the specimen frame (elastic_viscoplastic.py:
    # 175-182, 255-271)
    _device_fields = {
        "slip_resistances": (Scalar, "ps"),
        "plastic_strains": (Order2SymmetricTensor, "p"),
        "slip_system_shear_strains": (Scalar, "ps"),
        "accumulated_slip": (Scalar, "p"),
    }

    def __init__(self, stiffness, slip_systems, reference_slip_rate, rate_exponent,
                 slip_resistance, hardening_function, hardening_properties):
        if hardening_function not in KERNEL_LAWS:
            raise TypeError(
                "hardening_function must be one of perfect_plasticity, linear, voce, "
                "armstrong_frederick (the laws built into the device kernel)")
        self.stiffness = stiffness
        self._ref_modulus = float(stiffness.directional_modulus(Vector([1, 0, 0])).components)
        self.slip_systems = slip_systems
        self.reference_slip_rate = Scalar(reference_slip_rate)
        self.rate_exponent = Scalar(rate_exponent)
        self.initial_slip_resistance = slip_resistance
        self.hardening_function = hardening_function
        self.hardening_properties = hardening_properties
        self.state_variables = dict()
        self.available_state_variables = [
            "slip_resistances",
            "plastic_strains",
            "slip_system_shear_strains",
        ]
        n = len(slip_systems)
        self._state_variable_info = dict(zip(self.available_state_variables,
                                             [(Scalar, n), (Order2SymmetricTensor, 0), (Scalar, n)]))

    def _initial_slip_resistances(self):
        """A float, or one value per slip system: the initial slip resistance
        is added to a (points, systems) zero Scalar (elastic_viscoplastic.py:
        66-69), so a number or a Scalar over systems ("s") broadcasts onto
        every point; a per-point Scalar is not supported."""
        g0 = self.initial_slip_resistance
        n = len(self.slip_systems)
        if isinstance(g0, Scalar):
            if g0.dims_str == "":
                return float(g0.components)
            if g0.dims_str != "s":
                raise NotImplementedError(
                    f"initial slip resistance with dims {g0.dims_str!r}: only a number or "
                    "one value per slip system (dims 's') is supported")
            g0 = g0.components
        g0 = np.asarray(g0, dtype=np.float64)
        if g0.ndim == 0:
            return float(g0)
        if g0.shape != (n,):
            raise ValueError(f"initial slip resistance of shape {g0.shape}: expected ({n},)")
        return g0

    def _add_to(self, native):
        kind, props = self.kernel_hardening()
        return native.add_evp(self.stiffness.components, self.slip_systems.schmid_tensor.components,
                              float(self.reference_slip_rate.components),
                              float(self.rate_exponent.components),
                              self._initial_slip_resistances(), kind, props)

    def initialize(self, orientations):
        """elastic_viscoplastic.py:52-89: zero plastic state, slip resistances at
        their initial value; returns the elastic stiffness in the specimen frame."""
        tangent = super().initialize(orientations)
        schmid = self.slip_systems.schmid_tensor
        self.state_variables["stiffnesses"] = self.stiffness
        self.state_variables["schmid_tensors"] = schmid
        return tangent

    def kernel_hardening(self):
        kind, names = KERNEL_LAWS[self.hardening_function]
        props = self.hardening_properties or {}
        return kind, [float(props[k]) for k in names]

    def __repr__(self):
        return (f"ElasticViscoplastic(stiffness={self.stiffness.voigt}, "
                f"reference_slip_rate={self.reference_slip_rate.components}, "
                f"rate_exponent={self.rate_exponent.components}, "
                f"initial_slip_resistance={self.initial_slip_resistance}, "
                f"hardening_function={self.hardening_function.__name__}, "
                f"hardening_properties={self.hardening_properties})")
