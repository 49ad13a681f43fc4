"""Isotropic elastic-plastic (J2) model
(``small_strain_fft/isotropic_elastic_plastic.py:17-192``).

The radial-return update (elastic predictor, fixed-point iteration on the
equivalent plastic strain increment run to the reference's global iteration
count, consistent tangent) runs in the HIP constitutive kernels, inside
``SmallStrainFFT`` or through the reference's protocol on a point handle; this
class carries the parameters and names the state variables.
"""

from ...tensor import Order2SymmetricTensor, Order4SymmetricTensor, Scalar
from ._device_model import DeviceModel
from .hardening_laws import KERNEL_LAWS


class IsotropicElasticPlastic(DeviceModel):
    # trial state of the last update (isotropic_elastic_plastic.py:131-137)
    _device_fields = {
        "yield_stresses": (Scalar, "p"),
        "plastic_strains": (Order2SymmetricTensor, "p"),
        "eq_plastic_strains": (Scalar, "p"),
    }

    def __init__(self, modulus, shear_modulus, yield_stress, hardening_function,
                 hardening_properties):
        if hardening_function not in KERNEL_LAWS:
            raise TypeError(
                "hardening_function must be one of perfect_plasticity, linear, voce, "
                "armstrong_frederick (the laws built into the device kernel)")
        self.modulus = modulus
        self.shear_modulus = shear_modulus
        self.bulk_modulus = shear_modulus * modulus / (3 * (3 * shear_modulus - modulus))
        self.stiffness = Order4SymmetricTensor.from_isotropic_constants(modulus, shear_modulus)
        self.yield_stress = Scalar(yield_stress)
        self.hardening_function = hardening_function
        self.hardening_properties = hardening_properties
        self.state_variables = dict()
        self.available_state_variables = [
            "yield_stresses",
            "plastic_strains",
            "eq_plastic_strains",
        ]
        self._state_variable_info = dict(zip(
            self.available_state_variables,
            [(Scalar, 0), (Order2SymmetricTensor, 0), (Scalar, 0)]))

    def _add_to(self, native):
        kind, props = self.kernel_hardening()
        return native.add_j2(self.stiffness.components, self.shear_modulus, self.bulk_modulus,
                             float(self.yield_stress.components), kind, props)

    def initialize(self, orientations):
        """isotropic_elastic_plastic.py:55-75: zero plastic state, the initial
        yield stress; returns the isotropic stiffness of every point."""
        tangent = super().initialize(orientations)
        self.state_variables["stiffnesses"] = tangent
        return tangent

    def kernel_hardening(self):
        kind, names = KERNEL_LAWS[self.hardening_function]
        props = self.hardening_properties or {}
        return kind, [float(props[k]) for k in names]

    def __repr__(self):
        return (f"IsotropicElasticPlastic(modulus={self.modulus}, "
                f"shear_modulus={self.shear_modulus}, "
                f"yield_stress={self.yield_stress.components}, "
                f"hardening_function={self.hardening_function.__name__}, "
                f"hardening_properties={self.hardening_properties})")
