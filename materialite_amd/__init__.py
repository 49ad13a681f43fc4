"""materialite_amd: MI355X-native SmallStrainFFT (drop-in for materialite's hot path).

Import paths mirror the reference package: ``materialite_amd.Material``,
``materialite_amd.Order4SymmetricTensor`` ...,
``materialite_amd.models.small_strain_fft.SmallStrainFFT``.
"""

from .material import Box, Material, Sphere, import_evpfft
from .slip_system import SlipSystem
from .tensor import Order2SymmetricTensor, Order4SymmetricTensor, Orientation, Scalar, Vector

__all__ = ["Box", "Material", "Order2SymmetricTensor", "Order4SymmetricTensor", "Orientation",
           "Scalar", "SlipSystem", "Sphere", "Vector", "import_evpfft"]
