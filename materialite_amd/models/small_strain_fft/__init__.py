"""SmallStrainFFT and its constitutive models (mirrors materialite.models.small_strain_fft)."""

from .elastic import Elastic
from .elastic_viscoplastic import ElasticViscoplastic
from .hardening_laws import armstrong_frederick, linear, perfect_plasticity, voce
from .isotropic_elastic_plastic import IsotropicElasticPlastic
from .load_schedule import LoadSchedule
from .multiphase import Multiphase
from .small_strain_fft import SmallStrainFFT

__all__ = ["Elastic", "ElasticViscoplastic", "IsotropicElasticPlastic", "LoadSchedule", "Multiphase", "SmallStrainFFT",
           "armstrong_frederick", "linear", "perfect_plasticity", "voce"]
