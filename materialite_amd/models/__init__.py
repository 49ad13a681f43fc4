"""Models on the SmallStrainFFT path (import path mirrors materialite.models)."""
