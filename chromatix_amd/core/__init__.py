"""Core data model: Field containers and spectra (mirrors chromatix.core)."""
from .field import (  # noqa: F401
    ChromaticScalarField,
    Field,
    ScalarField,
    crop,
    freq_grid,
    grid,
    pad,
    shift_field,
    shift_grid,
)
from .spectrum import MonoSpectrum, Spectrum  # noqa: F401
