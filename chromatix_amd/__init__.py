"""chromatix_amd: the chromatix FFT-propagation hot path on MI355X (gfx950).

Hand-written HIP kernels (libcxm.so, C ABI in include/cxm.h) behind the
``chromatix.functional`` signatures and ``Field`` objects.  See
``chromatix_amd.functional`` for the operators and ``chromatix_amd.core``
for the data model.
"""
from .core import (  # noqa: F401
    ChromaticScalarField,
    Field,
    MonoSpectrum,
    ScalarField,
    Spectrum,
    crop,
    pad,
    shift_field,
    shift_grid,
)

__version__ = "0.1.0"
