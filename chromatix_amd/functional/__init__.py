"""Functional API of the FFT-propagation path (mirrors chromatix.functional)."""
from .propagation import *  # noqa: F401,F403
from .propagation import (  # noqa: F401
    get_phase_mode,
    propagate_intensity,
    set_phase_mode,
)
from .optics import (  # noqa: F401,E402
    compute_sas_precompensation,
    ff_lens,
    optical_fft,
    transform_propagate,
    transform_propagate_sas,
)
