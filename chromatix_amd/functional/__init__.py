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
    high_na_ff_lens,
    optical_fft,
    transform_propagate,
    transform_propagate_sas,
)
from .phase_masks import phase_change  # noqa: F401,E402
from .sources import objective_point_source  # noqa: F401,E402
from .psf import psf_4f, psf_4f_mse_grad, psf_output_spacing  # noqa: F401,E402
from .samples import (  # noqa: F401,E402
    fluorescent_multislice_thick_sample,
    multislice_mse_grad,
    multislice_thick_sample,
    thin_sample,
)
from .sensors import basic_sensor  # noqa: F401,E402
