"""optical_fft (src/chromatix/functional/convenience.py:9-59)."""
from .optics import optical_fft  # noqa: F401
