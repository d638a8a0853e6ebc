"""fft / ifft with optional centring shifts (src/chromatix/utils/fft.py:8-35)."""
from ..functional.optics import fft, ifft  # noqa: F401
