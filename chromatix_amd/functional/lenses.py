"""Lenses of the propagation path: ff_lens (src/chromatix/functional/lenses.py:64-95)."""
from .optics import ff_lens  # noqa: F401
