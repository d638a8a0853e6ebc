"""System-level caller of the PSF path (mirrors src/chromatix/systems/microscopes.py)."""
from .microscopes import Optical4FSystemPSF  # noqa: F401
