"""System-level callers (mirrors src/chromatix/systems/microscopes.py)."""
from .microscopes import Microscope, Optical4FSystemPSF  # noqa: F401
