"""Optical elements used by the imaging path (mirrors chromatix.elements)."""
from .sensors import BasicSensor  # noqa: F401
