"""Utilities of the propagation path (mirrors chromatix.utils)."""
