"""Experimental solvers on the propagation path (mirrors chromatix.experimental)."""
