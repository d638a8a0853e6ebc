"""Optical4FSystemPSF (src/chromatix/systems/microscopes.py:300-365) on the
fused libcxm PSF kernels.

The reference module takes a ``Microscope`` for the objective's focal
length ``f``, ``n``, ``NA``, ``spectrum`` and ``padding_ratio``; any object
with those attributes works here.
"""
from __future__ import annotations

import numpy as np
import torch

from ..core.field import ScalarField
from ..core.spectrum import MonoSpectrum, Spectrum
from ..functional.psf import psf_4f


class Optical4FSystemPSF:
    """PSF of a 4f system with a phase mask in the pupil plane.

    shape: (height, width) of the simulation; spacing: output spacing (e.g.
    the camera pitch); f_tube: tube-lens focal length; phase: (height,
    width) phase mask in radians (a torch tensor; set requires_grad for
    gradients through ``intensity``)."""

    def __init__(self, shape, spacing, f_tube, phase):
        self.shape = tuple(int(s) for s in shape)
        self.spacing = float(spacing)
        self.f_tube = float(f_tube)
        self.phase = phase

    @staticmethod
    def compute_required_spacing(height, output_spacing, f, n, spectrum):
        """Pupil-plane spacing giving `output_spacing` after the tube lens:
        f lam / (n height output_spacing) (microscopes.py:357-365)."""
        return f * spectrum / (n * height * output_spacing)

    def _setup(self, microscope):
        padding = tuple(int(s * getattr(microscope, "padding_ratio", 0)) for s in self.shape)
        padded = tuple(s + p for s, p in zip(self.shape, padding))
        spectrum = Spectrum.build(microscope.spectrum)
        if not isinstance(spectrum, MonoSpectrum):
            raise NotImplementedError("Optical4FSystemPSF: monochromatic spectra only")
        lam = float(spectrum.wavelength[0])
        dx = self.compute_required_spacing(padded[0], self.spacing, self.f_tube, microscope.n, lam)
        phase = torch.as_tensor(self.phase)
        if tuple(phase.shape) != padded:
            raise ValueError(f"phase must have the simulation shape {padded}")
        return phase, dx, lam

    def __call__(self, microscope, z) -> ScalarField:
        phase, dx, lam = self._setup(microscope)
        return psf_4f(phase, dx, lam, z, microscope.f, microscope.n, microscope.NA, self.f_tube,
                      output="field")

    def intensity(self, microscope, z) -> torch.Tensor:
        """|field|^2 of __call__, differentiable with respect to ``phase``."""
        phase, dx, lam = self._setup(microscope)
        return psf_4f(phase, dx, lam, z, microscope.f, microscope.n, microscope.NA, self.f_tube)
