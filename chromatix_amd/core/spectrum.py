"""Spectrum containers mirroring src/chromatix/core/spectrum.py.

Wavelengths and densities are host metadata (float64 numpy arrays); they are
rounded to float32 only where a kernel follows the reference's float32
arithmetic.
"""
from __future__ import annotations

import numpy as np


class Spectrum:
    """Wavelengths and their relative densities (core/spectrum.py:10-93)."""

    def __init__(self, wavelength, density=None):
        wavelength = np.atleast_1d(np.asarray(wavelength, dtype=np.float64))
        if wavelength.ndim != 1:
            raise ValueError("Wavelength must be 1D for polychromatic spectrum.")
        self.wavelength = wavelength
        if density is None:
            self.density = np.full_like(wavelength, 1.0 / wavelength.size)
        else:
            density = np.atleast_1d(np.asarray(density, dtype=np.float64))
            if density.shape != wavelength.shape:
                raise ValueError(
                    f"Density shape {density.shape} must match wavelength shape {wavelength.shape}."
                )
            self.density = density / density.sum()

    @classmethod
    def build(cls, spectrum) -> "Spectrum":
        """Same dispatch as core/spectrum.py:37-73: a scalar gives a
        MonoSpectrum, a tuple/list gives (wavelengths, densities), a size-1
        array gives a MonoSpectrum, any other array a Spectrum."""
        if isinstance(spectrum, Spectrum):
            return spectrum
        if isinstance(spectrum, (float, int)):
            return MonoSpectrum(spectrum)
        if isinstance(spectrum, (tuple, list)):
            return Spectrum(spectrum[0], spectrum[1])
        arr = np.asarray(spectrum)
        if arr.size == 1:
            return MonoSpectrum(arr.reshape(()))
        return Spectrum(arr)

    @property
    def size(self) -> int:
        return int(self.wavelength.size)

    @property
    def central_wavelength(self) -> float:
        return float(self.wavelength[0])

    @property
    def spectral_modulation(self) -> np.ndarray:
        return self.central_wavelength / self.wavelength

    def __repr__(self) -> str:
        return f"{type(self).__name__}(wavelength={self.wavelength}, density={self.density})"


class MonoSpectrum(Spectrum):
    """A single wavelength (core/spectrum.py:96-107)."""

    def __init__(self, wavelength):
        wl = np.atleast_1d(np.asarray(wavelength, dtype=np.float64))
        if wl.shape != (1,):
            raise ValueError(
                f"MonoSpectrum requires a scalar wavelength, got array of shape {wl.shape}"
            )
        self.wavelength = wl
        self.density = np.ones((1,))
