"""Oracle: field sources and phase masks of the 4f PSF path (test infrastructure).

Follows (paths relative to the chromatix repository root):
  objective_point_source  src/chromatix/functional/sources.py:122-209
  plane_wave              src/chromatix/functional/sources.py:212-293
  circular_pupil          src/chromatix/functional/pupils.py:17-23
  phase_change            src/chromatix/functional/phase_masks.py:29-68
  power                   src/chromatix/core/field.py:656-661
Scalar monochromatic fields, FP64.
"""
from __future__ import annotations

import numpy as np

from .propagation import grid


def power(u, spacing):
    """prod(dx) * sum |u|^2 over the last two axes (keepdims)."""
    return spacing[0] * spacing[1] * np.sum(np.abs(u) ** 2, axis=(-2, -1), keepdims=True)


def objective_point_source(shape, spacing, wavelength, z, f, n, NA, power_=1.0, amplitude=1.0,
                           offset=(0.0, 0.0)):
    """u = amplitude * -i/L^2 exp(-i pi (z/f) |r - offset|^2 / L^2), L^2 = lam f / n,
    times the circular pupil of diameter 2 f NA / n, normalised to `power_`.
    A z array of length Z gives (Z, H, W)."""
    z = np.asarray(z, np.float64)
    zz = z.reshape(-1)
    g = grid(shape, spacing)
    r2 = (g[..., 0] - offset[0]) ** 2 + (g[..., 1] - offset[1]) ** 2
    L2 = wavelength * f / n
    D = 2 * f * NA / n
    mask = (g[..., 0] ** 2 + g[..., 1] ** 2) <= (D / 2) ** 2
    u = np.stack([amplitude * -1j / L2 * np.exp(-1j * np.pi * (zi / f) * r2 / L2) * mask for zi in zz])
    if power_ is not None:
        u = u * np.sqrt(power_ / power(u, spacing))
    return u if zz.size > 1 else u[0]


def plane_wave(shape, spacing, wavelength, power_=1.0, amplitude=1.0, kykx=(0.0, 0.0), pupil=None):
    """amplitude * exp(i k.r), optional pupil(u, grid) -> u, normalised to `power_`."""
    g = grid(shape, spacing)
    u = amplitude * np.exp(1j * (kykx[0] * g[..., 0] + kykx[1] * g[..., 1]))
    if pupil is not None:
        u = pupil(u, g)
    if power_ is not None:
        u = u * np.sqrt(power_ / power(u, spacing))
    return u


def phase_change(u, phase, spectral_modulation=1.0):
    """u * exp(i * phase * lam0 / lam)."""
    return u * np.exp(1j * phase * spectral_modulation)
