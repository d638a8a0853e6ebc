"""Field sources on the PSF path (mirrors src/chromatix/functional/sources.py).

``objective_point_source`` (sources.py:122-209) is evaluated by a libcxm
kernel: defocus chirp, circular pupil and power normalisation in one pass.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ScalarField
from ..core.spectrum import MonoSpectrum, Spectrum
from .propagation import get_phase_mode

__all__ = ["objective_point_source"]


def _spacing(dx) -> tuple[float, float]:
    d = np.asarray(dx, dtype=np.float64).reshape(-1)
    if d.size == 1:
        return float(d[0]), float(d[0])
    if d.size == 2:
        return float(d[0]), float(d[1])
    raise ValueError("spacing must be a scalar or (dy, dx) for a monochromatic source")


def psf_desc(shape, dx, wavelength, z, f, n, NA, power, amplitude, offset, phase_mode):
    """libcxm descriptor of the point source / 4f PSF (host arrays kept alive
    by the returned tuple)."""
    d = L.Psf4fDesc()
    d.height, d.width = int(shape[0]), int(shape[1])
    zs = np.atleast_1d(np.asarray(z, dtype=np.float64)).reshape(-1)
    d.num_z = int(zs.size)
    d.phase_mode = L.PHASE_MODES[phase_mode or get_phase_mode()]
    d.normalize = 0 if power is None else 1
    d.dy, d.dx = _spacing(dx)
    d.wavelength = float(wavelength)
    d.f, d.n, d.NA = float(f), float(n), float(NA)
    d.power = 0.0 if power is None else float(power)
    amp = np.asarray(amplitude, dtype=np.float64).reshape(-1)
    if amp.size != 1:
        raise ValueError("scalar fields take a scalar amplitude")
    d.amplitude = float(amp[0])
    off = np.asarray(offset, dtype=np.float64).reshape(-1)
    d.offset_y, d.offset_x = float(off[0]), float(off[1])
    zbuf = L.doubles(zs)
    d.z = ctypes.cast(zbuf, ctypes.POINTER(ctypes.c_double))
    return d, zbuf, zs


def objective_point_source(shape, dx, spectrum, z, f, n, NA, power=1.0, amplitude=1.0,
                           offset=(0.0, 0.0), scalar: bool = True, *, device=None,
                           phase_mode=None) -> ScalarField:
    """Point source defocused by ``z`` in front of a lens of focal length
    ``f`` and numerical aperture ``NA``, just after the lens:
    ``-i A / L^2 exp(-i pi (z/f) |r - offset|^2 / L^2)`` with ``L^2 = lam f / n``,
    inside a circular pupil of diameter ``2 f NA / n``, normalised to total
    ``power`` per plane.  A ``z`` array of length Z gives ``u`` of shape
    (Z, H, W).  Monochromatic scalar fields."""
    if not scalar:
        raise NotImplementedError("vector fields are not on the propagation path")
    spectrum = Spectrum.build(spectrum)
    if not isinstance(spectrum, MonoSpectrum):
        raise NotImplementedError("objective_point_source: monochromatic spectra only")
    d, keep, zs = psf_desc(shape, dx, spectrum.wavelength[0], z, f, n, NA, power, amplitude,
                           offset, phase_mode)
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    out = torch.empty((d.num_z, d.height, d.width), dtype=torch.complex64, device=dev)
    L.ensure_device(out)
    L.check(L.lib().cxm_point_source_c64(L.stream_ptr(out.device), out.data_ptr(), ctypes.byref(d)))
    if zs.size == 1:  # the reference squeezes a single z (sources.py:196-199)
        out = out[0]
    return ScalarField(out, (d.dy, d.dx), 0.0, spectrum)
