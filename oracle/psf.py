"""Oracle: 4f widefield PSF stack and its phase-mask gradient (test
infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  Optical4FSystemPSF.__call__   src/chromatix/systems/microscopes.py:334-355
  (chromatic)                   with field.power / .intensity, core/field.py:729-740,
                                and spectrally_modulate_phase, phase_masks.py:105-124
  compute_required_spacing      src/chromatix/systems/microscopes.py:357-365
  objective_point_source        src/chromatix/functional/sources.py:122-209
  phase_change                  src/chromatix/functional/phase_masks.py:29-68
  ff_lens / optical_fft         src/chromatix/functional/lenses.py:64-95,
                                src/chromatix/functional/convenience.py:9-59
The gradient is the adjoint of that composition for L(I), written out by
hand (the reference gets it from JAX reverse mode); it is pinned by finite
differences of psf4f in tests/test_oracle_psf.py.
"""
from __future__ import annotations

import numpy as np

from .optics import ff_lens
from .sources import objective_point_source


def required_spacing(height, output_spacing, f_tube, n, wavelength):
    """microscopes.py:357-365: f * lam / (n * height * output_spacing)."""
    return f_tube * wavelength / (n * height * output_spacing)


def psf4f_field(phase, spacing, wavelength, z, f, n, NA, f_tube, power=1.0):
    """Complex field in the tube lens' back focal plane, (Z, H, W), and its
    spacing."""
    shape = np.shape(phase)
    u = objective_point_source(shape, spacing, wavelength, np.atleast_1d(z), f, n, NA, power)
    u = np.asarray(u).reshape((-1,) + shape)
    u = u * np.exp(1j * np.asarray(phase, np.float64))
    return ff_lens(u, spacing, wavelength, f_tube, n)


def psf4f_chromatic(phase, output_spacing, wavelengths, density, z, f, n, NA, f_tube, power=1.0):
    """Intensity stack (Z, H, W) of Optical4FSystemPSF for a chromatic
    spectrum (microscopes.py:334-355 with a ChromaticScalarField): per
    wavelength k the pupil spacing dx_k = required_spacing(..., lam_k); the
    point source of every wavelength, unnormalised, then normalised as one
    chromatic field (sources.py:206-208 with field.power, field.py:729-733:
    power_k = dx_k^2 sum_{y,x} sum_k' w_k' |u_k'|^2, so u_k *= sqrt(P /
    power_k)); the phase scaled by lam_0 / lam_k (phase_masks.py:105-124);
    ff_lens per wavelength; intensity sum_k w_k |out_k|^2 (field.py:735-740)
    with the density w normalised to sum 1 (Spectrum.build)."""
    shape = np.shape(phase)
    lam = np.asarray(wavelengths, np.float64).reshape(-1)
    w = np.asarray(density, np.float64).reshape(-1)
    w = w / w.sum()
    zs = np.atleast_1d(np.asarray(z, np.float64))
    dxs = [required_spacing(shape[0], output_spacing, f_tube, n, lk) for lk in lam]
    us = [np.asarray(objective_point_source(shape, (dk, dk), lk, zs, f, n, NA, None)).reshape((-1,) + shape)
          for lk, dk in zip(lam, dxs)]
    S = sum(wk * np.sum(np.abs(u) ** 2, axis=(-2, -1)) for wk, u in zip(w, us))  # per z plane
    out = 0.0
    for lk, dk, wk, u in zip(lam, dxs, w, us):
        if power is not None:
            u = u * np.sqrt(power / (dk * dk * S))[:, None, None]
        u = u * np.exp(1j * np.asarray(phase, np.float64) * (lam[0] / lk))
        o, _ = ff_lens(u, (dk, dk), lk, f_tube, n)
        out = out + wk * np.abs(o) ** 2
    return out


def psf4f(phase, spacing, wavelength, z, f, n, NA, f_tube, power=1.0):
    """Intensity stack (Z, H, W)."""
    out, _ = psf4f_field(phase, spacing, wavelength, z, f, n, NA, f_tube, power)
    return np.abs(out) ** 2


def psf4f_grad(phase, dL_dI, spacing, wavelength, z, f, n, NA, f_tube, power=1.0):
    """dL/dphase for a loss L(I) with dL/dI given, (H, W).

    With v = source * exp(i phase) and w = c * fftshift(fft2(ifftshift(v)))
    (c the lens norm), I = |w|^2: the cotangent of w is g_w = 2 dL/dI w,
    pulled back through the linear map by its adjoint
    g_v = conj(c) * fftshift(N^2 ifft2(ifftshift(g_w))) (shifts of even N are
    their own inverses; N^2 ifft2 is the adjoint of fft2), and since
    dv/dphase = i v, dL/dphase = sum_z Re(conj(g_v) i v) = sum_z Im(g_v conj(v))."""
    shape = np.shape(phase)
    src = objective_point_source(shape, spacing, wavelength, np.atleast_1d(z), f, n, NA, power)
    v = np.asarray(src).reshape((-1,) + shape) * np.exp(1j * np.asarray(phase, np.float64))
    w, _ = ff_lens(v, spacing, wavelength, f_tube, n)
    g_w = 2.0 * np.asarray(dL_dI, np.float64) * w
    c = -1j * spacing[0] * spacing[1] / (wavelength * f_tube / n)
    ax = (-2, -1)
    ny, nx = shape
    g_v = np.conj(c) * np.fft.fftshift(
        ny * nx * np.fft.ifft2(np.fft.ifftshift(g_w, axes=ax), axes=ax), axes=ax)
    return np.sum(np.imag(g_v * np.conj(v)), axis=0)


def psf4f_mse_grad(phase, target, spacing, wavelength, z, f, n, NA, f_tube, power=1.0):
    """(L, dL/dphase) for L = mean((I - target)^2)."""
    I = psf4f(phase, spacing, wavelength, z, f, n, NA, f_tube, power)
    r = I - target
    loss = np.mean(r**2)
    return loss, psf4f_grad(phase, 2.0 * r / r.size, spacing, wavelength, z, f, n, NA, f_tube, power)
