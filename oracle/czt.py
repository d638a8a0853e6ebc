"""Oracle: chirp-z transform and the CZT branch of kernel_propagate (test
infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  czt / cztn / zoomed_fft           src/chromatix/utils/czt.py:7-124
  kernel_propagate, CZT branch      src/chromatix/functional/propagation.py:395-443
  asm_propagate with output_dx      src/chromatix/functional/propagation.py:255-352
``czt_direct`` is the defining sum, evaluated by brute force; ``czt``
restates the reference's Bluestein steps (both FP64) so that each pins the
other.
"""
from __future__ import annotations

import numpy as np

from .propagation import asm_propagator, grid


def czt_direct(x, m, a, w, axis=-1):
    """y_k = sum_j x_j a^-j w^(j k)."""
    x = np.moveaxis(np.asarray(x, np.complex128), axis, -1)
    n = x.shape[-1]
    j = np.arange(n)[:, None]
    k = np.arange(m)[None, :]
    M = a ** (-j) * w ** (j * k)
    return np.moveaxis(x @ M, -1, axis)


def czt(x, m, a, w, axis=-1):
    """Bluestein, as czt.py:38-52, in FP64."""
    x = np.moveaxis(np.asarray(x, np.complex128), axis, -1)
    n = x.shape[-1]
    n_czt = m + n - 1
    k = np.arange(n_czt)
    wk2 = w ** (k**2 / 2)
    Awk2 = a ** -k[:n] * wk2[:n]
    Fwk2 = np.fft.fft(1 / np.hstack((wk2[n - 1: 0: -1], wk2[:m])), n_czt)
    y = np.fft.ifft(np.fft.fft(x * Awk2, n_czt, axis=-1) * Fwk2, axis=-1)
    y = y[..., n - 1: n + m - 1] * wk2[:m]
    return np.moveaxis(y, -1, axis)


def zoomed_fft(x, k_start, k_end, output_shape, axes=(-2, -1), include_end=True):
    renorm = [m - 1 for m in output_shape] if include_end else list(output_shape)
    for d, ax in enumerate(axes):
        w = np.exp(1j * (k_end - k_start) / renorm[d])
        x = czt_direct(x, output_shape[d], np.exp(-1j * k_start), w, ax)
    return x


def czt_propagate(u, spacing, wavelength, z, n, output_dx, output_shape, shift_yx=(0.0, 0.0),
                  kykx=(0.0, 0.0), pad=0, mode="full", remove_evanescent=False, bandlimit=False):
    """asm_propagate(..., output_dx, output_shape) with use_czt (FP64).
    Returns (u_out, dx_out, origin_out)."""
    py, px = (pad, pad) if np.isscalar(pad) else pad
    u = np.pad(np.asarray(u, np.complex128), [(0, 0)] * (np.ndim(u) - 2) + [(py, py), (px, px)])
    ny, nx = u.shape[-2:]
    H = asm_propagator((ny, nx), spacing, wavelength, z, n, kykx, bandlimit=bandlimit,
                       remove_evanescent=remove_evanescent)
    T = np.asarray(spacing, np.float64) * np.array([ny, nx])
    df = 1.0 / T
    ax = (-2, -1)
    spec = np.fft.fftshift(H * np.fft.fft2(np.fft.ifftshift(u, axes=ax), axes=ax), axes=ax)
    odx = np.broadcast_to(np.asarray(output_dx, np.float64), (2,))
    g = grid(output_shape, odx) + np.asarray(shift_yx, np.float64)
    lims = ((g[..., 0].min(), g[..., 0].max()), (g[..., 1].min(), g[..., 1].max()))
    alpha = odx / df
    for d in range(2):
        m = output_shape[d]
        lo, hi = lims[d]
        a = np.exp(-1j * 2 * np.pi / T[d] * lo)
        w = np.exp(1j * (2 * np.pi / T[d]) * (hi - lo) / (m - 1))
        spec = czt_direct(spec, m, a, w, axis=ax[d])
        N = (m - 1) // 2
        C = w ** (-N * np.arange(m)) * a**N
        spec = np.moveaxis(np.moveaxis(spec, ax[d], -1) * C, -1, ax[d])
    out = spec * np.prod(1 / alpha)
    if mode == "same":
        out = out[..., py: out.shape[-2] - py, px: out.shape[-1] - px]
    return out, odx, np.asarray(shift_yx, np.float64)
