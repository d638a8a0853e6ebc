"""Oracle: chirp-z transform and the CZT branch of kernel_propagate (test
infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  czt / cztn / zoomed_fft           src/chromatix/utils/czt.py:7-124
  kernel_propagate, CZT branch      src/chromatix/functional/propagation.py:395-443
  kernel_propagate, scaled-convolution branch (Yu et al. 2012)
                                    src/chromatix/functional/propagation.py:444-458
  asm_propagate with output_dx      src/chromatix/functional/propagation.py:255-352
  high_na_ff_lens (scalar fields)   src/chromatix/functional/lenses.py:98-156
``czt_direct`` is the defining sum, evaluated by brute force; ``czt``
restates the reference's Bluestein steps (both FP64) so that each pins the
other.  ``yu_axis_direct`` evaluates the scaled convolution as an explicit
double sum; ``yu_kernel_propagate`` follows the reference's steps with an
FFT convolution (scipy.signal.fftconvolve, mode "same"), so each pins the
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


def czt_kernel_propagate(u, H, spacing, output_dx, output_shape, shift_yx=(0.0, 0.0)):
    """The CZT branch of kernel_propagate (propagation.py:395-443) for a
    propagator H in FFT order, by the defining sums (FP64).  Returns
    (u_out, dx_out, origin_out)."""
    u = np.asarray(u, np.complex128)
    ny, nx = u.shape[-2:]
    T = np.broadcast_to(np.asarray(spacing, np.float64), (2,)) * np.array([ny, nx])
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
    return spec * np.prod(1 / alpha), odx, np.asarray(shift_yx, np.float64)


def czt_propagate(u, spacing, wavelength, z, n, output_dx, output_shape, shift_yx=(0.0, 0.0),
                  kykx=(0.0, 0.0), pad=0, mode="full", remove_evanescent=False, bandlimit=False):
    """asm_propagate(..., output_dx, output_shape) with use_czt (FP64).
    Returns (u_out, dx_out, origin_out)."""
    py, px = (pad, pad) if np.isscalar(pad) else pad
    u = np.pad(np.asarray(u, np.complex128), [(0, 0)] * (np.ndim(u) - 2) + [(py, py), (px, px)])
    ny, nx = u.shape[-2:]
    H = asm_propagator((ny, nx), spacing, wavelength, z, n, kykx, bandlimit=bandlimit,
                       remove_evanescent=remove_evanescent)
    out, odx, origin = czt_kernel_propagate(u, H, spacing, output_dx, output_shape, shift_yx)
    if mode == "same":
        out = out[..., py: out.shape[-2] - py, px: out.shape[-1] - px]
    return out, odx, origin


def _freq_1d(n, d):
    """fftshift(fftfreq(n)) / d (field.py:36-46)."""
    return np.fft.fftshift(np.fft.fftfreq(n)) / d


def yu_axis_direct(B, alpha, f_in, x_out, out_dx, axis):
    """One axis of the scaled convolution as an explicit double sum:
    y_j = out_dx exp(i pi x_j^2 / alpha) sum_k B_k (1/alpha) exp(i pi alpha
    f_k^2) kern[j + s - k], kern[m] = exp(-i pi alpha f_m^2) for 0 <= m < n
    (0 outside), s = (n - 1) // 2 (the "same" window of the full linear
    convolution)."""
    B = np.moveaxis(np.asarray(B, np.complex128), axis, -1)
    n = B.shape[-1]
    s = (n - 1) // 2
    pre = (1.0 / alpha) * np.exp(1j * np.pi * alpha * f_in**2)
    kern = np.exp(-1j * np.pi * alpha * f_in**2)
    j = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    m = j + s - k
    K = np.where((m >= 0) & (m < n), kern[np.clip(m, 0, n - 1)], 0.0)  # (j, k)
    y = (B * pre) @ K.T
    y = y * (out_dx * np.exp(1j * np.pi * x_out**2 / alpha))
    return np.moveaxis(y, -1, axis)


def yu_kernel_propagate(u, H, spacing, output_dx, shift_yx=(0.0, 0.0), direct=False):
    """kernel_propagate(field, H, output_dx=..., use_czt=False) for a field at
    the origin (propagation.py:395-410, 444-458), FP64.  The output shape is
    the input shape (the reference's product with the output-grid modulation
    broadcasts only then).  ``direct`` evaluates the convolutions by
    yu_axis_direct instead of FFTs."""
    from scipy.signal import fftconvolve

    u = np.asarray(u, np.complex128)
    ny, nx = u.shape[-2:]
    sp = np.asarray(spacing, np.float64)
    odx = np.broadcast_to(np.asarray(output_dx, np.float64), (2,))
    T = sp * np.array([ny, nx])
    df = 1.0 / T
    alpha = odx / df
    ax = (-2, -1)
    spec = np.fft.fftshift(H * np.fft.fft2(np.fft.ifftshift(u, axes=ax), axes=ax), axes=ax)
    fy, fx = _freq_1d(ny, sp[0]), _freq_1d(nx, sp[1])
    g = grid((ny, nx), odx) + np.asarray(shift_yx, np.float64)
    xy, xx = g[:, 0, 0], g[0, :, 1]
    if direct:
        out = yu_axis_direct(spec, alpha[0], fy, xy, odx[0], -2)
        return yu_axis_direct(out, alpha[1], fx, xx, odx[1], -1)
    # the reference's steps: B = u * prod((1/alpha) exp(i pi / alpha wn^2)),
    # f = prod(exp(-i pi / alpha wn^2)), wn = alpha f_grid;
    # u = prod(dx exp(i pi / alpha grid^2)) * fftconvolve(B, f, "same")
    wy, wx = alpha[0] * fy, alpha[1] * fx
    pre = ((1 / alpha[0]) * np.exp(1j * np.pi / alpha[0] * wy**2))[:, None] * \
          ((1 / alpha[1]) * np.exp(1j * np.pi / alpha[1] * wx**2))[None, :]
    f = np.exp(-1j * np.pi / alpha[0] * wy**2)[:, None] * np.exp(-1j * np.pi / alpha[1] * wx**2)[None, :]
    mod = (odx[0] * np.exp(1j * np.pi / alpha[0] * xy**2))[:, None] * \
          (odx[1] * np.exp(1j * np.pi / alpha[1] * xx**2))[None, :]
    B = spec * pre
    conv = fftconvolve(B, np.broadcast_to(f, B.shape), mode="same", axes=ax)
    return mod * conv


def high_na_ff_lens(u, spacing, wavelength, f, n, NA, output_shape=None, output_dx=None):
    """high_na_ff_lens for a scalar, square, single-wavelength field, FP64:
    u * defocus on the input frequency grid, then zoomed_fft over
    [-zoom pi, zoom pi] with include_end (lenses.py:98-156).  Returns
    (u_out, dx_out)."""
    u = np.asarray(u, np.complex128)
    ny, nx = u.shape[-2:]
    sp = np.asarray(spacing, np.float64)
    if output_dx is None:
        output_dx = sp[0]
    if output_shape is None:
        output_shape = (ny, nx)
    fov_out = output_shape[0] * output_dx
    zoom = 2 * NA * fov_out / ((nx - 1) * wavelength)
    fy, fx = _freq_1d(ny, sp[0]), _freq_1d(nx, sp[1])
    s2 = (fy[:, None] * wavelength / n) ** 2 + (fx[None, :] * wavelength / n) ** 2
    sz = np.sqrt(np.maximum(1 - NA**2 * s2, 0.0))
    k = 2 * np.pi * n / wavelength
    with np.errstate(divide="ignore", invalid="ignore"):
        defocus = np.where(sz != 0.0, np.exp(1j * k * sz * f) / sz, 0.0)
    out = zoomed_fft(u * defocus, -zoom * np.pi, zoom * np.pi, output_shape, include_end=True)
    return out, np.array([output_dx, output_dx], np.float64)
