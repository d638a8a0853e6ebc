"""Oracle: optical FFT, transform and SAS propagation (test infrastructure).

Follows (paths relative to the chromatix repository root):
  optical_fft                  src/chromatix/functional/convenience.py:9-59
  transform_propagate          src/chromatix/functional/propagation.py:34-82
  compute_sas_precompensation  src/chromatix/functional/propagation.py:85-103
  transform_propagate_sas      src/chromatix/functional/propagation.py:106-165
  circular_pupil               src/chromatix/functional/pupils.py:17-23
Fields are scalar (... H W); spacing is (dy, dx).  precision "fp64" uses
exact inputs; "ref32" repeats the reference's float32 chains for the
quantities it computes in float32 (norms, chirp phases, the SAS phase) and
keeps the FFTs exact, so that a GPU result in ref32 mode differs from it by
FFT rounding only.
"""
from __future__ import annotations

import numpy as np

from .propagation import _cis_f32, freq_1d, grid

F32 = np.float32


def _fft_c(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=(-2, -1)), axes=(-2, -1)), axes=(-2, -1))


def _grid_f32(shape, spacing, origin=(0.0, 0.0)):
    ny, nx = shape
    gy = F32(F32(spacing[0]) * (np.arange(ny, dtype=F32) - F32(ny / 2))) + F32(origin[0])
    gx = F32(F32(spacing[1]) * (np.arange(nx, dtype=F32) - F32(nx / 2))) + F32(origin[1])
    return gy[:, None], gx[None, :]


def new_spacing(shape, spacing, wavelength, z, n):
    L2 = abs(wavelength * z / n)
    return (L2 / (shape[0] * spacing[0]), L2 / (shape[1] * spacing[1]))


def optical_fft(u, spacing, wavelength, z, n, precision="fp64"):
    """Returns (u', du).  z >= 0: fftshift(fft2(ifftshift(-i dy dx / L2 u)));
    z < 0: conj(fftshift(fft2(ifftshift(-i L2 / (du_y du_x N) conj(u)))))."""
    u = np.asarray(u, np.complex128)
    shape = u.shape[-2:]
    du = new_spacing(shape, spacing, wavelength, z, n)
    if precision == "fp64":
        L2 = wavelength * z / n
        if z >= 0:
            norm = -1j * spacing[0] * spacing[1] / L2
        else:
            norm = -1j * (L2 / (du[0] * du[1])) / (shape[0] * shape[1])
    else:
        L2 = F32(F32(F32(wavelength) * F32(z)) / F32(n))
        if z >= 0:
            norm = -1j * np.float64(F32(F32(F32(spacing[0]) * F32(spacing[1])) / L2))
        else:
            dfy = F32(F32(1) / F32(F32(spacing[0]) * F32(shape[0])))
            dfx = F32(F32(1) / F32(F32(spacing[1]) * F32(shape[1])))
            pdu = F32(F32(dfy * abs(L2)) * F32(dfx * abs(L2)))
            norm = -1j * np.float64(F32(F32(L2 / pdu) / F32(shape[0] * shape[1])))
    if z >= 0:
        return _fft_c(norm * u), du
    return np.conj(_fft_c(norm * np.conj(u))), du


def chirp(shape, spacing, wavelength, z, n, origin=(0.0, 0.0), precision="fp64"):
    """exp(i pi |g|^2 / L2) on field.grid (propagation.py:75, :80)."""
    if precision == "fp64":
        g = grid(shape, spacing) + np.asarray(origin)
        L2 = wavelength * z / n
        return np.exp(1j * np.pi / L2 * (g[..., 0] ** 2 + g[..., 1] ** 2))
    L2 = F32(F32(F32(wavelength) * F32(z)) / F32(n))
    a = F32(F32(np.pi) / L2)
    gy, gx = _grid_f32(shape, spacing, origin)
    return _cis_f32(a * (gy * gy + gx * gx))


def transform_propagate(u, spacing, wavelength, z, n, pad_width, skip_initial_phase=False,
                        skip_final_phase=False, precision="fp64"):
    """Returns (u', du): pad, Q1 chirp, 1j * optical_fft, Q2 chirp, crop."""
    py, px = (pad_width, pad_width) if np.isscalar(pad_width) else pad_width
    up = np.pad(np.asarray(u, np.complex128), [(0, 0)] * (np.ndim(u) - 2) + [(py, py), (px, px)])
    shape = up.shape[-2:]
    if not skip_initial_phase:
        up = up * chirp(shape, spacing, wavelength, z, n, precision=precision)
    out, du = optical_fft(up, spacing, wavelength, z, n, precision)
    out = 1j * out
    if not skip_final_phase:
        du32 = du
        if precision == "ref32":
            L2 = F32(F32(F32(wavelength) * F32(z)) / F32(n))
            du32 = (F32(F32(F32(1) / F32(F32(spacing[0]) * F32(shape[0]))) * abs(L2)),
                    F32(F32(F32(1) / F32(F32(spacing[1]) * F32(shape[1]))) * abs(L2)))
        out = out * chirp(shape, du32, wavelength, z, n, precision=precision)
    ny, nx = shape
    return out[..., py: ny - py, px: nx - px], du


def sas_precompensation(shape, spacing, wavelength, z, n, precision="fp64"):
    """ifftshift(W exp(i kz (H_AS - H_Fr))) on a grid of `shape` (the padded
    field), s = lam f / n, t = L/2/|z| + |s| with L = 2 * extent."""
    if precision == "fp64":
        fy = freq_1d(shape[0], spacing[0])[:, None]
        fx = freq_1d(shape[1], spacing[1])[None, :]
        sy, sx = wavelength * fy / n, wavelength * fx / n
        ty = 2 * spacing[0] * shape[0] / 2 / abs(z) + np.abs(sy)
        tx = 2 * spacing[1] * shape[1] / 2 / abs(z) + np.abs(sx)
        W = (sy**2 * (2 + 1 / ty**2) <= 1) & (sx**2 * (2 + 1 / tx**2) <= 1)
        s2 = sy**2 + sx**2
        kz = 2 * z * np.pi * n / wavelength
        return W * np.exp(1j * kz * (np.sqrt(np.maximum(0, 1 - s2)) - (1 - s2 / 2)))
    fy = freq_1d(shape[0], spacing[0], "ref32")[:, None]
    fx = freq_1d(shape[1], spacing[1], "ref32")[None, :]
    lam, nf, zf = F32(wavelength), F32(n), F32(z)
    sy, sx = (lam * fy) / nf, (lam * fx) / nf
    sy2, sx2 = sy * sy, sx * sx
    ty = F32(F32(F32(F32(2) * F32(F32(spacing[0]) * F32(shape[0]))) / F32(2)) / abs(zf)) + np.abs(sy)
    tx = F32(F32(F32(F32(2) * F32(F32(spacing[1]) * F32(shape[1]))) / F32(2)) / abs(zf)) + np.abs(sx)
    W = (sy2 * (F32(2) + F32(1) / (ty * ty)) <= 1) & (sx2 * (F32(2) + F32(1) / (tx * tx)) <= 1)
    s2 = sy2 + sx2
    has = np.sqrt(np.maximum(F32(0), F32(1) - s2))
    hfr = F32(1) - s2 / F32(2)
    kz = F32(F32(F32(F32(F32(2) * zf) * F32(np.pi)) * nf) / lam)
    return W * _cis_f32(kz * (has - hfr)).astype(np.complex128)


def transform_propagate_sas(u, spacing, wavelength, z, n, skip_initial_phase=False,
                            skip_final_phase=False, precision="fp64"):
    """Returns (u', du).  Pad by half the size per side; z >= 0: dH then
    transform; z < 0: transform then dH (on the new grid); crop."""
    u = np.asarray(u, np.complex128)
    H, W = u.shape[-2:]
    py, px = H // 2, W // 2
    up = np.pad(u, [(0, 0)] * (u.ndim - 2) + [(py, py), (px, px)])
    shape = up.shape[-2:]

    def kprop(x, d):
        dH = sas_precompensation(shape, d, wavelength, z, n, precision)
        return np.fft.ifft2(np.fft.fft2(x, axes=(-2, -1)) * dH, axes=(-2, -1))

    if z >= 0:
        v = kprop(up, spacing)
        v, du = transform_propagate(v, spacing, wavelength, z, n, 0, skip_initial_phase,
                                    skip_final_phase, precision)
    else:
        v, du = transform_propagate(up, spacing, wavelength, z, n, 0, skip_initial_phase,
                                    skip_final_phase, precision)
        v = kprop(v, du)
    return v[..., py: py + H, px: px + W], du


def circular_mask(shape, spacing, diameter, origin=(0.0, 0.0)):
    g = grid(shape, spacing) + np.asarray(origin)
    return (g[..., 0] ** 2 + g[..., 1] ** 2) <= (diameter / 2) ** 2


def ff_lens(u, spacing, wavelength, f, n, NA=None, inverse=False, precision="fp64"):
    """Optional circular pupil of diameter 2 f NA / n, then optical_fft(+-f)
    (src/chromatix/functional/lenses.py:64-95)."""
    u = np.asarray(u, np.complex128)
    if NA is not None:
        u = u * circular_mask(u.shape[-2:], spacing, 2 * f * NA / n)
    return optical_fft(u, spacing, wavelength, -f if inverse else f, n, precision)
