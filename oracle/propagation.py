"""Oracle: AS / transfer propagation (test infrastructure, see oracle/__init__).

Follows src/chromatix/functional/propagation.py (paths relative to the
chromatix repository root):
  compute_transfer_propagator  :466-496
  compute_asm_propagator       :499-580
  kernel_propagate (default)   :391-394
  transfer_propagate           :168-252
  asm_propagate                :255-352
and the frequency grid of src/chromatix/core/field.py:36-46.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

F32 = np.float32


def _fft2(x, inverse=False, workers=1):
    if inverse:
        return sfft.ifft2(x, axes=(-2, -1), workers=workers)
    return sfft.fft2(x, axes=(-2, -1), workers=workers)


# ---- frequency grids ----------------------------------------------------------
def freq_1d(n: int, d: float, precision: str = "fp64") -> np.ndarray:
    """Frequencies of the FFT bins in natural (unshifted) order.

    fp64: fftfreq(n)/d.  ref32: the reference's float32 chain
    fl(fl(1/d) * fl(k/n)) (field.py:40-46: fftfreq = k / n in float32, then
    multiplied by dk = 1/spacing)."""
    k = np.fft.fftfreq(n) * n  # exact signed bin numbers
    if precision == "fp64":
        return k / (n * float(d))
    return F32(F32(1.0) / F32(d)) * (F32(k) / F32(n))


# ---- propagators (FFT order) ---------------------------------------------------
def _zs(z):
    z = np.asarray(z, dtype=np.float64)
    return z.reshape(-1), z.ndim > 0 and z.size > 1


def transfer_propagator(shape, spacing, wavelength, z, n, kykx=(0.0, 0.0), precision="fp64"):
    """H = exp(-i pi lam/n z |f - kykx|^2) (propagation.py:489-496).

    ref32 order: fl(fl(fl(fl(-pi*lam)/n)*z) * l2), l2 = fl(fl(fy'^2)+fl(fx'^2)).
    Returns (ny, nx) or (Z, ny, nx)."""
    ny, nx = shape
    dy, dx = spacing
    zz, batched = _zs(z)
    out = []
    for zi in zz:
        if precision == "fp64":
            fy = freq_1d(ny, dy)[:, None] - kykx[0]
            fx = freq_1d(nx, dx)[None, :] - kykx[1]
            ph = -np.pi * wavelength / n * zi * (fy**2 + fx**2)
            out.append(np.exp(1j * ph))
        else:
            fy = freq_1d(ny, dy, "ref32")[:, None] - F32(kykx[0])
            fx = freq_1d(nx, dx, "ref32")[None, :] - F32(kykx[1])
            l2 = fy * fy + fx * fx
            c = F32(F32(F32(F32(-np.pi) * F32(wavelength)) / F32(n)) * F32(zi))
            ph = c * l2
            out.append(_cis_f32(ph))
    H = np.stack(out)
    return H if batched else H[0]


def _cis_f32(ph32, amp=None):
    """exp(i*phase) of an exact float32 phase, rounded once to complex64."""
    p = ph32.astype(np.float64)
    v = np.cos(p) + 1j * np.sin(p)
    if amp is not None:
        v = v * amp
    return v.astype(np.complex64)


def _band_limit(shape, spacing, wavelength, z, shift_yx, precision):
    """Rect filter of Matsushima & Shimobaba (propagation.py:557-579)."""
    ny, nx = shape
    masks = []
    for ax, (nn, d) in enumerate(((ny, spacing[0]), (nx, spacing[1]))):
        s = shift_yx[ax]
        if precision == "fp64":
            ext = d * nn
            dfa = 1.0 / ext
            kp = ((s + 1 / (2 * dfa)) ** (-2) * z**2 + 1) ** (-0.5) / wavelength
            kn = ((s - 1 / (2 * dfa)) ** (-2) * z**2 + 1) ** (-0.5) / wavelength
            sgp, sgn = np.sign(s + ext), np.sign(s - ext)
            k0 = 0.5 * (sgp * kp + sgn * kn)
            kmax = (sgp * kp - sgn * kn) / 2
            f = freq_1d(nn, d)
        else:
            ext = F32(F32(d) * F32(nn))
            dfa = F32(F32(1) / ext)
            half = F32(F32(1) / F32(F32(2) * dfa))
            sp, sn = F32(F32(s) + half), F32(F32(s) - half)
            zf = F32(z)
            kp = F32(F32(1.0 / np.sqrt(np.float64(F32(F32(F32(1) / F32(sp * sp)) * F32(zf * zf)) + F32(1))))
                     / F32(wavelength))
            kn = F32(F32(1.0 / np.sqrt(np.float64(F32(F32(F32(1) / F32(sn * sn)) * F32(zf * zf)) + F32(1))))
                     / F32(wavelength))
            sgp, sgn = F32(np.sign(F32(s) + ext)), F32(np.sign(F32(s) - ext))
            k0 = F32(F32(0.5) * F32(sgp * kp + sgn * kn))
            kmax = F32(F32(sgp * kp - sgn * kn) / F32(2))
            f = freq_1d(nn, d, "ref32")
        masks.append(np.abs(f - k0) <= kmax)
    return masks[0][:, None] & masks[1][None, :]


def asm_propagator(shape, spacing, wavelength, z, n, kykx=(0.0, 0.0), bandlimit=False,
                   shift_yx=(0.0, 0.0), remove_evanescent=False, precision="fp64"):
    """Angular-spectrum H in FFT order (propagation.py:537-580).

    kernel = 1 - (lam/n)^2 |f - kykx|^2; delay = sqrt(max(kernel,0)) if
    remove_evanescent else the complex sqrt; phase = 2 pi (|z| n/lam) delay +
    2 pi f.shift; H = exp(i phase), conjugated for z < 0; optional band limit.
    Returns (ny, nx) or (Z, ny, nx)."""
    ny, nx = shape
    dy, dx = spacing
    zz, batched = _zs(z)
    out = []
    for zi in zz:
        if precision == "fp64":
            fy0 = freq_1d(ny, dy)[:, None]
            fx0 = freq_1d(nx, dx)[None, :]
            l2 = (fy0 - kykx[0]) ** 2 + (fx0 - kykx[1]) ** 2
            kern = 1 - (wavelength / n) ** 2 * l2
            if remove_evanescent:
                delay = np.sqrt(np.maximum(kern, 0.0)).astype(np.complex128)
            else:
                delay = np.sqrt(kern.astype(np.complex128))
            osh = 2 * np.pi * (fy0 * shift_yx[0] + fx0 * shift_yx[1])
            phase = 2 * np.pi * (abs(zi) * n / wavelength) * delay + osh
            H = np.exp(1j * phase)
            if zi < 0:
                H = np.conj(H)
        else:
            fy0 = freq_1d(ny, dy, "ref32")[:, None]
            fx0 = freq_1d(nx, dx, "ref32")[None, :]
            fy = fy0 - F32(kykx[0])
            fx = fx0 - F32(kykx[1])
            l2 = fy * fy + fx * fx
            ln = F32(F32(wavelength) / F32(n))
            kern = F32(1) - F32(ln * ln) * l2
            A = F32(F32(2 * np.pi) * F32(F32(F32(abs(F32(zi))) * F32(n)) / F32(wavelength)))
            osh = F32(2 * np.pi) * (fy0 * F32(shift_yx[0]) + fx0 * F32(shift_yx[1]))
            prop = (kern >= 0) | bool(remove_evanescent)
            d_re = np.sqrt(np.maximum(kern, F32(0)))
            d_im = np.sqrt(np.maximum(-kern, F32(0)))
            re = np.where(prop, A * d_re + osh, osh).astype(F32)
            amp = np.where(prop, 1.0, np.exp(-(A * d_im).astype(np.float64)))
            H = _cis_f32(re, amp)
            if zi < 0:
                H = np.conj(H)
        if bandlimit:
            H = H * _band_limit(shape, spacing, wavelength, zi, shift_yx, precision)
        out.append(H)
    H = np.stack(out)
    return H if batched else H[0]


# ---- propagation ----------------------------------------------------------------
def kernel_propagate(u, H, precision="fp64", workers=1):
    """ifft2(fft2(u) * H) over the last two axes (propagation.py:391-394)."""
    if precision == "fp64":
        return _fft2(_fft2(np.asarray(u, np.complex128), workers=workers) * H, True, workers)
    u32 = np.asarray(u, np.complex64)
    return _fft2(_fft2(u32, workers=workers) * H.astype(np.complex64), True, workers).astype(np.complex64)


def _pad(u, py, px):
    return np.pad(u, [(0, 0)] * (u.ndim - 2) + [(py, py), (px, px)])


def _crop(u, cy, cx):
    ny, nx = u.shape[-2:]
    return u[..., cy: ny - cy, cx: nx - cx]


def propagate(u, spacing, wavelength, z, n, pad_width, method="asm", mode="same",
              kykx=(0.0, 0.0), shift_yx=(0.0, 0.0), bandlimit=False, remove_evanescent=False,
              precision="fp64", workers=1, fft_precision=None):
    """transfer_propagate / asm_propagate for a scalar field u (... H W):
    pad, propagator on the padded grid, kernel_propagate, crop in "same"
    mode.  A z array of length Z gives (Z, ...) outputs (one field, or a
    field whose innermost batch is Z, as numpy broadcasting allows).
    fft_precision="fp64" with precision="ref32" isolates the propagator's
    float32 rounding (exact FFTs around the reference-order H)."""
    py, px = (pad_width, pad_width) if np.isscalar(pad_width) else pad_width
    up = _pad(np.asarray(u), py, px)
    shape = up.shape[-2:]
    if method == "asm":
        H = asm_propagator(shape, spacing, wavelength, z, n, kykx, bandlimit, shift_yx,
                           remove_evanescent, precision)
    else:
        H = transfer_propagator(shape, spacing, wavelength, z, n, kykx, precision)
    out = kernel_propagate(up, H, fft_precision or precision, workers)
    if mode == "same":
        out = _crop(out, py, px)
    return out


# ---- closed forms from the reference tests ----------------------------------------
def square_aperture_field(x, z, w, wavelength, n):
    """Fresnel diffraction of a square aperture of width w, unit plane wave
    (closed form of tests/test_propagate.py:17-33, via Fresnel integrals)."""
    from scipy.special import fresnel

    nf = (w / 2) ** 2 / (wavelength / n * z)

    def factor(xi):
        smin, cmin = fresnel(np.sqrt(2 * nf) * (1 - 2 * xi / w))
        splus, cplus = fresnel(np.sqrt(2 * nf) * (1 + 2 * xi / w))
        return (cmin + cplus) / np.sqrt(2) + 1j * (smin + splus) / np.sqrt(2)

    U = np.exp(1j * 2 * np.pi * z * n / wavelength) / 1j * factor(x[..., 0]) * factor(x[..., 1])
    return U / w


def grid(shape, spacing):
    """Sample coordinates (i - N/2) dx (field.py:23-33), shape (ny, nx, 2)."""
    ny, nx = shape
    gy, gx = np.meshgrid(np.arange(ny) - ny / 2, np.arange(nx) - nx / 2, indexing="ij")
    return np.stack([gy * spacing[0], gx * spacing[1]], axis=-1)
