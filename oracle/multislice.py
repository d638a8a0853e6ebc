"""Oracle: multislice propagation through a thick sample and its adjoint
(test infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  thin_sample              src/chromatix/functional/samples.py:85-134
  multislice_thick_sample  src/chromatix/functional/samples.py:137-282
  fluorescent_multislice_thick_sample
                           src/chromatix/functional/samples.py:285-441
  compute_asm_propagator   src/chromatix/functional/propagation.py:499-580
The adjoint is written out by hand (the reference differentiates with JAX
reverse mode through fori_loop) and pinned by finite differences in
tests/test_oracle_multislice.py.
"""
from __future__ import annotations

import numpy as np

from .propagation import asm_propagator, freq_1d


def _pad2(a, py, px):
    return np.pad(a, [(0, 0)] * (a.ndim - 2) + [(py, py), (px, px)])


F32 = np.float32


def sample_factor(dn, absorption, thickness, wavelength, precision="fp64"):
    """exp(i 2 pi (dn + i absorption) t / lam).  ref32: the reference's
    complex64 chain, re = fl(fl(-fl(2pi abs) t) / lam), im = fl(fl(fl(2pi dn) t) / lam),
    then the exponential of those float32 values."""
    ab = np.zeros_like(dn) if absorption is None else absorption
    if precision == "fp64":
        return np.exp(1j * 2 * np.pi * (dn + 1j * ab) * thickness / wavelength)
    tp = F32(2) * F32(np.pi)
    t, lam = F32(thickness), F32(wavelength)
    re = (-(tp * np.asarray(ab, F32)) * t) / lam
    im = ((tp * np.asarray(dn, F32)) * t) / lam
    return np.exp(re.astype(np.float64)) * np.exp(1j * im.astype(np.float64))


def na_mask(shape, spacing, wavelength, NA, precision="fp64"):
    """ifftshift(|f_grid|^2 <= (NA / lam)^2) in FFT order (samples.py:271-279)."""
    fy = freq_1d(shape[0], spacing[0], precision)[:, None]
    fx = freq_1d(shape[1], spacing[1], precision)[None, :]
    if precision == "fp64":
        return (fy**2 + fx**2) <= (NA / wavelength) ** 2
    r = F32(F32(NA) / F32(wavelength))
    return (fy * fy + fx * fx) <= r * r


def _prop(u, H):
    return np.fft.ifft2(np.fft.fft2(u, axes=(-2, -1)) * H, axes=(-2, -1))


class Multislice:
    """Forward and adjoint of multislice_thick_sample for fixed optics."""

    def __init__(self, shape, spacing, wavelength, n, thickness, depth, pad, NA=None,
                 kykx=(0.0, 0.0), reverse_distance=None, remove_evanescent=False,
                 bandlimit=False, precision="fp64"):
        self.h, self.w = shape
        self.py, self.px = (pad, pad) if np.isscalar(pad) else pad
        self.Ny, self.Nx = self.h + 2 * self.py, self.w + 2 * self.px
        self.spacing, self.lam, self.t = spacing, wavelength, thickness
        self.precision = precision
        self.H = asm_propagator((self.Ny, self.Nx), spacing, wavelength, thickness, n, kykx,
                                bandlimit=bandlimit, remove_evanescent=remove_evanescent,
                                precision=precision)
        if reverse_distance is None:
            reverse_distance = thickness * depth / 2
        Hb = asm_propagator((self.Ny, self.Nx), spacing, wavelength, -reverse_distance, n, kykx,
                            precision=precision)
        if NA is not None:
            Hb = Hb * na_mask((self.Ny, self.Nx), spacing, wavelength, NA, precision)
        self.Hb = Hb

    def _crop(self, u):
        return u[..., self.py: self.py + self.h, self.px: self.px + self.w]

    def fields(self, u_in, dn, absorption=None):
        """[u_0, u_1, ..., u_D] on the padded grid (u_{i+1} after slice i)."""
        u = _pad2(np.asarray(u_in, np.complex128), self.py, self.px)
        T = sample_factor(_pad2(np.asarray(dn, np.float64), self.py, self.px),
                          None if absorption is None else
                          _pad2(np.asarray(absorption, np.float64), self.py, self.px),
                          self.t, self.lam, self.precision)
        out = [u]
        for i in range(T.shape[0]):
            u = _prop(u, self.H) * T[i]
            out.append(u)
        return out, T

    def forward(self, u_in, dn, absorption=None, return_stack=False):
        fs, _ = self.fields(u_in, dn, absorption)
        if return_stack:
            return np.stack([self._crop(f) for f in fs[1:]])
        return self._crop(_prop(fs[-1], self.Hb))

    def grad(self, u_in, dn, cot, absorption=None, kind="upstream"):
        """(dL/d dn, dL/d absorption) for the final output.  kind "upstream":
        cot = dL/dI; "mse": cot = target, L = mean((I - target)^2); "complex":
        cot = dL/dRe(out) + i dL/dIm(out)."""
        fs, T = self.fields(u_in, dn, absorption)
        out = self._crop(_prop(fs[-1], self.Hb))
        if kind == "upstream":
            g_out = 2 * np.asarray(cot) * out
        elif kind == "mse":
            g_out = 2 * (2 * (np.abs(out) ** 2 - cot) / out.size) * out
        else:
            g_out = np.asarray(cot, np.complex128)
        g = _prop(_pad2(g_out, self.py, self.px), np.conj(self.Hb))
        k = 2 * np.pi * self.t / self.lam
        D = T.shape[0]
        gdn = np.zeros((D, self.Ny, self.Nx))
        gab = np.zeros((D, self.Ny, self.Nx))
        for i in range(D - 1, -1, -1):
            u1 = fs[i + 1]
            prod = g * np.conj(u1)
            gdn[i] = k * np.sum(np.imag(prod).reshape((-1, self.Ny, self.Nx)), axis=0)
            gab[i] = -k * np.sum(np.real(prod).reshape((-1, self.Ny, self.Nx)), axis=0)
            g = _prop(np.conj(T[i]) * g, np.conj(self.H))
        return self._crop(gdn), self._crop(gab)


def fluorescent_multislice(u, fluorescence, dn, spacing, wavelength, n, thickness, pad, phases,
                           kykx=(0.0, 0.0), remove_evanescent=False, bandlimit=False,
                           prop_fwd=None, prop_bwd=None):
    """fluorescent_multislice_thick_sample (samples.py:285-441) in FP64 with
    the random phases given, (num_samples, depth, H, W).  The steps and the
    carry of the field from one sample into the next are the reference's."""
    u = np.asarray(u, np.complex128)
    H, W = u.shape[-2:]
    D = fluorescence.shape[0]
    S = phases.shape[0]
    field = _pad2(u.reshape(H, W), pad, pad)
    Ny, Nx = field.shape
    dnp = _pad2(np.asarray(dn, np.float64), pad, pad)
    if prop_fwd is None:
        prop_fwd = asm_propagator((Ny, Nx), spacing, wavelength, thickness, n, kykx,
                                  bandlimit=bandlimit, remove_evanescent=remove_evanescent)
    if prop_bwd is None:
        prop_bwd = asm_propagator((Ny, Nx), spacing, wavelength, -thickness, n, kykx,
                                  bandlimit=bandlimit, remove_evanescent=remove_evanescent)
    out = np.zeros((D, H, W))
    for smp in range(S):
        fl = _pad2(np.asarray(fluorescence, np.float64) * np.exp(1j * phases[smp]), pad, pad)
        for i in range(D):
            field = field * np.exp(1j * 2 * np.pi * dnp[i] * thickness / wavelength)
            field = np.fft.ifft2(np.fft.fft2(fl[i] + field) * prop_fwd)
        field = np.fft.fft2(field)
        for i in range(D):
            field = field * prop_bwd
            fi = np.fft.ifft2(field)[pad:pad + H, pad:pad + W]
            out[D - 1 - i] += np.abs(fi) ** 2
    return out / S
