"""The output-resampling branches of kernel_propagate and the
output_dx / output_shape modes of transfer_propagate / asm_propagate
(src/chromatix/functional/propagation.py:168-352, 395-458).

u = fftshift(H * fft2(ifftshift(u_in))) is formed by the shifted raw FFT
(two passes) and one multiply by fftshift(H).  With use_czt, one Bluestein
pass per axis (cxm_czt_axis_c64) evaluates the field on the output grid,
with the reference's per-axis modulation w^(-N k) a^N, N = (m - 1) // 2,
and the scale prod(1 / alpha), alpha = output_dx / input_df.  Without it
(the scaled convolution of Yu et al. 2012), one pass per axis
(cxm_yu_axis_c64) applies the pre-chirp (1/alpha) e^{i pi alpha f^2}, the
"same" linear convolution with the finite chirp e^{-i pi alpha f^2} and the
output modulation dx e^{i pi x^2 / alpha}; the output keeps the input's
shape, as the reference's broadcast of the output-grid modulation requires.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field, ScalarField, crop, pad, shift_grid
from ..utils.czt import czt_axis, scaled_conv_axis, use_zoom_mfma, zoom_dft
from .optics import _raw as _raw_fft
from .optics import fft as _fft_shifted
from .propagation import (
    _pair,
    compute_asm_propagator,
    compute_transfer_propagator,
)

__all__ = ["czt_kernel_propagate", "czt_kernel_propagate_fields"]


def _shifted_mul(a, b):
    """a * fftshift(b) per plane (plane p of a uses b[p % len(b)])."""
    a = a.to(torch.complex64).resolve_conj().contiguous()
    ny, nx = a.shape[-2:]
    out = torch.empty_like(a)
    L.check(L.lib().cxm_cmul_c64(L.stream_ptr(a.device), a.data_ptr(), b.data_ptr(), out.data_ptr(),
                                 a.numel() // (ny * nx), ny, nx, b.shape[0], 1))
    return out


class _ShiftedMulFn(torch.autograd.Function):
    """The propagator multiply of the CZT branch, spec * fftshift(H); its
    reverse-mode rule in spec multiplies by conj(fftshift(H)) (H is a
    constant of the rule: a propagator that requires grad is refused)."""

    @staticmethod
    def forward(ctx, spec, H):
        ctx.save_for_backward(H)
        return _shifted_mul(spec, H)

    @staticmethod
    def backward(ctx, g):
        (H,) = ctx.saved_tensors
        return _shifted_mul(g, torch.conj_physical(H)), None


def czt_kernel_propagate(field: Field, propagator, output_dx=None, output_shape=None,
                         shift_yx=(0.0, 0.0), use_czt: bool = True) -> Field:
    """kernel_propagate with an output grid of spacing ``output_dx`` and shape
    ``output_shape`` centred on ``shift_yx`` (propagation.py:395-443)."""
    L.refuse_grad("kernel_propagate with an output grid (gradient with respect to the propagator)",
                  propagator)
    if isinstance(field, ChromaticScalarField):
        return _czt_chromatic(field, propagator, output_dx, output_shape, shift_yx, use_czt)
    ny, nx = field.spatial_shape
    if output_shape is None:
        output_shape = (ny, nx)
    output_shape = (int(output_shape[0]), int(output_shape[1]))
    if use_czt and min(output_shape) < 2:
        raise ValueError("output_shape must be at least 2 x 2")
    if not use_czt and output_shape != (ny, nx):
        # the reference multiplies the output-grid modulation (output_shape)
        # with the convolution (input shape): only equal shapes broadcast
        raise ValueError("the scaled-convolution branch keeps the input shape: output_shape "
                         f"{output_shape} != {(ny, nx)}")
    out_dx = field.dx if output_dx is None else np.broadcast_to(
        np.asarray(output_dx, dtype=np.float64), (2,)).copy()
    T = field.dx * np.array([ny, nx], dtype=np.float64)  # input extent
    df = 1.0 / T
    H = torch.as_tensor(propagator).to(device=field.u.device, dtype=torch.complex64)
    if H.shape[-2:] != (ny, nx):
        raise ValueError("propagator must have the field's spatial shape")
    H = H.reshape((-1, ny, nx)).resolve_conj().contiguous()
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    if H.shape[0] == 1 and not (torch.is_grad_enabled() and u.requires_grad):
        # the propagator multiply on the shifted transform's column store
        spec = _raw_fft(u, (-2, -1), True, False, spec_mul=H[0])
    else:
        spec = _ShiftedMulFn.apply(_fft_shifted(u, axes=(-2, -1), shift=True), H)
    out_field = Field.empty_like(shift_grid(field, shift_yx), dx=out_dx, shape=output_shape)
    alpha = out_field.dx / df
    if not use_czt:
        # first output coordinate per axis: (0 - n/2) dx + origin (field.py:23-33)
        origin = np.broadcast_to(np.asarray(out_field.origin, dtype=np.float64), (2,))
        res = spec
        for d in range(2):
            n_d = output_shape[d]
            x0 = -(n_d / 2.0) * float(out_field.dx[d]) + float(origin[d])
            res = scaled_conv_axis(res, d, float(alpha[d]), float(df[d]), float(out_field.dx[d]), x0)
        return out_field.replace(u=res)
    (y0, y1), (x0, x1) = out_field.spatial_limits
    lims = ((y0, y1), (x0, x1))
    pars = []
    for d in range(2):
        m = output_shape[d]
        lo, hi = lims[d]
        a = np.exp(-1j * 2 * np.pi / T[d] * lo)
        w = np.exp(1j * (2 * np.pi / T[d]) * (hi - lo) / (m - 1))
        scale = float(np.prod(1.0 / alpha)) if d == 1 else 1.0
        pars.append((m, a, w, (m - 1) // 2, scale))
    if use_zoom_mfma(output_shape):
        # a small window: both CZTs with their modulation as two complex GEMMs
        res = zoom_dft(spec, output_shape, (pars[0][1], pars[1][1]), (pars[0][2], pars[1][2]),
                       mod_n=(pars[0][3], pars[1][3]), scale=(pars[0][4], pars[1][4]))
        return out_field.replace(u=res)
    res = spec
    for d in range(2):
        m, a, w, nmod, scale = pars[d]
        res = czt_axis(res, m, a, w, d, mod_n=nmod, scale=scale)
    return out_field.replace(u=res)


def _czt_chromatic(field: ChromaticScalarField, propagator, output_dx, output_shape, shift_yx,
                   use_czt) -> ChromaticScalarField:
    """The output-grid branch per wavelength: each wavelength on its own
    input grid (dx[w], extent dx[w] N) and output grid (output_dx[w], or
    output_dx for all), a propagator with a trailing wavelength axis taken
    per wavelength.  The reference's branch indexes the field extent by axis
    (T[d], propagation.py:420-424), which for a (wv, 2) extent of a
    chromatic field selects a wavelength instead of an axis; per wavelength
    the result is the monochromatic one of the same grid."""
    nwv = field.spectrum.size
    ny, nx = field.spatial_shape
    H = torch.as_tensor(propagator)
    per_wv = H.ndim >= 3 and H.shape[-1] == nwv and tuple(H.shape[-3:-1]) == (ny, nx)
    if output_dx is None:
        odx = field.dx.copy()
    else:
        o = np.asarray(output_dx, dtype=np.float64)
        if o.ndim == 2 and o.shape != (nwv, 2):
            raise ValueError(f"output_dx of a chromatic field: (2,) or ({nwv}, 2), got {o.shape}")
        odx = np.broadcast_to(o if o.ndim == 2 else np.broadcast_to(o, (2,)), (nwv, 2)).copy()
    outs, first = [], None
    for w in range(nwv):
        mono = ScalarField(field.u[..., w], field.dx[w], field.origin[w],
                           float(field.spectrum.wavelength[w]))
        out = czt_kernel_propagate(mono, H[..., w] if per_wv else H, odx[w], output_shape, shift_yx,
                                   use_czt)
        if first is None:
            first = out
        outs.append(out.u)
    u = torch.stack(outs, dim=-1)
    return field.replace(u=u, dx=odx, origin=np.broadcast_to(first.origin, (nwv, 2)).copy())


def czt_kernel_propagate_fields(method: str, field: Field, z, n, pad_width, cval, kykx, shift_yx,
                                output_dx, output_shape, use_czt, mode, remove_evanescent=False,
                                bandlimit=False, *, phase_mode=None) -> Field:
    """transfer_propagate / asm_propagate with output_dx or output_shape:
    pad, propagator on the padded grid (no output shift when the CZT places
    the window), CZT branch of kernel_propagate, crop in "same" mode
    (propagation.py:229-252, 325-352)."""
    py, px = _pair(pad_width)
    padded = pad(field, (py, px), cval=cval)
    z = float(np.asarray(z).reshape(()))
    if method == "asm":
        H = compute_asm_propagator(padded, z, n, kykx, bandlimit,
                                   (0.0, 0.0) if use_czt else shift_yx,
                                   remove_evanescent, phase_mode=phase_mode)
    else:
        H = compute_transfer_propagator(padded, z, n, kykx, phase_mode=phase_mode)
    out = czt_kernel_propagate(padded, H, output_dx, output_shape, shift_yx, use_czt)
    if mode == "same":
        out = crop(out, (py, px))
    return out
