"""Intensity and power of a Field through libcxm kernels.

Mirrors ``ScalarField.intensity/power`` and
``ChromaticScalarField.intensity/power`` (src/chromatix/core/field.py:656-665,
728-740): intensity sums the spectral density times |u|^2 over the
wavelength axis; power is ``prod(dx) * sum_{y,x} intensity`` with a
reproducible FP64 plane reduction.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib as L


def _density(field) -> tuple[int, object]:
    from ..core.field import ChromaticScalarField

    if isinstance(field, ChromaticScalarField):
        d = field.spectrum.density
        return int(d.size), L.doubles(d)
    return 1, None


def _intensity_fwd(u, nwv, dens, out_shape):
    out = torch.empty(out_shape, dtype=torch.float32, device=u.device)
    L.check(L.lib().cxm_intensity_c64(L.stream_ptr(u.device), u.data_ptr(), out.data_ptr(),
                                      out.numel(), nwv, dens))
    return out


class _Intensity(torch.autograd.Function):
    """|u|^2 with its reverse-mode rule dL/du = 2 density dL/dI u
    (cxm_intensity_vjp_c64)."""

    @staticmethod
    def forward(ctx, u, nwv, density, out_shape):
        ctx.nwv, ctx.density = nwv, density
        ctx.save_for_backward(u)
        return _intensity_fwd(u, nwv, L.doubles(density) if density is not None else None,
                              out_shape)

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        g = g.to(torch.float32).resolve_conj().contiguous()
        gu = torch.empty_like(u)
        dens = L.doubles(ctx.density) if ctx.density is not None else None
        L.check(L.lib().cxm_intensity_vjp_c64(L.stream_ptr(u.device), u.data_ptr(), g.data_ptr(),
                                              gu.data_ptr(), g.numel(), ctx.nwv, dens))
        return gu, None, None, None


def intensity(field) -> torch.Tensor:
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    chrom = _is_chromatic(field)
    nwv = field.spectrum.size if chrom else 1
    out_shape = tuple(u.shape[:-1]) if chrom else tuple(u.shape)
    if u.requires_grad and torch.is_grad_enabled():
        return _Intensity.apply(u, nwv, field.spectrum.density if chrom else None, out_shape)
    _, dens = _density(field)
    return _intensity_fwd(u, nwv, dens, out_shape)


def _is_chromatic(field) -> bool:
    from ..core.field import ChromaticScalarField

    return isinstance(field, ChromaticScalarField)


def plane_sumsq(field) -> torch.Tensor:
    """FP64 sum over (y, x[, wv]) of density-weighted |u|^2, one per batch
    element (shape = batch shape)."""
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    nwv, dens = _density(field)
    chrom = _is_chromatic(field)
    spatial = 3 if chrom else 2
    batch_shape = tuple(u.shape[: u.ndim - spatial])
    n_planes = int(np.prod(batch_shape)) if batch_shape else 1
    plane_elems = u.numel() // max(1, n_planes)
    out = torch.empty(max(1, n_planes), dtype=torch.float64, device=u.device)
    ws_bytes = L.lib().cxm_plane_sumsq_workspace_size(n_planes)
    ws = torch.empty(max(1, ws_bytes), dtype=torch.uint8, device=u.device)
    L.check(L.lib().cxm_plane_sumsq_c64(L.stream_ptr(u.device), u.data_ptr(), out.data_ptr(),
                                        n_planes, plane_elems, nwv if chrom else 1, dens,
                                        ws.data_ptr(), ws_bytes))
    return out.reshape(batch_shape)


def power(field) -> torch.Tensor:
    """prod(dx) * sum of intensity; shape (... 1 1) for scalar fields and
    (... 1 1 wv) for chromatic ones, as in the reference."""
    s = plane_sumsq(field)
    if _is_chromatic(field):
        area = torch.as_tensor(np.prod(field.dx, axis=-1), dtype=torch.float64, device=s.device)
        p = s[..., None, None, None] * area
    else:
        area = float(np.prod(field.dx))
        p = s[..., None, None] * area
    return p.to(torch.float32)


def plane_sumsq_raw(u: torch.Tensor) -> float:
    """sum |u|^2 over the whole complex64 tensor, reduced in FP64 on the device."""
    u = u.resolve_conj().contiguous()
    L.ensure_device(u)
    out = torch.empty(1, dtype=torch.float64, device=u.device)
    ws_bytes = L.lib().cxm_plane_sumsq_workspace_size(1)
    ws = torch.empty(max(1, ws_bytes), dtype=torch.uint8, device=u.device)
    L.check(L.lib().cxm_plane_sumsq_c64(L.stream_ptr(u.device), u.data_ptr(), out.data_ptr(), 1,
                                        u.numel(), 1, None, ws.data_ptr(), ws_bytes))
    return float(out.item())
