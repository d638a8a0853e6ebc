"""Lens / far-field FFT family and scalable angular spectrum on libcxm.

Mirrors, with the reference signatures:
  optical_fft                  src/chromatix/functional/convenience.py:9-59
  ff_lens                      src/chromatix/functional/lenses.py:64-95
  transform_propagate          src/chromatix/functional/propagation.py:34-82
  compute_sas_precompensation  src/chromatix/functional/propagation.py:85-103
  transform_propagate_sas      src/chromatix/functional/propagation.py:106-165
  fft / ifft (with shift)      src/chromatix/utils/fft.py:8-35
  high_na_ff_lens (scalar)     src/chromatix/functional/lenses.py:98-156

Each call is one fused two-pass (optical FFT, transform) or four/five-pass
(SAS) kernel chain; the new sample spacing is computed on the host in FP64.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field
from .propagation import _Plan, _pair, get_phase_mode

__all__ = [
    "optical_fft",
    "ff_lens",
    "transform_propagate",
    "compute_sas_precompensation",
    "transform_propagate_sas",
    "fft",
    "ifft",
    "high_na_ff_lens",
]


def _desc(field: Field, mode: int, z: float, n: float, phase_mode, out_kind=L.CXM_OUT_COMPLEX):
    d = L.OptFftDesc()
    ny, nx = field.spatial_shape
    d.height, d.width = ny, nx
    d.mode = mode
    d.z, d.n = float(z), float(n)
    d.phase_mode = L.PHASE_MODES[phase_mode or get_phase_mode()]
    d.out_kind = out_kind
    d.central_wavelength = float(field.spectrum.wavelength[0])
    keep = []
    if isinstance(field, ChromaticScalarField):
        d.num_wv = field.spectrum.size
        wl = L.doubles(field.spectrum.wavelength)
        sp = L.doubles(np.asarray(field.dx, np.float64).reshape(-1))
        org = L.doubles(np.asarray(field.origin, np.float64).reshape(-1))
        keep += [wl, sp, org]
        d.wavelengths = ctypes.cast(wl, ctypes.POINTER(ctypes.c_double))
        d.spacing = ctypes.cast(sp, ctypes.POINTER(ctypes.c_double))
        d.origin = ctypes.cast(org, ctypes.POINTER(ctypes.c_double))
        d.wavelength = float(field.spectrum.wavelength[0])
        d.dy, d.dx = float(field.dx[0, 0]), float(field.dx[0, 1])
    else:
        d.num_wv = 1
        d.wavelength = float(field.spectrum.wavelength[0])
        d.dy, d.dx = float(field.dx[0]), float(field.dx[1])
        d.oy, d.ox = float(field.origin[0]), float(field.origin[1])
    return d, keep


def _run(field: Field, d, entry: str, out_hw: tuple[int, int]) -> torch.Tensor:
    u = field.u
    L.ensure_device(u)
    plan = _Plan(field, np.zeros(1))
    u_pl = plan.planar_input(u)
    d.batch = u_pl.shape[0]
    dtype = torch.float32 if d.out_kind == L.CXM_OUT_INTENSITY else torch.complex64
    out = torch.empty((d.batch,) + tuple(out_hw), dtype=dtype, device=u.device)
    size_fn = getattr(L.lib(), f"cxm_{entry}_workspace_size")
    ws_bytes = size_fn(ctypes.byref(d))
    if ws_bytes == 0 and d.batch > 0:
        L.check(getattr(L.lib(), f"cxm_{entry}_c64")(L.stream_ptr(u.device), None, None, None, 0,
                                                      ctypes.byref(d)))
    ws = torch.empty(max(1, ws_bytes), dtype=torch.uint8, device=u.device)
    L.check(getattr(L.lib(), f"cxm_{entry}_c64")(L.stream_ptr(u.device), u_pl.data_ptr(),
                                                  out.data_ptr(), ws.data_ptr(), ws_bytes,
                                                  ctypes.byref(d)))
    return plan.unplanar(out)


class _OpticalFn(torch.autograd.Function):
    """An optical-FFT-family call with its adjoint: the same descriptor with
    `adjoint` set (the conjugated chirps, scale and pupil around the DFT of
    the other direction, pads and crops exchanged) applied to the output's
    cotangent.  The reference gets these rules from JAX's FFT transpose."""

    @staticmethod
    def forward(ctx, u, field, d, keep, entry, out_hw):
        ctx.field, ctx.d, ctx.keep, ctx.entry = field, d, keep, entry
        ctx.in_hw = tuple(field.spatial_shape)
        return _run(field.replace(u=u.detach()), d, entry, out_hw)

    @staticmethod
    def backward(ctx, g):
        da = L.OptFftDesc.from_buffer_copy(ctx.d)
        da.adjoint = 1
        fg = ctx.field.replace(u=g.to(torch.complex64).resolve_conj().contiguous())
        return _run(fg, da, ctx.entry, ctx.in_hw), None, None, None, None, None


def _apply(field: Field, d, keep, entry: str, out_hw) -> torch.Tensor:
    if torch.is_grad_enabled() and field.u.requires_grad:
        if d.out_kind != L.CXM_OUT_COMPLEX:
            raise NotImplementedError("gradients flow through the complex output; take "
                                      "`.intensity` of the returned field")
        return _OpticalFn.apply(field.u, field, d, keep, entry, tuple(out_hw))
    return _run(field, d, entry, out_hw)


def _new_spacing(field: Field, z: float, n: float, grid: tuple[int, int]) -> np.ndarray:
    """du = df * |lam z / n| with df = 1 / (N dx) on the (padded) grid
    (convenience.py:33-41)."""
    lam = np.asarray(field.spectrum.wavelength, np.float64)
    L2 = np.abs(lam * z / n)
    N = np.asarray(grid, np.float64)
    if isinstance(field, ChromaticScalarField):
        return L2[:, None] / (N[None, :] * field.dx)
    return L2[0] / (N * field.dx)


def optical_fft(field: Field, z, n, *, phase_mode=None) -> Field:
    """Far-field / lens FFT with the sampling change du = df |lam z / n|
    (differentiable in the field)."""
    z = float(z)
    d, keep = _desc(field, L.CXM_OFFT_OPTICAL, z, n, phase_mode)
    u = _apply(field, d, keep, "optical_fft", field.spatial_shape)
    return field.replace(u=u, dx=_new_spacing(field, z, n, field.spatial_shape))


def ff_lens(field: Field, f, n, NA=None, inverse: bool = False, *, phase_mode=None) -> Field:
    """Thin lens a distance f after the field, observed in its back focal
    plane: optional circular pupil of diameter 2 f NA / n, then optical_fft
    with f (or -f when inverse)."""
    f = float(f)
    z = -f if inverse else f
    d, keep = _desc(field, L.CXM_OFFT_OPTICAL, z, n, phase_mode)
    if NA is not None:
        d.pupil = 1
        d.pupil_d = 2.0 * f * float(NA) / float(n)
    u = _apply(field, d, keep, "optical_fft", field.spatial_shape)
    return field.replace(u=u, dx=_new_spacing(field, z, n, field.spatial_shape))


def transform_propagate(field: Field, z, n, pad_width, cval: float = 0,
                        skip_initial_phase: bool = False, skip_final_phase: bool = False, *,
                        phase_mode=None) -> Field:
    """Single-FFT Fresnel propagation (SFT-FR) with the sampling change
    (differentiable in the field)."""
    py, px = _pair(pad_width)
    if cval != 0:
        from ..core.field import pad as _pad

        field = _pad(field, (py, px), cval=cval)
        grid_pad = (0, 0)
    else:
        grid_pad = (py, px)
    z = float(z)
    d, keep = _desc(field, L.CXM_OFFT_TRANSFORM, z, n, phase_mode)
    d.pad_y, d.pad_x = grid_pad
    d.crop_y, d.crop_x = py, px
    d.skip_initial_phase = int(bool(skip_initial_phase))
    d.skip_final_phase = int(bool(skip_final_phase))
    ny, nx = field.spatial_shape
    grid = (ny + 2 * grid_pad[0], nx + 2 * grid_pad[1])
    out_hw = (grid[0] - 2 * py, grid[1] - 2 * px)
    u = _apply(field, d, keep, "optical_fft", out_hw)
    return field.replace(u=u, dx=_new_spacing(field, z, n, grid))


def transform_propagate_sas(field: Field, z, n, cval: float = 0,
                            skip_initial_phase: bool = False, skip_final_phase: bool = False, *,
                            phase_mode=None) -> Field:
    """Scalable angular spectrum propagation (padding factor 2, as the
    reference supports)."""
    L.refuse_grad("transform_propagate_sas", field.u)
    if cval != 0:
        raise NotImplementedError("transform_propagate_sas with cval != 0")
    z = float(z)
    d, keep = _desc(field, L.CXM_OFFT_TRANSFORM, z, n, phase_mode)
    d.skip_initial_phase = int(bool(skip_initial_phase))
    d.skip_final_phase = int(bool(skip_final_phase))
    ny, nx = field.spatial_shape
    u = _run(field, d, "sas", (ny, nx))
    grid = (ny + 2 * (ny // 2), nx + 2 * (nx // 2))
    return field.replace(u=u, dx=_new_spacing(field, z, n, grid))


def compute_sas_precompensation(field: Field, z, n, *, phase_mode=None) -> torch.Tensor:
    """dH of the SAS method for `field`'s grid, in FFT order; shape
    (H, W) or (H, W, wv) like the reference."""
    d, keep = _desc(field, L.CXM_OFFT_TRANSFORM, float(z), n, phase_mode)
    ny, nx = field.spatial_shape
    nwv = d.num_wv
    L.ensure_device(field.u)
    out = torch.empty((nwv, ny, nx), dtype=torch.complex64, device=field.u.device)
    L.check(L.lib().cxm_sas_precompensation_c64(L.stream_ptr(field.u.device), out.data_ptr(),
                                                ctypes.byref(d)))
    if isinstance(field, ChromaticScalarField):
        return out.permute(1, 2, 0)
    return out[0]


def _raw(x: torch.Tensor, axes, shift: bool, inverse: bool) -> torch.Tensor:
    x = torch.as_tensor(x)
    ax = tuple(a % x.ndim for a in axes)
    if len(ax) != 2:
        raise ValueError("fft/ifft take two axes")
    xm = x.movedim(ax, (-2, -1)).to(torch.complex64)
    from ..core.field import ScalarField

    f = ScalarField(xm, 1.0, 0.0, 1.0)
    d, keep = _desc(f, L.CXM_OFFT_RAW_INV if inverse else L.CXM_OFFT_RAW_FWD, 0.0, 1.0, None)
    d.shift = int(bool(shift))
    out = _apply(f, d, keep, "optical_fft", f.spatial_shape)
    return out.movedim((-2, -1), ax)


def fft(x, axes=(1, 2), shift: bool = False) -> torch.Tensor:
    """fft2 over `axes`; with shift, fftshift(fft2(ifftshift(x)))."""
    return _raw(x, axes, shift, inverse=False)


def ifft(x, axes=(1, 2), shift: bool = False) -> torch.Tensor:
    """ifft2 over `axes`; with shift, fftshift(ifft2(ifftshift(x)))."""
    return _raw(x, axes, shift, inverse=True)


def high_na_ff_lens(field: Field, f, n, NA, output_shape=None, output_dx=None) -> Field:
    """High-NA lens a distance ``f`` after a scalar, single-wavelength, square
    field (lenses.py:98-156): the field times the defocus factor
    exp(i k sz f) / sz (0 where sz = 0), sz = sqrt(max(1 - NA^2 |f_grid
    lambda / n|^2, 0)), k = 2 pi n / lambda (cxm_high_na_defocus_c64, FP64),
    then the zoomed FFT over [-zoom pi, zoom pi] with the end point,
    zoom = 2 NA (output_shape[0] output_dx) / ((N - 1) lambda) (two CZT
    passes).  The output spacing is ``output_dx`` (default: the field's y
    spacing) on both axes.  Vectorial fields (the reference's
    cartesian_to_spherical branch) are not part of this library."""
    from ..utils.czt import zoomed_fft

    L.refuse_grad("high_na_ff_lens", field.u)
    if isinstance(field, ChromaticScalarField):
        raise NotImplementedError("high_na_ff_lens: single-wavelength scalar fields only "
                                  "(the reference assumes one wavelength)")
    ny, nx = field.spatial_shape
    if output_dx is None:
        output_dx = float(field.dx[0])
    output_dx = float(output_dx)
    if output_shape is None:
        output_shape = (ny, nx)
    output_shape = (int(output_shape[0]), int(output_shape[1]))
    wl = float(field.spectrum.wavelength[0])
    zoom = 2.0 * float(NA) * output_shape[0] * output_dx / ((nx - 1) * wl)
    u = field.u
    L.ensure_device(u)
    u = u.to(torch.complex64).resolve_conj().contiguous()
    n_planes = u.numel() // (ny * nx)
    tmp = torch.empty_like(u)
    L.check(L.lib().cxm_high_na_defocus_c64(L.stream_ptr(u.device), u.data_ptr(), tmp.data_ptr(),
                                            n_planes, ny, nx, float(field.dx[0]), float(field.dx[1]),
                                            wl, float(n), float(NA), float(f)))
    out = zoomed_fft(tmp, -zoom * np.pi, zoom * np.pi, output_shape, include_end=True)
    return field.replace(u=out, dx=np.full(2, output_dx))
