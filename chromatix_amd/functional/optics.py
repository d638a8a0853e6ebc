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
    reference supports); differentiable in the field."""
    if cval != 0:
        raise NotImplementedError("transform_propagate_sas with cval != 0")
    z = float(z)
    d, keep = _desc(field, L.CXM_OFFT_TRANSFORM, z, n, phase_mode)
    d.skip_initial_phase = int(bool(skip_initial_phase))
    d.skip_final_phase = int(bool(skip_final_phase))
    ny, nx = field.spatial_shape
    grid = (ny + 2 * (ny // 2), nx + 2 * (nx // 2))
    if torch.is_grad_enabled() and field.u.requires_grad:
        u = _SasFn.apply(field.u, field, d, keep, float(n), phase_mode)
    else:
        u = _run(field, d, "sas", (ny, nx))
    return field.replace(u=u, dx=_new_spacing(field, z, n, grid))


def _spatial_axes(field: Field) -> tuple[int, int]:
    return (-3, -2) if isinstance(field, ChromaticScalarField) else (-2, -1)


def _sas_conj_dh_mul(field: Field, v: torch.Tensor, z: float, n: float, spacing,
                     phase_mode) -> torch.Tensor:
    """ifft2(conj(dH) fft2(v)) on v's (padded) grid with the SAS
    precompensation of `spacing` (conj(dH(z)) = dH(-z): the window depends on
    |z| only): the adjoint of the precompensation step."""
    ax = _spatial_axes(field)
    ny, nx = v.shape[ax[0]], v.shape[ax[1]]
    chrom = isinstance(field, ChromaticScalarField)
    d, keep = _desc(field, L.CXM_OFFT_TRANSFORM, -z, n, phase_mode)
    d.height, d.width = ny, nx
    sp = L.doubles(np.asarray(spacing, np.float64).reshape(-1))
    keep.append(sp)
    if chrom:
        d.spacing = ctypes.cast(sp, ctypes.POINTER(ctypes.c_double))
    else:
        d.dy, d.dx = float(sp[0]), float(sp[1])
    nwv = d.num_wv
    dh = torch.empty((nwv, ny, nx), dtype=torch.complex64, device=v.device)
    L.check(L.lib().cxm_sas_precompensation_c64(L.stream_ptr(v.device), dh.data_ptr(),
                                                ctypes.byref(d)))
    spec = fft(v, axes=ax, shift=False)
    pl = spec.movedim(-1, -3).resolve_conj().contiguous() if chrom else spec.resolve_conj().contiguous()
    L.check(L.lib().cxm_cmul_c64(L.stream_ptr(v.device), pl.data_ptr(), dh.data_ptr(),
                                 pl.data_ptr(), pl.numel() // (ny * nx), ny, nx, nwv, 0))
    spec = pl.movedim(-3, -1) if chrom else pl
    return ifft(spec, axes=ax, shift=False)


def _sas_adjoint(field: Field, d, g: torch.Tensor, n: float, phase_mode) -> torch.Tensor:
    """Adjoint of transform_propagate_sas (propagation.py:106-165) on the
    cotangent g of its output: z >= 0: crop(K^H(T^H(pad(g)))), z < 0:
    crop(T^H(K'^H(pad(g)))), with T the transform step on the padded grid
    (optical-FFT adjoint), K the precompensation on the input grid and K'
    on the transformed grid."""
    ny, nx = field.spatial_shape
    py, px = ny // 2, nx // 2
    Ny, Nx = ny + 2 * py, nx + 2 * px
    ax = _spatial_axes(field)
    z = float(d.z)
    chrom = isinstance(field, ChromaticScalarField)
    dt = L.OptFftDesc.from_buffer_copy(d)
    dt.height, dt.width = Ny, Nx
    dt.pad_y = dt.pad_x = 0
    dt.adjoint = 1
    g = g.to(torch.complex64).resolve_conj().contiguous()

    def crop(v):
        return v.narrow(ax[0] % v.ndim, py, ny).narrow(ax[1] % v.ndim, px, nx)

    if z >= 0:
        dt.crop_y, dt.crop_x = py, px
        v = _run(field.replace(u=g), dt, "optical_fft", (Ny, Nx))
        v = _sas_conj_dh_mul(field, v, z, n, field.dx, phase_mode)
        return crop(v).resolve_conj().contiguous()
    pads = [0, 0] * (1 if chrom else 0) + [px, px, py, py]
    v = torch.nn.functional.pad(g, pads)
    du = _new_spacing(field, z, n, (Ny, Nx))
    v = _sas_conj_dh_mul(field, v, z, n, du, phase_mode)
    dt.crop_y = dt.crop_x = 0
    v = _run(field.replace(u=v.resolve_conj().contiguous()), dt, "optical_fft", (Ny, Nx))
    return crop(v).resolve_conj().contiguous()


class _SasFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, field, d, keep, n, phase_mode):
        ctx.field, ctx.d, ctx.keep, ctx.n, ctx.pm = field, d, keep, n, phase_mode
        ny, nx = field.spatial_shape
        return _run(field.replace(u=u.detach()), d, "sas", (ny, nx))

    @staticmethod
    def backward(ctx, g):
        return _sas_adjoint(ctx.field, ctx.d, g, ctx.n, ctx.pm), None, None, None, None, None


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


def _raw(x: torch.Tensor, axes, shift: bool, inverse: bool, spec_mul=None) -> torch.Tensor:
    """fft2 / ifft2 over `axes`; spec_mul (a contiguous complex64 device
    array of the two axes' shape, FFT order; not differentiated) multiplies
    the output by fftshift(spec_mul) (shift) or spec_mul in the same pass."""
    x = torch.as_tensor(x)
    ax = tuple(a % x.ndim for a in axes)
    if len(ax) != 2:
        raise ValueError("fft/ifft take two axes")
    xm = x.movedim(ax, (-2, -1)).to(torch.complex64)
    from ..core.field import ScalarField

    f = ScalarField(xm, 1.0, 0.0, 1.0)
    d, keep = _desc(f, L.CXM_OFFT_RAW_INV if inverse else L.CXM_OFFT_RAW_FWD, 0.0, 1.0, None)
    d.shift = int(bool(shift))
    if spec_mul is not None:
        if torch.is_grad_enabled() and xm.requires_grad:
            raise RuntimeError("the fused spectrum multiply has no reverse-mode rule")
        d.spec_mul = spec_mul.data_ptr()
        keep = (keep, spec_mul)
    out = _apply(f, d, keep, "optical_fft", f.spatial_shape)
    return out.movedim((-2, -1), ax)


def fft(x, axes=(1, 2), shift: bool = False) -> torch.Tensor:
    """fft2 over `axes`; with shift, fftshift(fft2(ifftshift(x)))."""
    return _raw(x, axes, shift, inverse=False)


def ifft(x, axes=(1, 2), shift: bool = False) -> torch.Tensor:
    """ifft2 over `axes`; with shift, fftshift(ifft2(ifftshift(x)))."""
    return _raw(x, axes, shift, inverse=True)


def _defocus(u, pars):
    u = u.to(torch.complex64).resolve_conj().contiguous()
    ny, nx = u.shape[-2:]
    out = torch.empty_like(u)
    L.check(L.lib().cxm_high_na_defocus_c64(L.stream_ptr(u.device), u.data_ptr(), out.data_ptr(),
                                            u.numel() // (ny * nx), ny, nx, *pars))
    return out


class _DefocusFn(torch.autograd.Function):
    """The defocus factor of high_na_ff_lens; its reverse-mode rule
    multiplies by the conjugate factor (the factor regenerated from a plane
    of ones)."""

    @staticmethod
    def forward(ctx, u, pars):
        ctx.pars = pars
        return _defocus(u, pars)

    @staticmethod
    def backward(ctx, g):
        g = g.to(torch.complex64).resolve_conj().contiguous()
        ny, nx = g.shape[-2:]
        fac = torch.conj_physical(_defocus(torch.ones((1, ny, nx), dtype=torch.complex64,
                                                      device=g.device), ctx.pars))
        out = torch.empty_like(g)
        L.check(L.lib().cxm_cmul_c64(L.stream_ptr(g.device), g.data_ptr(), fac.data_ptr(),
                                     out.data_ptr(), g.numel() // (ny * nx), ny, nx, 1, 0))
        return out, None


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
    L.ensure_device(field.u)
    pars = (float(field.dx[0]), float(field.dx[1]), wl, float(n), float(NA), float(f))
    tmp = _DefocusFn.apply(field.u, pars)
    out = zoomed_fft(tmp, -zoom * np.pi, zoom * np.pi, output_shape, include_end=True)
    return field.replace(u=out, dx=np.full(2, output_dx))
