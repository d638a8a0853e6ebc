"""Free-space propagation with the signatures of
src/chromatix/functional/propagation.py, computed by libcxm HIP kernels.

The default (non-CZT) branch of ``kernel_propagate`` together with the
propagators of ``transfer_propagate`` / ``asm_propagate`` runs as one fused
three-pass kernel sequence per call (zero-pad, FFT, H on the fly, IFFT,
crop); no propagator array is materialised.

Phase evaluation: ``"ref32"`` (default) reproduces the reference's float32
operation order for the propagator phase, so results agree with the
reference to FFT rounding; ``"accurate"`` evaluates the phase in double-float
and agrees with an FP64 computation to ~1e-7.  Select per call with the
keyword-only ``phase_mode`` argument or globally with ``set_phase_mode``.
"""
from __future__ import annotations

import ctypes
from typing import Literal

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field, pad, shift_grid

__all__ = [
    "transform_propagate",
    "compute_sas_precompensation",
    "transform_propagate_sas",
    "transfer_propagate",
    "asm_propagate",
    "kernel_propagate",
    "compute_transfer_propagator",
    "compute_asm_propagator",
    "compute_padding_transform",
    "compute_padding_transfer",
    "compute_padding_exact",
    "set_phase_mode",
    "get_phase_mode",
    "propagate_intensity",
]

_PHASE_MODE = "ref32"
_BOUNDARIES = ("tukey", "super_gaussian")


def set_phase_mode(mode: str) -> None:
    """Global default phase evaluation: "ref32" or "accurate"."""
    global _PHASE_MODE
    if mode not in L.PHASE_MODES:
        raise ValueError(f"phase mode must be one of {tuple(L.PHASE_MODES)}")
    _PHASE_MODE = mode


def get_phase_mode() -> str:
    return _PHASE_MODE


def _pair(v) -> tuple[int, int]:
    if isinstance(v, (int, np.integer)):
        return int(v), int(v)
    if isinstance(v, (tuple, list, np.ndarray)) and len(v) == 2:
        return int(v[0]), int(v[1])
    raise TypeError("pad_width must be an int or a pair of ints (it must not be an array)")


def _vec2(v) -> tuple[float, float]:
    a = np.atleast_1d(np.asarray(v, dtype=np.float64)).reshape(-1)
    if a.size == 1:
        return float(a[0]), float(a[0])
    if a.size != 2:
        raise ValueError("expected a pair (y, x)")
    return float(a[0]), float(a[1])


class _Plan:
    """Batch/z broadcasting of one propagation call (reference semantics:
    a (Z,1,1[,1]) propagator broadcast against the field's batch)."""

    def __init__(self, field: Field, z: np.ndarray):
        self.chrom = isinstance(field, ChromaticScalarField)
        self.B = field.batch_shape
        self.nz = int(z.size)
        if self.nz == 1:
            self.mode, self.Bp = "scalar", self.B
        elif len(self.B) == 0:
            self.mode, self.Bp = "outer", ()
        elif self.B[-1] == 1:
            self.mode, self.Bp = "outer", self.B[:-1]
        elif self.B[-1] == self.nz:
            self.mode, self.Bp = "zip", self.B
        else:
            raise ValueError(
                f"cannot broadcast {self.nz} distances against batch shape {self.B}"
            )
        self.nwv = field.spectrum.size if self.chrom else 1

    def planar_input(self, u: torch.Tensor) -> torch.Tensor:
        if self.mode == "outer" and len(self.B) > 0:
            u = u.squeeze(len(self.B) - 1)
        if self.chrom:
            u = u.movedim(-1, -3)
        h, w = u.shape[-2], u.shape[-1]
        return u.reshape(-1, h, w).resolve_conj().contiguous()

    def out_planes(self, n_in: int) -> int:
        return n_in * self.nz if self.mode == "outer" else n_in

    def unplanar(self, out: torch.Tensor) -> torch.Tensor:
        ho, wo = out.shape[-2], out.shape[-1]
        if self.mode == "outer":
            lead = self.Bp + ((self.nwv,) if self.chrom else ())
            out = out.reshape(*lead, self.nz, ho, wo)
            if self.chrom:  # (..., wv, Z, H, W) -> (..., Z, H, W, wv)
                out = out.movedim(-4, -1)
            return out
        lead = self.Bp + ((self.nwv,) if self.chrom else ())
        out = out.reshape(*lead, ho, wo)
        if self.chrom:
            out = out.movedim(-3, -1)
        return out


def _desc_common(field: Field, kind: int, z: np.ndarray, n: float, pad_yx, crop_yx,
                 phase_mode: str | None, out_kind: int, kykx=(0.0, 0.0), shift_yx=(0.0, 0.0),
                 remove_evanescent=False, bandlimit=False, plan: _Plan | None = None):
    d = L.Prop2dDesc()
    ny, nx = field.spatial_shape
    d.height, d.width = ny, nx
    d.pad_y, d.pad_x = pad_yx
    d.crop_y, d.crop_x = crop_yx
    d.kind = kind
    d.phase_mode = L.PHASE_MODES[phase_mode or _PHASE_MODE]
    d.out_kind = out_kind
    d.remove_evanescent = int(bool(remove_evanescent))
    d.bandlimit = int(bool(bandlimit))
    d.n = float(n)
    d.ky, d.kx = _vec2(kykx)
    d.shift_y, d.shift_x = _vec2(shift_yx)
    keep = []
    zarr = L.doubles(z)
    keep.append(zarr)
    d.z = ctypes.cast(zarr, ctypes.POINTER(ctypes.c_double))
    d.num_z = int(z.size)
    d.z_zip = 1 if (plan is not None and plan.mode == "zip") else 0
    if isinstance(field, ChromaticScalarField):
        d.num_wv = field.spectrum.size
        wl = L.doubles(field.spectrum.wavelength)
        sp = L.doubles(np.asarray(field.dx, dtype=np.float64).reshape(-1))
        keep += [wl, sp]
        d.wavelengths = ctypes.cast(wl, ctypes.POINTER(ctypes.c_double))
        d.spacing = ctypes.cast(sp, ctypes.POINTER(ctypes.c_double))
        d.wavelength = float(field.spectrum.wavelength[0])
        d.dy, d.dx = float(field.dx[0, 0]), float(field.dx[0, 1])
    else:
        d.num_wv = 1
        d.wavelength = float(field.spectrum.wavelength[0])
        d.dy, d.dx = float(field.dx[0]), float(field.dx[1])
    return d, keep


def _launch(field: Field, d, plan: _Plan, out_kind: int, propagator: torch.Tensor | None = None):
    u = field.u
    L.ensure_device(u)
    u_pl = plan.planar_input(u)
    d.batch = u_pl.shape[0]
    if propagator is not None:
        d.propagator = propagator.data_ptr()
    ho = d.height + 2 * d.pad_y - 2 * d.crop_y
    wo = d.width + 2 * d.pad_x - 2 * d.crop_x
    n_out = plan.out_planes(d.batch)
    dtype = torch.complex64 if out_kind == L.CXM_OUT_COMPLEX else torch.float32
    out = torch.empty((n_out, ho, wo), dtype=dtype, device=u.device)
    ws_bytes = L.lib().cxm_prop2d_workspace_size(ctypes.byref(d))
    if ws_bytes == 0 and d.batch > 0:
        # the size query validates; surface the reason
        L.check(L.lib().cxm_prop2d_c64(L.stream_ptr(u.device), None, None, None, 0, ctypes.byref(d)))
    ws = torch.empty(max(1, ws_bytes), dtype=torch.uint8, device=u.device)
    L.check(L.lib().cxm_prop2d_c64(L.stream_ptr(u.device), u_pl.data_ptr(), out.data_ptr(),
                                   ws.data_ptr(), ws_bytes, ctypes.byref(d)))
    return plan.unplanar(out)


def _fused_propagate(field: Field, kind: int, z, n, pad_width, cval, mode, kykx=(0.0, 0.0),
                     shift_yx=(0.0, 0.0), remove_evanescent=False, bandlimit=False,
                     phase_mode=None, out_kind=L.CXM_OUT_COMPLEX):
    if mode not in ("full", "same"):
        raise ValueError('mode must be "full" or "same"')
    pad_yx = _pair(pad_width)
    if cval != 0:
        field = pad(field, pad_yx, cval=cval)
        grid_pad = (0, 0)
    else:
        grid_pad = pad_yx
    crop_yx = pad_yx if mode == "same" else (0, 0)
    z = np.atleast_1d(np.asarray(z, dtype=np.float64)).reshape(-1)
    plan = _Plan(field, z)
    args = (field, kind, z, n, grid_pad, crop_yx, phase_mode, kykx, shift_yx, remove_evanescent,
            bandlimit)
    if field.u.requires_grad and torch.is_grad_enabled():
        if out_kind != L.CXM_OUT_COMPLEX:
            raise NotImplementedError("gradients flow through the complex output; take "
                                      "`.intensity` of the propagated field")
        if plan.mode == "outer":
            raise NotImplementedError("gradients of one field propagated to several distances "
                                      "are not implemented (batch the field over z instead)")
        return _PropagateFn.apply(field.u, args)
    return _propagate_planes(field.u, args, out_kind)


def _propagate_planes(u, args, out_kind, adjoint: bool = False):
    field, kind, z, n, grid_pad, crop_yx, phase_mode, kykx, shift_yx, rem, bl = args
    f = field.replace(u=u)
    if adjoint:
        # crop_c(P_z(pad_p(u)))^H = crop_p(P_{-z}(pad_c(.))): conj(H(z)) = H(-z) for the
        # transfer and angular-spectrum propagators as the reference defines them
        # (propagation.py:493-496, 556), the band limit depending on z^2 only
        z, grid_pad, crop_yx = -z, crop_yx, grid_pad
    plan = _Plan(f, z)
    d, keep = _desc_common(f, kind, z, n, grid_pad, crop_yx, phase_mode, out_kind, kykx,
                           shift_yx, rem, bl, plan)
    out = _launch(f, d, plan, out_kind)
    del keep
    return out


class _PropagateFn(torch.autograd.Function):
    """Fused propagation with its adjoint: the same kernels with -z and the
    pad / crop roles exchanged (the reference gets this from JAX's FFT
    transpose rules)."""

    @staticmethod
    def forward(ctx, u, args):
        ctx.args = args
        return _propagate_planes(u.detach(), args, L.CXM_OUT_COMPLEX)

    @staticmethod
    def backward(ctx, g):
        return _propagate_planes(g.to(torch.complex64).resolve_conj().contiguous(), ctx.args, L.CXM_OUT_COMPLEX,
                                 adjoint=True), None


def _check_boundary(absorbing_boundary):
    if absorbing_boundary is not None and absorbing_boundary not in _BOUNDARIES:
        raise AssertionError(f"The absorbing_boundary must be None or in {_BOUNDARIES}.")


def transfer_propagate(
    field: Field,
    z,
    n,
    pad_width,
    cval: float = 0,
    absorbing_boundary: Literal["tukey", "super_gaussian"] | None = None,
    absorbing_boundary_width: float = 0.65,
    kykx=(0.0, 0.0),
    shift_yx=(0.0, 0.0),
    output_dx=None,
    output_shape=None,
    use_czt: bool = True,
    mode: Literal["full", "same"] = "full",
    *,
    phase_mode: str | None = None,
) -> Field:
    """Fresnel (transfer-function) propagation; propagation.py:168-252."""
    _check_boundary(absorbing_boundary)
    if output_dx is not None or output_shape is not None:
        from .czt_propagation import czt_kernel_propagate_fields

        return czt_kernel_propagate_fields("transfer", field, z, n, pad_width, cval, kykx,
                                           shift_yx, output_dx, output_shape, use_czt, mode,
                                           phase_mode=phase_mode)
    u = _fused_propagate(field, L.CXM_H_TRANSFER, z, n, pad_width, cval, mode, kykx=kykx,
                         phase_mode=phase_mode)
    # the reference pads, propagates on the padded grid, then crops in "same" mode;
    # shift_grid sets the origin (kernel_propagate, propagation.py:391-394)
    return shift_grid(field.replace(u=u), shift_yx)


def asm_propagate(
    field: Field,
    z,
    n,
    pad_width,
    cval: float = 0,
    absorbing_boundary: Literal["tukey", "super_gaussian"] | None = None,
    absorbing_boundary_width: float = 0.65,
    kykx=(0.0, 0.0),
    remove_evanescent: bool = False,
    bandlimit: bool = False,
    shift_yx=(0.0, 0.0),
    output_dx=None,
    output_shape=None,
    use_czt: bool = True,
    mode: Literal["full", "same"] = "full",
    *,
    phase_mode: str | None = None,
) -> Field:
    """Angular-spectrum propagation; propagation.py:255-352."""
    _check_boundary(absorbing_boundary)
    if output_dx is not None or output_shape is not None:
        from .czt_propagation import czt_kernel_propagate_fields

        return czt_kernel_propagate_fields("asm", field, z, n, pad_width, cval, kykx, shift_yx,
                                           output_dx, output_shape, use_czt, mode,
                                           remove_evanescent=remove_evanescent,
                                           bandlimit=bandlimit, phase_mode=phase_mode)
    u = _fused_propagate(field, L.CXM_H_ASM, z, n, pad_width, cval, mode, kykx=kykx,
                         shift_yx=shift_yx, remove_evanescent=remove_evanescent,
                         bandlimit=bandlimit, phase_mode=phase_mode)
    return shift_grid(field.replace(u=u), shift_yx)


def propagate_intensity(
    field: Field,
    z,
    n,
    pad_width,
    method: Literal["asm", "transfer"] = "asm",
    mode: Literal["full", "same"] = "same",
    kykx=(0.0, 0.0),
    remove_evanescent: bool = False,
    bandlimit: bool = False,
    shift_yx=(0.0, 0.0),
    *,
    phase_mode: str | None = None,
) -> torch.Tensor:
    """``asm_propagate(...).intensity`` (or transfer) with |u|^2 fused into the
    last pass; returns the intensity tensor (mono fields)."""
    if isinstance(field, ChromaticScalarField):
        return (asm_propagate if method == "asm" else transfer_propagate)(
            field, z, n, pad_width, kykx=kykx, mode=mode, phase_mode=phase_mode).intensity
    kind = L.CXM_H_ASM if method == "asm" else L.CXM_H_TRANSFER
    return _fused_propagate(field, kind, z, n, pad_width, 0, mode, kykx=kykx, shift_yx=shift_yx,
                            remove_evanescent=remove_evanescent, bandlimit=bandlimit,
                            phase_mode=phase_mode, out_kind=L.CXM_OUT_INTENSITY)


def kernel_propagate(
    field: Field,
    propagator,
    absorbing_boundary: Literal["tukey", "super_gaussian"] | None = None,
    absorbing_boundary_width: float = 0.65,
    output_dx=None,
    output_shape=None,
    shift_yx=(0.0, 0.0),
    use_czt: bool = False,
) -> Field:
    """Fourier-convolution propagation with a given propagator in FFT order
    (propagation.py:355-463).

    As in the reference, an absorbing boundary is validated but does not
    change the returned field: the reference applies the pupil to a copy of
    the input field and then replaces its ``u`` with the propagated field
    (propagation.py:460-463).
    """
    _check_boundary(absorbing_boundary)
    if output_dx is not None or output_shape is not None or use_czt:
        from .czt_propagation import czt_kernel_propagate

        return czt_kernel_propagate(field, propagator, output_dx, output_shape, shift_yx, use_czt)
    H = propagator if isinstance(propagator, torch.Tensor) else torch.as_tensor(np.asarray(propagator))
    H = H.to(device=field.u.device, dtype=torch.complex64)
    ny, nx = field.spatial_shape
    if isinstance(field, ChromaticScalarField) and H.ndim >= 3 and H.shape[-1] == field.spectrum.size:
        H = H.movedim(-1, -3)  # (..., wv, Ny, Nx)
    if H.shape[-2:] != (ny, nx):
        raise ValueError(f"propagator spatial shape {tuple(H.shape[-2:])} != field {(ny, nx)}")
    lead = H.shape[:-2]
    if all(s == 1 for s in lead):
        out = _kernel_propagate_array(field, H.reshape(ny, nx).resolve_conj().contiguous())
    else:
        out = _kernel_propagate_broadcast(field, H)
    return shift_grid(field.replace(u=out), shift_yx)


def _kernel_propagate_array(field: Field, H: torch.Tensor) -> torch.Tensor:
    z = np.zeros(1)
    plan = _Plan(field, z)
    d, keep = _desc_common(field, L.CXM_H_ARRAY, z, 1.0, (0, 0), (0, 0), None, L.CXM_OUT_COMPLEX,
                           plan=plan)
    out = _launch(field, d, plan, L.CXM_OUT_COMPLEX, propagator=H)
    del keep
    return out


def _kernel_propagate_broadcast(field: Field, H: torch.Tensor) -> torch.Tensor:
    """Propagator with its own batch dimensions: one launch per distinct
    propagator plane, broadcast as numpy would."""
    chrom = isinstance(field, ChromaticScalarField)
    u = field.u.movedim(-1, -3) if chrom else field.u  # planar (..., [wv,] H, W)
    shape = torch.broadcast_shapes(u.shape, H.shape)
    ub = u.expand(shape)
    Hb = H.expand(shape[:-2] + H.shape[-2:])
    outs = torch.empty(shape, dtype=torch.complex64, device=u.device)
    lead = shape[:-2]
    for idx in np.ndindex(*lead):
        mono = field.replace(u=ub[idx]) if not chrom else None
        if chrom:
            from ..core.field import ScalarField

            w = idx[-1]
            mono = ScalarField(ub[idx], field.dx[w], 0.0, float(field.spectrum.wavelength[w]))
        outs[idx] = _kernel_propagate_array(mono, Hb[idx].resolve_conj().contiguous())
    return outs.movedim(-3, -1) if chrom else outs


def _materialise(field: Field, kind: int, z, n, kykx=(0.0, 0.0), shift_yx=(0.0, 0.0),
                 remove_evanescent=False, bandlimit=False, phase_mode=None) -> torch.Tensor:
    """The propagator array of the reference in FFT (ifftshifted) order."""
    z = np.atleast_1d(np.asarray(z, dtype=np.float64)).reshape(-1)
    d, keep = _desc_common(field, kind, z, n, (0, 0), (0, 0), phase_mode, L.CXM_OUT_COMPLEX,
                           kykx, shift_yx, remove_evanescent, bandlimit)
    ny, nx = field.spatial_shape
    nwv = d.num_wv
    out = torch.empty((nwv, z.size, ny, nx), dtype=torch.complex64, device=field.u.device)
    L.ensure_device(field.u)
    L.check(L.lib().cxm_propagator_c64(L.stream_ptr(field.u.device), out.data_ptr(),
                                       ctypes.byref(d)))
    del keep
    chrom = isinstance(field, ChromaticScalarField)
    if chrom:
        out = out.permute(1, 2, 3, 0)  # (Z, Ny, Nx, wv)
        return out[0] if z.size == 1 else out
    return out[0, 0] if z.size == 1 else out[0]


def compute_transfer_propagator(field: Field, z, n, kykx=(0.0, 0.0), *, phase_mode=None):
    """Fresnel propagator array (propagation.py:466-496)."""
    return _materialise(field, L.CXM_H_TRANSFER, z, n, kykx=kykx, phase_mode=phase_mode)


def compute_asm_propagator(field: Field, z, n, kykx=(0.0, 0.0), bandlimit: bool = False,
                           shift_yx=(0.0, 0.0), remove_evanescent: bool = False, *,
                           phase_mode=None):
    """Angular-spectrum propagator array (propagation.py:499-580)."""
    return _materialise(field, L.CXM_H_ASM, z, n, kykx=kykx, shift_yx=shift_yx,
                        remove_evanescent=remove_evanescent, bandlimit=bandlimit,
                        phase_mode=phase_mode)


# ---- padding calculators (propagation.py:583-656) ---------------------------
# Host arithmetic; the formulas are the reference's so that the same padding
# results (a calculator that returned different pads would change every
# downstream grid).
def _fresnel_q(height, wavelength, dx, z):
    D = height * dx
    Nf = np.max((D / 2) ** 2 / (wavelength * z))
    M = height
    return 2 * np.maximum(1.0, M / (4 * Nf)), M


def compute_padding_transform(height: int, wavelength: float, dx: float, z: float) -> int:
    Q, M = _fresnel_q(height, wavelength, dx, z)
    N = (np.ceil((Q * M) / 2) * 2).astype(int)
    return int(N - M)


def compute_padding_transfer(height: int, wavelength: float, dx: float, z: float) -> int:
    Q, M = _fresnel_q(height, wavelength, dx, z)
    N = (np.ceil((Q * M) / 2) * 2).astype(int)
    return int(N - M)


def compute_padding_exact(height: int, wavelength: float, dx: float, z: float) -> int:
    Q, M = _fresnel_q(height, wavelength, dx, z)
    scale = np.max(wavelength / (2 * dx))
    Q = Q / np.sqrt(1 - scale**2)
    N = (np.ceil((Q * M) / 2) * 2).astype(int)
    return int(N - M)


def transform_propagate(field, z, n, pad_width, cval=0, skip_initial_phase=False,
                        skip_final_phase=False, *, phase_mode=None):
    """Single-FFT Fresnel propagation (propagation.py:34-82); see optics.py."""
    from .optics import transform_propagate as _tp

    return _tp(field, z, n, pad_width, cval, skip_initial_phase, skip_final_phase,
               phase_mode=phase_mode)


def compute_sas_precompensation(field, z, n, *, phase_mode=None):
    """SAS precompensation dH (propagation.py:85-103); see optics.py."""
    from .optics import compute_sas_precompensation as _c

    return _c(field, z, n, phase_mode=phase_mode)


def transform_propagate_sas(field, z, n, cval=0, skip_initial_phase=False, skip_final_phase=False,
                            *, phase_mode=None):
    """Scalable angular spectrum propagation (propagation.py:106-165); see optics.py."""
    from .optics import transform_propagate_sas as _s

    return _s(field, z, n, cval, skip_initial_phase, skip_final_phase, phase_mode=phase_mode)
