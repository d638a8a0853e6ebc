"""Thin and thick samples on libcxm (mirrors src/chromatix/functional/samples.py).

``thin_sample`` (samples.py:85-134) is one elementwise kernel.
``multislice_thick_sample`` (samples.py:137-282) runs the whole slice loop
as two fused passes per slice (cxm_multislice_fwd).  With gradients
requested for ``dn_stack`` / ``absorption_stack`` the forward also stores
the field state every ``checkpoint_every`` slices, and the backward
(cxm_multislice_bwd) recomputes one segment at a time from those states and
runs the hand-written adjoint: no per-slice residuals are kept.
``fluorescent_multislice_thick_sample`` (samples.py:285-441) runs each
Monte-Carlo sample as one library call (cxm_fluor_multislice_sample_c64):
two fused passes per slice forward, three per slice for the backward
intensity sweep.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field
from .propagation import _pair, _vec2, get_phase_mode

__all__ = ["thin_sample", "multislice_thick_sample", "multislice_mse_grad",
           "fluorescent_multislice_thick_sample"]


def _thin_fwd(u, dn, ab, meta):
    n_planes, H, W, nwv, wl, t = meta
    out = torch.empty_like(u)
    L.check(L.lib().cxm_thin_sample_c64(L.stream_ptr(u.device), u.data_ptr(), dn.data_ptr(),
                                        ab.data_ptr(), out.data_ptr(), n_planes, H, W, nwv,
                                        L.doubles(wl), t))
    return out


class _ThinSampleFn(torch.autograd.Function):
    """thin_sample with its vector-Jacobian product (cxm_modulation_vjp_c64):
    grad_u = g conj(T), grad_dn = sum k Im(g conj(out)), grad_abs = -sum k
    Re(g conj(out)), k = 2 pi t / lambda."""

    @staticmethod
    def forward(ctx, u, dn, ab, meta):
        out = _thin_fwd(u.detach(), dn.detach(), ab.detach(), meta)
        ctx.meta = meta
        ctx.save_for_backward(dn, ab, out)
        return out

    @staticmethod
    def backward(ctx, g):
        dn, ab, out = ctx.saved_tensors
        n_planes, H, W, nwv, wl, t = ctx.meta
        g = g.to(torch.complex64).resolve_conj().contiguous()
        gu = torch.empty_like(g) if ctx.needs_input_grad[0] else None
        gdn = torch.empty_like(dn) if ctx.needs_input_grad[1] else None
        gab = torch.empty_like(ab) if ctx.needs_input_grad[2] else None
        L.check(L.lib().cxm_modulation_vjp_c64(
            L.stream_ptr(g.device), 1, g.data_ptr(), out.data_ptr(), dn.data_ptr(), ab.data_ptr(),
            L.dptr(gu), L.dptr(gdn), L.dptr(gab), n_planes, H, W, nwv, L.doubles(wl), t))
        return gu, gdn, gab, None


def thin_sample(field: Field, absorption, dn, thickness) -> Field:
    """``field * exp(1j * 2 pi (dn + 1j * absorption) * thickness / lambda)``
    with (H, W) ``absorption`` and ``dn`` broadcast over the batch (and the
    wavelengths of a chromatic field).  Differentiable in the field, dn and
    the absorption."""
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    H, W = field.spatial_shape
    dn_t = torch.as_tensor(dn, device=u.device).to(torch.float32).resolve_conj().contiguous()
    ab_t = torch.as_tensor(absorption, device=u.device).to(torch.float32).resolve_conj().contiguous()
    if dn_t.ndim != 2 or tuple(dn_t.shape) != (H, W):
        raise ValueError(f"Refractive index must have shape {(H, W)}, got {tuple(dn_t.shape)}.")
    if ab_t.ndim != 2 or tuple(ab_t.shape) != (H, W):
        raise ValueError(f"Absorption must have shape {(H, W)}, got {tuple(ab_t.shape)}")
    if np.ndim(thickness) != 0:
        raise NotImplementedError("thin_sample: scalar thickness only")
    nwv = field.spectrum.size if isinstance(field, ChromaticScalarField) else 1
    meta = (u.numel() // (H * W * nwv), H, W, nwv,
            [float(w) for w in np.atleast_1d(field.spectrum.wavelength)], float(thickness))
    if torch.is_grad_enabled() and (u.requires_grad or dn_t.requires_grad or ab_t.requires_grad):
        return field.replace(u=_ThinSampleFn.apply(u, dn_t, ab_t, meta))
    return field.replace(u=_thin_fwd(u, dn_t, ab_t, meta))


def _desc(field: Field, depth, n, thickness, pad_width, NA, kykx, reverse_distance,
          remove_evanescent, bandlimit, return_stack, phase_mode, checkpoint_every):
    if isinstance(field, ChromaticScalarField):
        raise NotImplementedError("multislice_thick_sample: monochromatic scalar fields only")
    d = L.MultisliceDesc()
    H, W = field.spatial_shape
    d.batch = int(np.prod(field.shape[:-2])) if field.u.ndim > 2 else 1
    d.height, d.width = H, W
    d.depth = int(depth)
    d.pad_y, d.pad_x = _pair(pad_width)
    d.phase_mode = L.PHASE_MODES[phase_mode or get_phase_mode()]
    d.remove_evanescent = int(bool(remove_evanescent))
    d.bandlimit = int(bool(bandlimit))
    d.use_na = 0 if NA is None else 1
    d.NA = 0.0 if NA is None else float(NA)
    d.return_stack = int(bool(return_stack))
    d.out_kind = L.CXM_OUT_COMPLEX
    d.checkpoint_every = int(checkpoint_every)
    d.dy, d.dx = float(field.dx[0]), float(field.dx[1])
    d.wavelength = float(field.spectrum.wavelength[0])
    d.n = float(n)
    d.thickness = float(thickness)
    d.ky, d.kx = _vec2(kykx)
    if reverse_distance is None:
        # thickness_per_slice * depth / 2 (samples.py:266-267)
        reverse_distance = float(thickness) * depth / 2
    d.reverse_distance = float(reverse_distance)
    return d


def _stacks(field, absorption_stack, dn_stack):
    dev = field.u.device
    dn = torch.as_tensor(dn_stack, device=dev)
    ab = torch.as_tensor(absorption_stack, device=dev)
    if dn.shape != ab.shape:
        raise ValueError("absorption_stack and dn_stack must have the same shape")
    if dn.ndim != 3 or tuple(dn.shape[1:]) != field.spatial_shape:
        raise ValueError(f"stacks must have shape (depth, {field.spatial_shape[0]}, "
                         f"{field.spatial_shape[1]})")
    return dn, ab


def _fwd(u, dn, ab, d, want_ckpt):
    dev = u.device
    B = d.batch
    H, W = d.height, d.width
    if d.return_stack:
        out = torch.empty((d.depth, B, H, W), dtype=torch.complex64, device=dev)
    else:
        out = torch.empty((B, H, W), dtype=torch.complex64, device=dev)
    ws = torch.empty(L.lib().cxm_multislice_workspace_size(ctypes.byref(d)), dtype=torch.uint8,
                     device=dev)
    ck = None
    if want_ckpt:
        nb = L.lib().cxm_multislice_checkpoint_size(ctypes.byref(d))
        ck = torch.empty(max(1, nb), dtype=torch.uint8, device=dev)
    L.check(L.lib().cxm_multislice_fwd(L.stream_ptr(dev), u.data_ptr(), dn.data_ptr(),
                                       ab.data_ptr(), out.data_ptr(),
                                       None if ck is None else ck.data_ptr(), ws.data_ptr(),
                                       ws.numel(), ctypes.byref(d)))
    return out, ck


def _bwd(u, dn, ab, ck, cot, d, kind):
    dev = u.device
    d.grad_kind = kind
    gdn = torch.empty(dn.shape, dtype=torch.float32, device=dev)
    gab = torch.empty(ab.shape, dtype=torch.float32, device=dev)
    ws = torch.empty(L.lib().cxm_multislice_bwd_workspace_size(ctypes.byref(d)), dtype=torch.uint8,
                     device=dev)
    L.check(L.lib().cxm_multislice_bwd(L.stream_ptr(dev), u.data_ptr(), dn.data_ptr(), ab.data_ptr(),
                                       None if ck is None else ck.data_ptr(), cot.data_ptr(),
                                       gdn.data_ptr(), gab.data_ptr(), ws.data_ptr(), ws.numel(),
                                       ctypes.byref(d)))
    return gdn, gab


class _MultisliceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, dn, ab, d):
        out, ck = _fwd(u, dn, ab, d, want_ckpt=True)
        ctx.d = d
        ctx.ck = ck
        ctx.save_for_backward(u, dn, ab)
        return out

    @staticmethod
    def backward(ctx, g):
        u, dn, ab = ctx.saved_tensors
        g = g.to(torch.complex64).resolve_conj().contiguous()
        gdn, gab = _bwd(u, dn, ab, ctx.ck, g, ctx.d, L.CXM_GRAD_COMPLEX)
        return None, gdn, gab, None


def multislice_thick_sample(field: Field, absorption_stack, dn_stack, n, thickness_per_slice,
                            pad_width, NA=None, propagator=None, kykx=(0.0, 0.0),
                            reverse_propagate_distance=None, return_stack: bool = False,
                            remove_evanescent: bool = False, bandlimit: bool = False, *,
                            phase_mode=None, checkpoint_every: int = 16) -> Field:
    """Thick sample as ``depth`` thin slices: per slice an angular-spectrum
    propagation over ``thickness_per_slice`` then the slice's thin-sample
    factor; afterwards propagation by ``-reverse_propagate_distance``
    (default: half the sample thickness) through an optional NA pupil, and
    cropping of ``pad_width``.  ``return_stack`` returns the fields after
    every slice, (depth, ..., H, W).  Differentiable with respect to
    ``dn_stack`` and ``absorption_stack`` (checkpointed adjoint)."""
    if propagator is not None:
        raise NotImplementedError("multislice_thick_sample: the propagator is evaluated in the "
                                  "kernel; pass n, kykx, remove_evanescent, bandlimit instead")
    dn, ab = _stacks(field, absorption_stack, dn_stack)
    d = _desc(field, dn.shape[0], n, thickness_per_slice, pad_width, NA, kykx,
              reverse_propagate_distance, remove_evanescent, bandlimit, return_stack, phase_mode,
              checkpoint_every)
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    batch_shape = tuple(field.shape[:-2])
    H, W = field.spatial_shape
    needs_grad = (dn.requires_grad or ab.requires_grad) and torch.is_grad_enabled()
    dn32 = dn if dn.dtype == torch.float32 else dn.to(torch.float32)
    ab32 = ab if ab.dtype == torch.float32 else ab.to(torch.float32)
    dn32, ab32 = dn32.resolve_conj().contiguous(), ab32.resolve_conj().contiguous()
    if needs_grad:
        if return_stack:
            raise NotImplementedError("gradients through return_stack are not implemented")
        out = _MultisliceFn.apply(u.detach(), dn32, ab32, d)
    else:
        out, _ = _fwd(u, dn32.detach(), ab32.detach(), d, want_ckpt=False)
    if return_stack:
        return field.replace(u=out.reshape((dn.shape[0],) + batch_shape + (H, W)))
    return field.replace(u=out.reshape(batch_shape + (H, W)))


def multislice_mse_grad(field: Field, absorption_stack, dn_stack, target, n, thickness_per_slice,
                        pad_width, NA=None, kykx=(0.0, 0.0), reverse_propagate_distance=None,
                        remove_evanescent: bool = False, bandlimit: bool = False, *,
                        phase_mode=None, checkpoint_every: int = 16, checkpoints=None):
    """(d/d dn_stack, d/d absorption_stack) of mean((|out|^2 - target)^2) for
    the output of multislice_thick_sample, with the cotangent formed inside
    the adjoint.  ``checkpoints`` from :func:`multislice_checkpoints` skip the
    forward recomputation of the stored states."""
    dn, ab = _stacks(field, absorption_stack, dn_stack)
    d = _desc(field, dn.shape[0], n, thickness_per_slice, pad_width, NA, kykx,
              reverse_propagate_distance, remove_evanescent, bandlimit, False, phase_mode,
              checkpoint_every)
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    t = torch.as_tensor(target, dtype=torch.float32, device=u.device).resolve_conj().contiguous()
    if t.numel() != d.batch * d.height * d.width:
        raise ValueError("target must have the output's shape")
    return _bwd(u, dn.to(torch.float32).resolve_conj().contiguous(), ab.to(torch.float32).resolve_conj().contiguous(), checkpoints,
                t, d, L.CXM_GRAD_MSE)


def fluorescent_multislice_thick_sample(field: Field, fluorescence_stack, dn_stack, n,
                                        thickness_per_slice, pad_width, key=0, num_samples: int = 1,
                                        propagator_forward=None, propagator_backward=None,
                                        kykx=(0.0, 0.0), remove_evanescent: bool = False,
                                        bandlimit: bool = False, *, random_phases=None,
                                        phase_mode=None) -> torch.Tensor:
    """Scattered fluorescence intensity (depth, H, W) of a thick, phase-only,
    fluorescent sample (samples.py:285-441), by Monte-Carlo over random
    emitter phases.

    Per sample: the field goes forward through the slices (thin sample, plus
    fluorescence * exp(i r) with r uniform in [0, 2 pi), then an
    angular-spectrum step of +thickness), is Fourier transformed, and is
    propagated back step by step (-thickness), the intensity of each step
    accumulating into its slice.  As in the reference, the field a sample
    ends with (its backward-swept spectrum) is where the next sample starts,
    and the stack is averaged over the samples.

    ``key``: an integer seed or a ``torch.Generator`` for the random phases
    (drawn on the device with torch; JAX's key stream is not reproduced).
    ``random_phases``: optional (num_samples, depth, H, W) phases to use
    instead.  ``propagator_forward`` / ``propagator_backward``: optional
    (H + 2 pad, W + 2 pad) propagators in FFT order.
    """
    if isinstance(field, ChromaticScalarField):
        raise NotImplementedError("fluorescent multislice: monochromatic scalar fields only")
    u = field.u
    L.ensure_device(u)
    H, W = field.spatial_shape
    if u.numel() != H * W:
        raise NotImplementedError("fluorescent multislice: one field (the reference's intensity "
                                  "stack has no batch axis)")
    dev = u.device
    fl = torch.as_tensor(fluorescence_stack, dtype=torch.float32, device=dev).resolve_conj().contiguous()
    dn = torch.as_tensor(dn_stack, dtype=torch.float32, device=dev).resolve_conj().contiguous()
    if fl.ndim != 3 or tuple(fl.shape) != tuple(dn.shape) or tuple(fl.shape[1:]) != (H, W):
        raise ValueError("fluorescence_stack and dn_stack must both have shape (depth, H, W) "
                         f"with (H, W) = {(H, W)}")
    D = fl.shape[0]
    S = int(num_samples)
    if S < 1:
        raise ValueError("num_samples must be >= 1")
    if isinstance(pad_width, (tuple, list)) or np.ndim(pad_width) != 0:
        raise ValueError("pad_width must be an integer (the reference pads both axes equally)")
    pad = int(pad_width)
    Ny, Nx = H + 2 * pad, W + 2 * pad
    if random_phases is not None:
        ph_all = torch.as_tensor(random_phases, dtype=torch.float32, device=dev)
        if tuple(ph_all.shape) != (S, D, H, W):
            raise ValueError(f"random_phases must have shape {(S, D, H, W)}")
    else:
        gen = key if isinstance(key, torch.Generator) else None
        if gen is None:
            gen = torch.Generator(device=dev)
            gen.manual_seed(int(key))
        ph_all = None
    d = L.FluorDesc()
    d.height, d.width, d.depth = H, W, D
    d.pad_y = d.pad_x = pad
    d.phase_mode = L.PHASE_MODES[phase_mode or get_phase_mode()]
    d.remove_evanescent = int(bool(remove_evanescent))
    d.bandlimit = int(bool(bandlimit))
    d.dy, d.dx = float(field.dx[0]), float(field.dx[1])
    d.wavelength = float(field.spectrum.wavelength[0])
    d.n = float(n)
    d.thickness = float(thickness_per_slice)
    d.ky, d.kx = _vec2(kykx)
    d.intensity_scale = 1.0 / S
    keep = []
    for name, P in (("prop_fwd", propagator_forward), ("prop_bwd", propagator_backward)):
        if P is not None:
            Pt = torch.as_tensor(P).to(device=dev, dtype=torch.complex64).reshape(-1)
            if Pt.numel() != Ny * Nx:
                raise ValueError(f"propagators must have the padded shape {(Ny, Nx)}")
            Pt = Pt.resolve_conj().contiguous()
            keep.append(Pt)
            setattr(d, name, Pt.data_ptr())
    state = torch.zeros((Ny, Nx), dtype=torch.complex64, device=dev)
    state[pad:pad + H, pad:pad + W] = u.reshape(H, W)
    out = torch.zeros((D, H, W), dtype=torch.float32, device=dev)
    nbytes = L.lib().cxm_fluor_multislice_workspace_size(ctypes.byref(d))
    if nbytes == 0:
        L.check(L.lib().cxm_fluor_multislice_sample_c64(None, None, None, None, None, None, None, 0,
                                                         ctypes.byref(d)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    stream = L.stream_ptr(dev)
    for smp in range(S):
        if ph_all is not None:
            ph = ph_all[smp].resolve_conj().contiguous()
        else:
            ph = torch.empty((D, H, W), device=dev, dtype=torch.float32).uniform_(0.0, 2 * np.pi,
                                                                                   generator=gen)
        L.check(L.lib().cxm_fluor_multislice_sample_c64(stream, state.data_ptr(), fl.data_ptr(),
                                                         dn.data_ptr(), ph.data_ptr(), out.data_ptr(),
                                                         ws.data_ptr(), nbytes, ctypes.byref(d)))
    return out
