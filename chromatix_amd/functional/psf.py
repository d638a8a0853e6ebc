"""Fused 4f widefield PSF stack and its phase-mask gradient on libcxm.

``psf_4f`` computes, for every defocus z_j, the composition of
Optical4FSystemPSF.__call__ (src/chromatix/systems/microscopes.py:334-355):
objective_point_source -> phase_change(phase) -> ff_lens(f_tube, n), and
optionally ``.intensity``, as two fused passes per plane (source generated
in registers).  The intensity output is differentiable with respect to
``phase`` through a hand-written adjoint (three passes per plane, no stored
residuals): ``torch.autograd`` calls ``cxm_psf4f_bwd`` with dL/dI.
``psf_4f_mse_grad`` fuses the mean-squared-error cotangent into the adjoint
(the target is read where I is formed; I is never stored).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib as L
from ..core.field import ScalarField
from .sources import psf_desc

__all__ = ["psf_4f", "psf_4f_mse_grad", "psf_output_spacing"]


def psf_output_spacing(shape, spacing, wavelength, f_tube, n) -> np.ndarray:
    """Spacing in the tube lens' back focal plane: lam f_tube / (n N dx)
    (convenience.py:33-41)."""
    dy, dx = (spacing, spacing) if np.ndim(spacing) == 0 else tuple(np.asarray(spacing).reshape(-1))
    return np.array([wavelength * f_tube / (n * shape[0] * dy),
                     wavelength * f_tube / (n * shape[1] * dx)])


def _desc(phase, spacing, wavelength, z, f, n, NA, f_tube, power, amplitude, offset, phase_mode,
          modulation=1.0):
    H, W = phase.shape[-2:]
    d, keep, zs = psf_desc((H, W), spacing, wavelength, z, f, n, NA, power, amplitude, offset,
                           phase_mode)
    d.f_tube = float(f_tube)
    d.phase_scale = float(modulation)
    return d, keep, zs


def _workspace(d, device) -> torch.Tensor:
    nbytes = L.lib().cxm_psf4f_workspace_size(ctypes.byref(d))
    if nbytes == 0:
        L.check(L.lib().cxm_psf4f_fwd(None, None, None, None, 0, ctypes.byref(d)))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


# defocus planes per library call (the per-plane parameters travel with the
# launch); longer z lists run in chunks
MAX_Z_PER_CALL = 32


def _z_chunks(args):
    """(plane slice, args with that chunk's z) for the z list in ``args``."""
    z = args[2]
    zl = np.atleast_1d(np.asarray(z, np.float64)).reshape(-1)
    if zl.size <= MAX_Z_PER_CALL:
        return [(slice(0, zl.size), args)]
    return [(slice(i, min(i + MAX_Z_PER_CALL, zl.size)),
             args[:2] + (zl[i: i + MAX_Z_PER_CALL],) + args[3:])
            for i in range(0, zl.size, MAX_Z_PER_CALL)]


def _forward_all(phase, args, out_kind) -> torch.Tensor:
    """Forward over every z chunk into one (Z, H, W) output."""
    chunks = _z_chunks(args)
    if len(chunks) == 1:
        d, keep, zs = _desc(phase, *args)
        return _forward(phase, d, out_kind)
    H, W = phase.shape[-2:]
    dtype = torch.float32 if out_kind == L.CXM_OUT_INTENSITY else torch.complex64
    out = torch.empty((chunks[-1][0].stop, H, W), dtype=dtype, device=phase.device)
    for sl, a in chunks:
        d, keep, zs = _desc(phase, *a)
        _forward(phase, d, out_kind, out[sl])
    return out


def _forward(phase, d, out_kind, out=None) -> torch.Tensor:
    d.out_kind = out_kind
    dtype = torch.float32 if out_kind == L.CXM_OUT_INTENSITY else torch.complex64
    if out is None:
        out = torch.empty((d.num_z, d.height, d.width), dtype=dtype, device=phase.device)
    ws = _workspace(d, phase.device)
    L.check(L.lib().cxm_psf4f_fwd(L.stream_ptr(phase.device), phase.data_ptr(), out.data_ptr(),
                                  ws.data_ptr(), ws.numel(), ctypes.byref(d)))
    return out


def _backward(phase, cot, d, kind, intensity=None) -> torch.Tensor:
    """dL/dphase; with ``intensity`` (a (num_z, H, W) float32 tensor) the
    forward intensity stack is stored on the way (cxm_psf4f_value_and_grad)."""
    d.grad_kind = kind
    cot = cot.to(torch.float32).resolve_conj().contiguous()
    if tuple(cot.shape) != (d.num_z, d.height, d.width):
        raise ValueError(f"cotangent / target must have shape {(d.num_z, d.height, d.width)}")
    grad = torch.empty((d.height, d.width), dtype=torch.float32, device=phase.device)
    ws = _workspace(d, phase.device)
    if intensity is None:
        L.check(L.lib().cxm_psf4f_bwd(L.stream_ptr(phase.device), phase.data_ptr(), cot.data_ptr(),
                                      grad.data_ptr(), ws.data_ptr(), ws.numel(), ctypes.byref(d)))
    else:
        L.check(L.lib().cxm_psf4f_value_and_grad(L.stream_ptr(phase.device), phase.data_ptr(),
                                                 cot.data_ptr(), grad.data_ptr(), intensity.data_ptr(),
                                                 ws.data_ptr(), ws.numel(), ctypes.byref(d)))
    return grad


class _PsfIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phase, args):
        ctx.args = args
        ctx.save_for_backward(phase)
        return _forward_all(phase, args, L.CXM_OUT_INTENSITY)

    @staticmethod
    def backward(ctx, g):
        (phase,) = ctx.saved_tensors
        grad = None
        g = g.reshape((-1,) + tuple(phase.shape[-2:]))
        for sl, a in _z_chunks(ctx.args):
            d, keep, zs = _desc(phase, *a)
            gc = _backward(phase, g[sl], d, L.CXM_GRAD_UPSTREAM)
            grad = gc if grad is None else grad.add_(gc)
        return grad, None


def _phase_tensor(phase) -> torch.Tensor:
    ph = torch.as_tensor(phase)
    if ph.dtype != torch.float32:
        ph = ph.to(torch.float32)
    if ph.ndim != 2:
        raise ValueError("phase must be a (height, width) array")
    L.ensure_device(ph)
    return ph


def psf_4f(phase, spacing, wavelength, z, f, n, NA, f_tube, power=1.0, amplitude=1.0,
           offset=(0.0, 0.0), *, output: str = "intensity", phase_mode=None, spectral_modulation=1.0):
    """PSF stack of a 4f system with ``phase`` (radians, (H, W), on the GPU)
    in the pupil plane of an objective (focal length ``f``, numerical
    aperture ``NA``, immersion index ``n``) followed by a tube lens of focal
    length ``f_tube``; ``spacing`` is the pupil-plane sample spacing.

    output="intensity": (Z, H, W) float32, differentiable w.r.t. ``phase``.
    output="field": a ScalarField (Z, H, W) with the back-focal-plane spacing.
    A scalar ``z`` drops the Z axis, like objective_point_source.
    ``spectral_modulation``: the phase enters as phase * m (phase_change's
    lam_0 / lam for one wavelength of a chromatic field)."""
    ph = phase if isinstance(phase, torch.Tensor) and phase.dtype == torch.float32 else _phase_tensor(phase)
    if ph.ndim != 2:
        raise ValueError("phase must be a (height, width) array")
    L.ensure_device(ph)
    ph_c = ph.resolve_conj().contiguous()
    args = (spacing, wavelength, z, f, n, NA, f_tube, power, amplitude, offset, phase_mode,
            spectral_modulation)
    single = np.ndim(np.asarray(z)) == 0 or np.asarray(z).size == 1
    if output == "intensity":
        if ph_c.requires_grad:
            out = _PsfIntensity.apply(ph_c, args)
        else:
            out = _forward_all(ph_c, args, L.CXM_OUT_INTENSITY)
        return out[0] if single else out
    if output == "field":
        d, keep, zs = _desc(ph_c, *args)
        u = _forward_all(ph_c.detach(), args, L.CXM_OUT_COMPLEX)
        du = psf_output_spacing(ph_c.shape, (d.dy, d.dx), wavelength, f_tube, n)
        return ScalarField(u[0] if single else u, du, 0.0, float(wavelength))
    raise ValueError('output must be "intensity" or "field"')


def psf_4f_mse_grad(phase, target, spacing, wavelength, z, f, n, NA, f_tube, power=1.0,
                    amplitude=1.0, offset=(0.0, 0.0), *, phase_mode=None, return_intensity=False,
                    spectral_modulation=1.0):
    """d/dphase of mean((psf_4f(phase, ...) - target)^2), fused: the forward
    runs inside the adjoint.  With ``return_intensity`` the forward's
    intensity stack (what psf_4f returns) comes back too, stored by the
    adjoint's column pass as it forms it -- a value-and-gradient step in one
    sweep: returns (grad, intensity)."""
    ph = _phase_tensor(phase).detach().resolve_conj().contiguous()
    args = (spacing, wavelength, z, f, n, NA, f_tube, power, amplitude, offset, phase_mode,
            spectral_modulation)
    chunks = _z_chunks(args)
    nz = chunks[-1][0].stop
    t = torch.as_tensor(target, device=ph.device).reshape((nz,) + tuple(ph.shape[-2:]))
    inten = (torch.empty((nz,) + tuple(ph.shape[-2:]), dtype=torch.float32, device=ph.device)
             if return_intensity else None)
    grad = None
    for sl, a in chunks:
        d, keep, zs = _desc(ph, *a)
        gc = _backward(ph, t[sl], d, L.CXM_GRAD_MSE, None if inten is None else inten[sl])
        if len(chunks) > 1:
            # the library's mean is over its chunk's planes
            gc.mul_((sl.stop - sl.start) / nz)
        grad = gc if grad is None else grad.add_(gc)
    if return_intensity:
        single = np.ndim(np.asarray(z)) == 0 or np.asarray(z).size == 1
        return grad, (inten[0] if single else inten)
    return grad
