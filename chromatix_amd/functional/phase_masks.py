"""Phase masks (mirrors src/chromatix/functional/phase_masks.py:29-68,
105-124): ``field * exp(i phase * lam0 / lam)`` by a libcxm kernel."""
from __future__ import annotations

import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field

__all__ = ["phase_change"]


def _phase_fwd(u, ph, meta):
    n_planes, H, W, nwv, mod = meta
    out = torch.empty_like(u)
    L.check(L.lib().cxm_phase_change_c64(L.stream_ptr(u.device), u.data_ptr(), ph.data_ptr(),
                                         out.data_ptr(), n_planes, H, W, nwv, L.doubles(mod)))
    return out


class _PhaseChangeFn(torch.autograd.Function):
    """phase_change with its vector-Jacobian product (cxm_modulation_vjp_c64):
    grad_u = g exp(-i phi m), grad_phi = sum m Im(g conj(out))."""

    @staticmethod
    def forward(ctx, u, ph, meta):
        out = _phase_fwd(u.detach(), ph.detach(), meta)
        ctx.meta = meta
        ctx.save_for_backward(ph, out)
        return out

    @staticmethod
    def backward(ctx, g):
        ph, out = ctx.saved_tensors
        n_planes, H, W, nwv, mod = ctx.meta
        g = g.to(torch.complex64).resolve_conj().contiguous()
        gu = torch.empty_like(g) if ctx.needs_input_grad[0] else None
        gph = torch.empty_like(ph) if ctx.needs_input_grad[1] else None
        L.check(L.lib().cxm_modulation_vjp_c64(
            L.stream_ptr(g.device), 0, g.data_ptr(), out.data_ptr(), ph.data_ptr(), None,
            L.dptr(gu), L.dptr(gph), None, n_planes, H, W, nwv, L.doubles(mod), 0.0))
        return gu, gph, None


def phase_change(field: Field, phase, spectrally_modulate: bool = True) -> Field:
    """Perturbs ``field`` by ``phase`` (radians, (H, W), broadcast over the
    batch and, scaled by the spectral modulation, over wavelengths).
    Differentiable in the field and the phase."""
    u = field.u.resolve_conj().contiguous()
    L.ensure_device(u)
    ph = torch.as_tensor(phase, device=u.device).to(torch.float32).resolve_conj().contiguous()
    H, W = field.spatial_shape
    if tuple(ph.shape) != (H, W):
        raise ValueError("Phase must have same height and width as incoming ``Field``.")
    nwv = field.spectrum.size if isinstance(field, ChromaticScalarField) else 1
    mod = field.spectrum.spectral_modulation if spectrally_modulate else [1.0] * nwv
    if not isinstance(field, ChromaticScalarField):
        mod = [float(mod[0])]
    meta = (u.numel() // (H * W * nwv), H, W, nwv, [float(m) for m in mod])
    if torch.is_grad_enabled() and (u.requires_grad or ph.requires_grad):
        return field.replace(u=_PhaseChangeFn.apply(u, ph, meta))
    return field.replace(u=_phase_fwd(u, ph, meta))
