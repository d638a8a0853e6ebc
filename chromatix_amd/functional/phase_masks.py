"""Phase masks (mirrors src/chromatix/functional/phase_masks.py:29-68,
105-124): ``field * exp(i phase * lam0 / lam)`` by a libcxm kernel."""
from __future__ import annotations

import torch

from .. import _lib as L
from ..core.field import ChromaticScalarField, Field

__all__ = ["phase_change"]


def phase_change(field: Field, phase, spectrally_modulate: bool = True) -> Field:
    """Perturbs ``field`` by ``phase`` (radians, (H, W), broadcast over the
    batch and, scaled by the spectral modulation, over wavelengths)."""
    L.refuse_grad("phase_change", field.u, phase)
    u = field.u.contiguous()
    L.ensure_device(u)
    ph = torch.as_tensor(phase, dtype=torch.float32, device=u.device).contiguous()
    H, W = field.spatial_shape
    if tuple(ph.shape) != (H, W):
        raise ValueError("Phase must have same height and width as incoming ``Field``.")
    nwv = field.spectrum.size if isinstance(field, ChromaticScalarField) else 1
    mod = field.spectrum.spectral_modulation if spectrally_modulate else [1.0] * nwv
    if not isinstance(field, ChromaticScalarField):
        mod = [float(mod[0])]
    n_planes = u.numel() // (H * W * nwv)
    out = torch.empty_like(u)
    L.check(L.lib().cxm_phase_change_c64(L.stream_ptr(u.device), u.data_ptr(), ph.data_ptr(),
                                         out.data_ptr(), n_planes, H, W, nwv, L.doubles(mod)))
    return field.replace(u=out)
