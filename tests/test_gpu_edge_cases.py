"""Edge cases of the propagation path on the GPU: empty batches, grids of
one row or column and odd lengths (the chirp-z passes), z = 0, a
single-wavelength chromatic field, and the largest fused grid, against
the FP64 oracle or the properties that fix the result (identity, round
trip, energy)."""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from oracle import propagation as op
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
LAM, NM = 0.532, 1.33


def _field(shape, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("method", ["asm", "transfer"])
def test_empty_batch(method):
    u = torch.zeros((0, 64, 64), dtype=torch.complex64, device="cuda")
    f = cx.ScalarField(u, 0.3, 0.0, LAM)
    fn = cf.asm_propagate if method == "asm" else cf.transfer_propagate
    out = fn(f, 10.0, NM, 16, mode="same")
    assert tuple(out.u.shape) == (0, 64, 64)
    full = fn(f, 10.0, NM, 16, mode="full")
    assert tuple(full.u.shape) == (0, 96, 96)
    assert tuple(out.intensity.shape) == (0, 64, 64)


@pytest.mark.parametrize("shape,pad", [((1, 64), 0), ((64, 1), 0), ((3, 5), 0), ((1, 1), 2),
                                       ((7, 9), (1, 3))])
@pytest.mark.parametrize("method", ["asm", "transfer"])
def test_tiny_and_odd_grids_match_oracle(shape, pad, method):
    u = _field(shape, sum(shape))
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.4, 0.0, LAM)
    fn = cf.asm_propagate if method == "asm" else cf.transfer_propagate
    out = fn(f, 3.0, NM, pad, mode="full", phase_mode="accurate").u.cpu().numpy()
    ref = op.propagate(u, (0.4, 0.4), LAM, 3.0, NM, pad, method=method, mode="full")
    err = rel_l2(out, ref)
    print(f"{method} {shape} pad {pad}: {err:.2e}")
    assert out.shape == ref.shape and err < 1e-5


@pytest.mark.parametrize("shape", [(256, 256), (100, 100)])
def test_zero_distance_is_identity(shape):
    u = _field(shape, 3)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM)
    out = cf.asm_propagate(f, 0.0, NM, 0, mode="same").u.cpu().numpy()
    assert rel_l2(out, u) < 1e-6


def test_single_wavelength_chromatic_equals_mono():
    u = _field((128, 96), 4)
    mono = cf.asm_propagate(cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM), 20.0, NM, 8)
    chrom = cf.asm_propagate(cx.ChromaticScalarField(torch.as_tensor(u[..., None]).cuda(), 0.3, 0.0,
                                                     cx.Spectrum([LAM])), 20.0, NM, 8)
    assert tuple(chrom.u.shape) == tuple(mono.u.shape) + (1,) == (144, 112, 1)
    assert torch.equal(chrom.u[..., 0], mono.u)


def test_largest_fused_grid_round_trip_and_energy():
    """8192^2, the largest fused FFT grid: a band-limited field propagated
    by z and back is the identity, and the energy is conserved (|H| = 1 on
    the propagating band)."""
    n = 8192
    x = (np.arange(n) - n / 2) * 0.3
    g = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2 * 40.0**2)).astype(np.complex64)
    f = cx.ScalarField(torch.as_tensor(g).cuda(), 0.3, 0.0, LAM)
    fwd = cf.asm_propagate(f, 200.0, NM, 0, mode="same", phase_mode="accurate")
    e0 = float(f.u.abs().square().double().sum())
    e1 = float(fwd.u.abs().square().double().sum())
    back = cf.asm_propagate(fwd, -200.0, NM, 0, mode="same", phase_mode="accurate")
    err = float((back.u - f.u).abs().square().double().sum() ** 0.5) / e0**0.5
    print(f"8192^2 round trip {err:.2e}, energy {abs(e1 - e0) / e0:.2e}")
    assert abs(e1 - e0) / e0 < 1e-5 and err < 1e-5
