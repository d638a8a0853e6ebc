"""GPU parity of thin_sample and the multislice forward / checkpointed
adjoint (cxm_thin_sample_c64, cxm_multislice_fwd / _bwd) against the
oracle: ref32 against the reference-order float32 oracle, accurate against
FP64.  Bar: relative L2 <= 1e-5 for fields and gradients."""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from oracle import multislice as om
from oracle import propagation as op
from oracle.multislice import Multislice, sample_factor
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5
LAM, NM, T = 0.532, 1.33, 0.5


def _case(B=2, h=96, w=48, D=8, seed=0):
    rng = np.random.default_rng(seed)
    u = (np.exp(2j * np.pi * rng.random((B, h, w))) * (1 + 0.2 * rng.random((B, h, w)))).astype(np.complex64)
    dn = (0.03 * rng.standard_normal((D, h, w))).astype(np.float32)
    ab = (0.005 * rng.random((D, h, w))).astype(np.float32)
    return u, dn, ab, rng


def _oracle(pm, shape, D, pad, NA=None, kykx=(0.05, -0.1), bandlimit=False, rd=None):
    return Multislice(shape, (0.3, 0.25), LAM, NM, T, D, pad, NA=NA, kykx=kykx, bandlimit=bandlimit,
                      reverse_distance=rd, precision="ref32" if pm == "ref32" else "fp64")


def _run(u, dn, ab, pm, pad, **kw):
    f = cx.ScalarField(torch.as_tensor(u).cuda(), (0.3, 0.25), 0.0, LAM)
    return cf.multislice_thick_sample(f, torch.as_tensor(ab).cuda(), torch.as_tensor(dn).cuda(), NM, T,
                                      pad, kykx=(0.05, -0.1), phase_mode=pm, **kw)


def test_thin_sample():
    u, dn, ab, _ = _case(D=1)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM)
    out = cf.thin_sample(f, torch.as_tensor(ab[0]).cuda(), torch.as_tensor(dn[0]).cuda(), T).u.cpu().numpy()
    ref = u * sample_factor(dn[0].astype(np.float64), ab[0].astype(np.float64), T, LAM, "ref32")
    assert rel_l2(out, ref) < 1e-6


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("NA", [None, 0.8])
def test_forward(pm, NA):
    u, dn, ab, _ = _case()
    out = _run(u, dn, ab, pm, (16, 8), NA=NA).u.cpu().numpy()
    ref = _oracle(pm, (96, 48), 8, (16, 8), NA=NA).forward(u, dn.astype(np.float64), ab.astype(np.float64))
    assert out.shape == (2, 96, 48)
    err = rel_l2(out, ref)
    print(f"multislice fwd {pm} NA={NA}: {err:.2e}")
    assert err < TOL


def test_forward_stack_and_reverse_distance():
    u, dn, ab, _ = _case(B=1, D=5)
    out = _run(u[0], dn, ab, "accurate", (16, 8), return_stack=True).u.cpu().numpy()
    ref = _oracle("accurate", (96, 48), 5, (16, 8)).forward(u[0], dn, ab, return_stack=True)
    assert out.shape == (5, 96, 48)
    assert rel_l2(out, ref) < TOL
    out2 = _run(u[0], dn, ab, "accurate", (16, 8), reverse_propagate_distance=-3.0).u.cpu().numpy()
    ref2 = _oracle("accurate", (96, 48), 5, (16, 8), rd=-3.0).forward(u[0], dn, ab)
    assert rel_l2(out2, ref2) < TOL


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("k", [3, 16])
def test_autograd_complex_cotangent(pm, k):
    u, dn, ab, rng = _case()
    dn_t = torch.as_tensor(dn).cuda().requires_grad_(True)
    ab_t = torch.as_tensor(ab).cuda().requires_grad_(True)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), (0.3, 0.25), 0.0, LAM)
    out = cf.multislice_thick_sample(f, ab_t, dn_t, NM, T, (16, 8), NA=0.8, kykx=(0.05, -0.1),
                                     phase_mode=pm, checkpoint_every=k)
    c = (rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)).astype(np.complex64)
    loss = torch.sum(torch.real(torch.conj(torch.as_tensor(c).cuda()) * out.u))
    loss.backward()
    gdn, gab = _oracle(pm, (96, 48), 8, (16, 8), NA=0.8).grad(u, dn.astype(np.float64), c,
                                                           ab.astype(np.float64), "complex")
    e1, e2 = rel_l2(dn_t.grad.cpu().numpy(), gdn), rel_l2(ab_t.grad.cpu().numpy(), gab)
    print(f"multislice grad {pm} k={k}: dn {e1:.2e} abs {e2:.2e}")
    assert e1 < TOL and e2 < TOL


def test_autograd_through_intensity():
    u, dn, ab, rng = _case(B=1, D=6)
    dn_t = torch.as_tensor(dn).cuda().requires_grad_(True)
    f = cx.ScalarField(torch.as_tensor(u[0]).cuda(), (0.3, 0.25), 0.0, LAM)
    out = cf.multislice_thick_sample(f, torch.as_tensor(ab).cuda(), dn_t, NM, T, (16, 8),
                                     kykx=(0.05, -0.1), phase_mode="accurate", checkpoint_every=4)
    w = rng.standard_normal((96, 48)).astype(np.float32)
    (out.intensity * torch.as_tensor(w).cuda()).sum().backward()
    gdn, _ = _oracle("accurate", (96, 48), 6, (16, 8)).grad(u[0], dn, w.astype(np.float64), ab, "upstream")
    assert rel_l2(dn_t.grad.cpu().numpy(), gdn) < TOL


@pytest.mark.parametrize("stored", [True, False])
def test_mse_gradient_fused(stored):
    u, dn, ab, rng = _case(B=2, D=7)
    m = _oracle("accurate", (96, 48), 7, (16, 8))
    I = np.abs(m.forward(u, dn, ab)) ** 2
    target = (I * rng.uniform(0.8, 1.2, I.shape)).astype(np.float32)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), (0.3, 0.25), 0.0, LAM)
    ck = None
    if stored:
        import ctypes
        from chromatix_amd import _lib as L
        from chromatix_amd.functional import samples as S
        d = S._desc(f, 7, NM, T, (16, 8), None, (0.05, -0.1), None, False, False, False, "accurate", 3)
        _, ck = S._fwd(f.u, torch.as_tensor(dn).cuda(), torch.as_tensor(ab).cuda(), d, want_ckpt=True)
    gdn, gab = cf.multislice_mse_grad(f, torch.as_tensor(ab).cuda(), torch.as_tensor(dn).cuda(),
                                      torch.as_tensor(target).cuda(), NM, T, (16, 8),
                                      kykx=(0.05, -0.1), phase_mode="accurate", checkpoint_every=3,
                                      checkpoints=ck)
    rdn, rab = m.grad(u, dn, target.astype(np.float64), ab, "mse")
    # the cotangent 2 (I - target) amplifies the forward's float32 error by
    # kappa = |I| / |I - target| (~10 here); the bar is 1e-5 times kappa
    tol = TOL * _kappa(I, target)
    assert rel_l2(gdn.cpu().numpy(), rdn) < tol
    assert rel_l2(gab.cpu().numpy(), rab) < tol


def _kappa(I, target):
    return max(1.0, float(np.linalg.norm(I) / np.linalg.norm(I - target)))


def test_c3_shape_against_oracle():
    """1024^2, 24 slices of the C3 optics (dx 0.3 um is an assumption of the
    test), one field; and the adjoint at k = 8."""
    rng = np.random.default_rng(3)
    N, D = 1024, 24
    u = np.ones((N, N), np.complex64)
    dn = (0.01 * rng.standard_normal((D, N, N))).astype(np.float32)
    ab = np.zeros_like(dn)
    m = Multislice((N, N), (0.3, 0.3), LAM, NM, T, D, 0, precision="fp64")
    ref = m.forward(u, dn, ab)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM)
    out = cf.multislice_thick_sample(f, torch.as_tensor(ab).cuda(), torch.as_tensor(dn).cuda(), NM, T, 0,
                                     phase_mode="accurate").u.cpu().numpy()
    assert rel_l2(out, ref) < TOL
    target = (np.abs(ref) ** 2 * rng.uniform(0.9, 1.1, ref.shape)).astype(np.float32)
    gdn, _ = cf.multislice_mse_grad(f, torch.as_tensor(ab).cuda(), torch.as_tensor(dn).cuda(),
                                    torch.as_tensor(target).cuda(), NM, T, 0, phase_mode="accurate",
                                    checkpoint_every=8)
    rdn, _ = m.grad(u, dn, target.astype(np.float64), ab, "mse")
    assert rel_l2(gdn.cpu().numpy(), rdn) < TOL * _kappa(np.abs(ref) ** 2, target)


def test_uniform_slab_closed_form_256_slices():
    """Plane wave, no padding, uniform dn, 256 slices of 1024^2: the output is
    exp(i 2 pi / lam (D t n + D t dn - (D t / 2) n)) everywhere.  The error
    of the same loop run in complex64 with scipy.fft (a float32 restatement
    of the reference algorithm) is printed beside it for scale."""
    import scipy.fft as sfft

    from oracle.propagation import asm_propagator

    N, D, dn0 = 1024, 256, np.float32(0.02)
    f = cx.ScalarField(torch.ones((N, N), dtype=torch.complex64).cuda(), 0.3, 0.0, LAM)
    dn = torch.full((D, N, N), float(dn0), dtype=torch.float32).cuda()
    out = cf.multislice_thick_sample(f, torch.zeros_like(dn), dn, NM, T, 0, phase_mode="accurate").u
    out = out.cpu().numpy()
    phase = 2 * np.pi / LAM * (D * T * NM + D * T * np.float64(dn0) - D * T / 2 * NM)
    ref = np.full((N, N), np.exp(1j * phase))
    err = rel_l2(out, ref)
    # complex64 restatement
    H = asm_propagator((N, N), (0.3, 0.3), LAM, T, NM).astype(np.complex64)
    Hb = asm_propagator((N, N), (0.3, 0.3), LAM, -D * T / 2, NM).astype(np.complex64)
    Ts = np.complex64(np.exp(1j * 2 * np.pi * np.float64(dn0) * T / LAM))
    w = os_workers()
    v = np.ones((N, N), np.complex64)
    for _ in range(D):
        v = sfft.ifft2(sfft.fft2(v, workers=w) * H, workers=w) * Ts
    v = sfft.ifft2(sfft.fft2(v, workers=w) * Hb, workers=w)
    err32 = rel_l2(v, ref)
    print(f"uniform slab D=256: GPU {err:.2e}, complex64 scipy {err32:.2e}, "
          f"|out| mean {np.abs(out).mean():.8f}")
    assert err < TOL


def os_workers():
    import os

    return max(1, min(16, len(os.sched_getaffinity(0))))


def _fluor_case(H, D, seed):
    rng = np.random.default_rng(seed)
    u = np.exp(2j * np.pi * 0.05 * rng.standard_normal((H, H))).astype(np.complex64)
    fl = rng.random((D, H, H)).astype(np.float32)
    dn = (0.02 * rng.standard_normal((D, H, H))).astype(np.float32)
    return u, fl, dn, rng


@pytest.mark.parametrize("samples", [1, 2])
def test_fluorescent_multislice_matches_oracle(samples):
    """fluorescent_multislice_thick_sample with given random phases against
    the FP64 oracle (accurate propagators), including the reference's carry
    of the field from one Monte-Carlo sample into the next."""
    H, D, pad, t = 256, 6, 32, 1.5
    u, fl, dn, rng = _fluor_case(H, D, samples)
    ph = rng.uniform(0, 2 * np.pi, (samples, D, H, H)).astype(np.float32)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
    got = cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, t, pad, num_samples=samples,
                                                 random_phases=ph, phase_mode="accurate")
    ref = om.fluorescent_multislice(u, fl.astype(np.float64), dn.astype(np.float64), (0.3, 0.3),
                                    0.532, 1.33, t, pad, ph.astype(np.float64))
    err = rel_l2(got.cpu().numpy(), ref)
    print(f"fluorescent multislice {H}^2 x {D}, {samples} sample(s): {err:.2e}")
    assert got.shape == (D, H, H)
    assert err < TOL


def test_fluorescent_multislice_ref32_and_given_propagators():
    """Default (ref32) propagators against the oracle run with the
    reference-order float32 propagators, and caller-supplied propagators
    used as given."""
    H, D, pad, t = 192, 4, 32, 2.0
    u, fl, dn, rng = _fluor_case(H, D, 7)
    ph = rng.uniform(0, 2 * np.pi, (1, D, H, H)).astype(np.float32)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.25, 0.0, 0.532)
    Ny = H + 2 * pad
    Pf = op.asm_propagator((Ny, Ny), (0.25, 0.25), 0.532, t, 1.33, precision="ref32")
    Pb = op.asm_propagator((Ny, Ny), (0.25, 0.25), 0.532, -t, 1.33, precision="ref32")
    ref = om.fluorescent_multislice(u, fl.astype(np.float64), dn.astype(np.float64), (0.25, 0.25),
                                    0.532, 1.33, t, pad, ph.astype(np.float64), prop_fwd=Pf, prop_bwd=Pb)
    got = cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, t, pad, random_phases=ph)
    assert rel_l2(got.cpu().numpy(), ref) < TOL
    # the same propagators passed in (complex64) give the same result
    got2 = cf.fluorescent_multislice_thick_sample(
        f, fl, dn, 1.33, t, pad, random_phases=ph,
        propagator_forward=torch.as_tensor(Pf.astype(np.complex64)).cuda(),
        propagator_backward=torch.as_tensor(Pb.astype(np.complex64)).cuda())
    assert rel_l2(got2.cpu().numpy(), ref) < TOL


def test_fluorescent_multislice_seeded_monte_carlo():
    """Phases drawn on the device: the same seed gives the same stack, another
    seed another one; the stack is non-negative."""
    H, D = 128, 3
    u, fl, dn, _ = _fluor_case(H, D, 3)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
    a = cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, 1.0, 16, key=5, num_samples=3)
    b = cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, 1.0, 16, key=5, num_samples=3)
    c = cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, 1.0, 16, key=6, num_samples=3)
    assert torch.equal(a, b)
    assert not torch.equal(a, c)
    assert bool((a >= 0).all()) and bool(torch.isfinite(a).all())


@pytest.mark.parametrize("pm", ["ref32", "accurate"])
@pytest.mark.parametrize("shape,pad", [((480, 448), (16, 32)), ((1024, 1024), 0), ((2000, 2016), (24, 16))])
def test_dedicated_passes_match_pipeline_passes(pm, shape, pad):
    """The dedicated slice row passes and the H tables (CXM_OPT_MS_PASSES 0)
    against the operator-pipeline passes with H per bin (1): forward field
    and the dn / absorption gradients, with absorption, kykx and an NA mask
    on the final back-propagation (padded grids 512^2, 1024^2, 2048^2)."""
    from chromatix_amd import _lib as L

    rng = np.random.default_rng(21)
    h, w = shape
    D = 6
    u = (np.exp(2j * np.pi * rng.random((h, w)))).astype(np.complex64)
    dn = (0.02 * rng.standard_normal((D, h, w))).astype(np.float32)
    ab = (0.002 * rng.random((D, h, w))).astype(np.float32)
    c = (rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))).astype(np.complex64)
    outs, gdns, gabs = [], [], []
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_MS_PASSES, opt))
            dn_t = torch.as_tensor(dn).cuda().requires_grad_(True)
            ab_t = torch.as_tensor(ab).cuda().requires_grad_(True)
            f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, LAM)
            out = cf.multislice_thick_sample(f, ab_t, dn_t, NM, T, pad, NA=0.8, kykx=(0.05, -0.1),
                                             phase_mode=pm, checkpoint_every=4)
            loss = torch.sum(torch.real(torch.conj(torch.as_tensor(c).cuda()) * out.u))
            loss.backward()
            outs.append(out.u.detach().cpu().numpy())
            gdns.append(dn_t.grad.cpu().numpy())
            gabs.append(ab_t.grad.cpu().numpy())
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_MS_PASSES, 0))
    e = (rel_l2(outs[0], outs[1]), rel_l2(gdns[0], gdns[1]), rel_l2(gabs[0], gabs[1]))
    print(f"dedicated vs pipeline {pm} {shape} pad {pad}: fwd {e[0]:.2e} dn {e[1]:.2e} abs {e[2]:.2e}")
    assert max(e) < 1e-6


@pytest.mark.parametrize("H,pad", [(448, 32), (512, 64), (1024, 0)])
def test_dedicated_fluorescent_passes_match_pipeline(H, pad):
    """The dedicated fluorescent passes (the slice's thin sample and
    fluorescence between the row transforms; the backward sweep batched
    over its steps) give the operator pipeline's bits (CXM_OPT_MS_PASSES 1)
    at the pipeline's row shapes (CXM_OPT_FL_ROW_E 20), at padded lengths
    512, 640 and 1024, two Monte-Carlo samples; the default 640-point shape
    (10 values per thread) agrees to float32 rounding; and the 512 case
    matches the oracle."""
    from chromatix_amd import _lib as L

    D, samples, t = 3, 2, 1.5
    u, fl, dn, rng = _fluor_case(H, D, samples)
    ph = rng.uniform(0, 2 * np.pi, (samples, D, H, H)).astype(np.float32)
    outs = []
    try:
        for ms, row_e in ((0, 20), (1, 0), (0, 0)):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_MS_PASSES, ms))
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_ROW_E, row_e))
            f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
            outs.append(cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, t, pad, num_samples=samples,
                                                               random_phases=ph, phase_mode="accurate"))
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_MS_PASSES, 0))
        L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_ROW_E, 0))
    e = rel_l2(outs[2].cpu().numpy(), outs[1].cpu().numpy())
    print(f"dedicated vs pipeline fluorescent {H + 2 * pad}: {bool(torch.equal(outs[0], outs[1]))}, "
          f"default shape {e:.2e}")
    assert torch.equal(outs[0], outs[1])
    assert e < 1e-6
    if H + 2 * pad == 512:
        ref = om.fluorescent_multislice(u, fl.astype(np.float64), dn.astype(np.float64), (0.3, 0.3),
                                        0.532, 1.33, t, pad, ph.astype(np.float64))
        assert rel_l2(outs[0].cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("chunk,batch,block", [(1, 0, 0), (3, 0, 3), (2, 3, 0), (3, 2, 2), (1, 1, 1)])
def test_batched_fluorescent_backward_chunking(chunk, batch, block):
    """The batched backward sweep gives the step-by-step pipeline's bits for
    every split of its steps: steps per workgroup (CXM_OPT_FL_BWD_CHUNK,
    ragged last chunk) and steps per launch pair (CXM_OPT_FL_BWD_BATCH: 3
    batches of a 7-slice sample end with the carried spectrum in the second
    buffer and copied back, 4 end in the caller's; one step per batch), two
    Monte-Carlo samples so the carried spectrum is checked too; and the
    slice factors precomputed in blocks of 1, 2, 3 or all 7 slices
    (CXM_OPT_FL_FACTOR_BLOCK; the pipeline's 640-point row shape,
    CXM_OPT_FL_ROW_E 20)."""
    from chromatix_amd import _lib as L

    H, pad, D, samples, t = 448, 32, 7, 2, 1.5
    u, fl, dn, rng = _fluor_case(H, D, samples)
    ph = rng.uniform(0, 2 * np.pi, (samples, D, H, H)).astype(np.float32)
    outs = []
    try:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_ROW_E, 20))
        for ms, c, b in ((1, 0, 0), (0, chunk, batch)):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_MS_PASSES, ms))
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_BWD_CHUNK, c))
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_BWD_BATCH, b))
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FL_FACTOR_BLOCK, block if ms == 0 else 0))
            f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
            outs.append(cf.fluorescent_multislice_thick_sample(f, fl, dn, 1.33, t, pad, num_samples=samples,
                                                               random_phases=ph, phase_mode="accurate"))
    finally:
        for key in (L.CXM_OPT_MS_PASSES, L.CXM_OPT_FL_BWD_CHUNK, L.CXM_OPT_FL_BWD_BATCH,
                    L.CXM_OPT_FL_ROW_E, L.CXM_OPT_FL_FACTOR_BLOCK):
            L.check(L.lib().cxm_set_option(key, 0))
    print(f"batched backward chunk {chunk} batch {batch} factor block {block}: "
          f"{bool(torch.equal(outs[0], outs[1]))}")
    assert torch.equal(outs[0], outs[1])
