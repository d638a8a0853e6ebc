"""GPU parity of the fused 2D propagation (cxm_prop2d_c64) against the oracle.

ref32 mode is compared with the oracle's reference-order float32 propagator
wrapped in exact FFTs (so the difference is the GPU's complex64 FFT
rounding); accurate mode is compared with the FP64 oracle.  Bar: relative
L2 <= 1e-5 for complex64.
"""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from chromatix_amd import _lib as L
from oracle import propagation as op
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5
LAM, NM = 0.532, 1.33


def _phase_field(shape, seed=0, batch=()):
    rng = np.random.default_rng(seed)
    return np.exp(1j * rng.uniform(0, 2 * np.pi, size=batch + tuple(shape))).astype(np.complex64)


def _mono(u, dx):
    return cx.ScalarField(torch.as_tensor(u).cuda(), dx, 0.0, LAM)


def _oracle(u, dx, z, pad, method, mode, precision, **kw):
    return op.propagate(u, (dx, dx) if np.isscalar(dx) else dx, LAM, z, NM, pad, method=method,
                        mode=mode, precision=precision, fft_precision="fp64", **kw)


def _run(u, dx, z, pad, method, mode, phase_mode, **kw):
    f = _mono(u, dx)
    fn = cf.asm_propagate if method == "asm" else cf.transfer_propagate
    out = fn(f, z, NM, pad, mode=mode, phase_mode=phase_mode, **kw)
    return out.u.cpu().numpy()


CASES = [
    # shape, dx, z, pad, method, mode
    ((64, 64), 0.3, 10.0, 0, "asm", "same"),
    ((100, 60), 0.3, 25.0, (14, 2), "asm", "same"),
    ((100, 60), 0.3, 25.0, (14, 2), "transfer", "full"),
    ((512, 512), 0.3, 100.0, 256, "asm", "same"),  # C1 geometry
    ((512, 512), 0.3, 100.0, 256, "transfer", "same"),
    ((256, 256), 40 / 256, 100.0, 512, "asm", "same"),  # tests/test_propagate.py shape -> 1280^2 grid
    ((160, 160), 0.25, 15.0, 0, "transfer", "full"),
    ((2, 2), 0.1, 1.0, 0, "asm", "full"),
    ((4, 8), 0.2, 3.0, (2, 0), "asm", "full"),
    ((64, 32), 0.1, 2.0, 0, "asm", "same"),  # evanescent regime
    ((128, 128), 0.3, -40.0, 0, "asm", "same"),  # negative z
    ((1024, 1024), 0.3, 100.0, 0, "asm", "same"),  # headline plane
    ((80, 2560), 0.3, 50.0, 0, "asm", "same"),
    ((8192, 16), 0.3, 50.0, 0, "transfer", "same"),
    ((16, 4096), 0.3, 50.0, 0, "asm", "same"),
]


@pytest.mark.parametrize("phase_mode", ["ref32", "accurate"])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0][0]}x{c[0][1]}-{c[4]}-{c[5]}-z{c[2]}" for c in CASES])
def test_propagation_matches_oracle(case, phase_mode):
    shape, dx, z, pad, method, mode = case
    u = _phase_field(shape, seed=hash(shape) % 1000)
    got = _run(u, dx, z, pad, method, mode, phase_mode)
    ref = _oracle(u, dx, z, pad, method, mode, "ref32" if phase_mode == "ref32" else "fp64")
    err = rel_l2(got, ref)
    print(f"{case} {phase_mode}: rel_l2 = {err:.3e}")
    assert got.shape == ref.shape
    assert err < TOL


@pytest.mark.parametrize("phase_mode", ["ref32", "accurate"])
@pytest.mark.parametrize("opts", [
    dict(kykx=(0.4, -0.7)),
    dict(shift_yx=(3.0, -5.5)),
    dict(bandlimit=True),
    dict(remove_evanescent=True),
    dict(bandlimit=True, shift_yx=(20.0, 10.0)),
])
def test_asm_options(opts, phase_mode):
    shape, dx, z = (128, 96), 0.15, 800.0
    u = _phase_field(shape, seed=3)
    got = _run(u, dx, z, 16, "asm", "full", phase_mode, **opts)
    ref = _oracle(u, dx, z, 16, "asm", "full", "ref32" if phase_mode == "ref32" else "fp64", **opts)
    err = rel_l2(got, ref)
    print(f"{opts} {phase_mode}: rel_l2 = {err:.3e}")
    assert err < TOL


def test_transfer_kykx():
    u = _phase_field((64, 64), seed=4)
    got = _run(u, 0.3, 30.0, 0, "transfer", "same", "ref32", kykx=(0.2, 0.1))
    ref = _oracle(u, 0.3, 30.0, 0, "transfer", "same", "ref32", kykx=(0.2, 0.1))
    assert rel_l2(got, ref) < TOL


def test_z_batch_outer_and_zip():
    zs = np.array([-20.0, 5.0, 12.5, 40.0])
    u = _phase_field((112, 48), seed=5)
    f = _mono(u, 0.3)
    out = cf.asm_propagate(f, zs, NM, 8, mode="same")
    assert out.shape == (4, 112, 48)
    ref = np.stack([_oracle(u, 0.3, z, 8, "asm", "same", "ref32") for z in zs])
    assert rel_l2(out.u.cpu().numpy(), ref) < TOL
    # zip: innermost batch equals the number of distances
    ub = _phase_field((128, 64), seed=6, batch=(2, 4))
    fb = _mono(ub, 0.3)
    outb = cf.transfer_propagate(fb, zs, NM, 0, mode="same")
    assert outb.shape == (2, 4, 128, 64)
    refb = np.stack([[_oracle(ub[a, k], 0.3, zs[k], 0, "transfer", "same", "ref32") for k in range(4)]
                     for a in range(2)])
    assert rel_l2(outb.u.cpu().numpy(), refb) < TOL
    # outer with a trailing singleton batch
    u1 = _phase_field((32, 32), seed=7, batch=(3, 1))
    o1 = cf.asm_propagate(_mono(u1, 0.3), zs, NM, 0)
    assert o1.shape == (3, 4, 32, 32)


def test_chromatic_fields():
    wl = np.array([0.45, 0.532, 0.65])
    rng = np.random.default_rng(8)
    u = np.exp(2j * np.pi * rng.random((64, 48, 3))).astype(np.complex64)
    dx = np.array([[0.3, 0.3], [0.31, 0.29], [0.35, 0.35]])
    f = cx.ChromaticScalarField(torch.as_tensor(u).cuda(), dx, 0.0, cx.Spectrum(wl))
    for pm, prec in (("ref32", "ref32"), ("accurate", "fp64")):
        out = cf.asm_propagate(f, 30.0, NM, 8, mode="same", phase_mode=pm).u.cpu().numpy()
        ref = np.stack([op.propagate(u[..., w], dx[w], wl[w], 30.0, NM, 8, precision=prec,
                                     fft_precision="fp64") for w in range(3)], axis=-1)
        assert rel_l2(out, ref) < TOL
    # the reference tests' (wavelength, 1.0) spectrum: a chromatic field with wv = 1
    f1 = cx.Field.build(torch.as_tensor(u[..., :1]).cuda(), 0.3, (0.532, 1.0))
    assert isinstance(f1, cx.ChromaticScalarField)
    o1 = cf.transfer_propagate(f1, 10.0, NM, 8, mode="same").u.cpu().numpy()
    r1 = op.propagate(u[..., 0], (0.3, 0.3), 0.532, 10.0, NM, 8, method="transfer",
                      precision="ref32", fft_precision="fp64")
    assert rel_l2(o1[..., 0], r1) < TOL


def test_chromatic_z_batch():
    wl = np.array([0.5, 0.6])
    zs = np.array([10.0, 20.0, 30.0])
    rng = np.random.default_rng(9)
    u = np.exp(2j * np.pi * rng.random((32, 32, 2))).astype(np.complex64)
    f = cx.ChromaticScalarField(torch.as_tensor(u).cuda(), 0.25, 0.0, cx.Spectrum(wl))
    out = cf.transfer_propagate(f, zs, NM, 0).u.cpu().numpy()
    assert out.shape == (3, 32, 32, 2)
    for k, z in enumerate(zs):
        for w in range(2):
            r = op.propagate(u[..., w], (0.25, 0.25), wl[w], z, NM, 0, method="transfer",
                             mode="full", precision="ref32", fft_precision="fp64")
            assert rel_l2(out[k, ..., w], r) < TOL


def test_materialised_propagator_matches_oracle():
    # dx = 0.22 keeps every bin away from the evanescent boundary |f| = n/lam,
    # where the propagator is ill-conditioned for any implementation
    f = _mono(_phase_field((128, 80)), 0.22)
    for pm, prec in (("ref32", "ref32"), ("accurate", "fp64")):
        H = cf.compute_asm_propagator(f, 55.0, NM, shift_yx=(1.0, -2.0), phase_mode=pm).cpu().numpy()
        Hr = op.asm_propagator((128, 80), (0.22, 0.22), LAM, 55.0, NM, shift_yx=(1.0, -2.0), precision=prec)
        assert np.max(np.abs(H - Hr)) < 2e-6
        Ht = cf.compute_transfer_propagator(f, [3.0, -8.0], NM, phase_mode=pm).cpu().numpy()
        Htr = op.transfer_propagator((128, 80), (0.22, 0.22), LAM, [3.0, -8.0], NM, precision=prec)
        assert Ht.shape == (2, 128, 80)
        assert np.max(np.abs(Ht - Htr)) < 2e-6


def test_kernel_propagate_with_array():
    rng = np.random.default_rng(10)
    u = (rng.standard_normal((3, 64, 40)) + 1j * rng.standard_normal((3, 64, 40))).astype(np.complex64)
    H = np.exp(2j * np.pi * rng.random((64, 40))).astype(np.complex64)
    f = _mono(u, 0.3)
    out = cf.kernel_propagate(f, torch.as_tensor(H).cuda()).u.cpu().numpy()
    ref = op.kernel_propagate(u, H)
    assert rel_l2(out, ref) < TOL
    # per-plane propagators broadcast against the batch
    Hb = np.exp(2j * np.pi * rng.random((3, 64, 40))).astype(np.complex64)
    outb = cf.kernel_propagate(f, torch.as_tensor(Hb).cuda()).u.cpu().numpy()
    assert rel_l2(outb, op.kernel_propagate(u, Hb)) < TOL


def test_fused_intensity_c1():
    """Config C1: 512^2 exp(i phi) (seed 0), pad 256 -> 1024^2, intensity."""
    u = _phase_field((512, 512), seed=0)
    f = _mono(u, 0.3)
    I = cf.propagate_intensity(f, 100.0, NM, 256, method="asm").cpu().numpy()
    I2 = cf.asm_propagate(f, 100.0, NM, 256, mode="same").intensity.cpu().numpy()
    ref32 = np.abs(_oracle(u, 0.3, 100.0, 256, "asm", "same", "ref32")) ** 2
    ref64 = np.abs(_oracle(u, 0.3, 100.0, 256, "asm", "same", "fp64")) ** 2
    e32 = rel_l2(I, ref32)
    print(f"C1 intensity ref32 vs ref32-oracle {e32:.2e}; vs fp64 {rel_l2(I, ref64):.2e}")
    assert e32 < TOL
    assert rel_l2(I2, I) < 1e-6
    Ia = cf.propagate_intensity(f, 100.0, NM, 256, phase_mode="accurate").cpu().numpy()
    assert rel_l2(Ia, ref64) < TOL


def test_intensity_and_power():
    rng = np.random.default_rng(11)
    u = (rng.standard_normal((2, 33, 17)) + 1j * rng.standard_normal((2, 33, 17))).astype(np.complex64)
    f = _mono(u, (0.3, 0.2))
    np.testing.assert_allclose(f.intensity.cpu().numpy(), np.abs(u) ** 2, rtol=1e-6)
    p = f.power.cpu().numpy()
    assert p.shape == (2, 1, 1)
    np.testing.assert_allclose(p[:, 0, 0], 0.06 * np.sum(np.abs(u) ** 2, axis=(1, 2)), rtol=1e-6)
    uc = (rng.standard_normal((5, 6, 3)) + 1j * rng.standard_normal((5, 6, 3))).astype(np.complex64)
    fc = cx.ChromaticScalarField(torch.as_tensor(uc).cuda(), 0.5, 0.0, cx.Spectrum([0.4, 0.5, 0.6], [1, 2, 1]))
    dens = np.array([0.25, 0.5, 0.25])
    np.testing.assert_allclose(fc.intensity.cpu().numpy(), np.sum(dens * np.abs(uc) ** 2, -1), rtol=1e-6)


def test_headline_batch_properties():
    """64 x 1024^2 (the benchmark batch): spot planes vs the oracle and the
    z / -z round trip (remove_evanescent) over the whole batch."""
    B = 64
    rng = np.random.default_rng(0)
    u = torch.exp(1j * torch.as_tensor(rng.uniform(0, 2 * np.pi, (B, 1024, 1024)), dtype=torch.float32)).cuda()
    f = _mono(u, 0.3)
    out = cf.asm_propagate(f, 100.0, NM, 0, mode="same")
    for b in (0, 37, B - 1):
        ref = _oracle(u[b].cpu().numpy(), 0.3, 100.0, 0, "asm", "same", "ref32")
        assert rel_l2(out.u[b].cpu().numpy(), ref) < TOL
    fwd = cf.asm_propagate(f, 100.0, NM, 0, remove_evanescent=True)
    back = cf.asm_propagate(fwd, -100.0, NM, 0, remove_evanescent=True)
    err = (torch.linalg.vector_norm(back.u - u) / torch.linalg.vector_norm(u)).item()
    assert err < TOL


@pytest.mark.parametrize("phase_mode", ["ref32", "accurate"])
@pytest.mark.parametrize("method,shape,pad,opts", [
    ("asm", (100, 60), (14, 2), {}),
    ("asm", (128, 96), 16, dict(bandlimit=True, shift_yx=(20.0, 10.0))),
    ("asm", (64, 32), 0, {}),  # evanescent regime
    ("transfer", (256, 256), 0, dict(kykx=(0.2, 0.1))),
])
def test_h_table_is_bit_identical(method, shape, pad, opts, phase_mode):
    """A batch sharing one parameter set reads H from a table evaluated once
    per call; the column pass evaluating H on the fly gives the same bits."""
    u = _phase_field(shape, seed=11, batch=(L.CXM_H_TABLE_MIN_PLANES + 1,))
    z, dx = (2.0, 0.1) if shape == (64, 32) else (40.0, 0.3)
    lib = L.lib()
    try:
        outs = []
        for opt in (1, 0, 2):  # never, automatic (batch >= min), always
            L.check(lib.cxm_set_option(L.CXM_OPT_H_TABLE, opt))
            uu = u if opt != 2 else u[:1]
            outs.append(_run(uu, dx, z, pad, method, "same", phase_mode, **opts))
    finally:
        L.check(lib.cxm_set_option(L.CXM_OPT_H_TABLE, 0))
    fly, tab, tab1 = outs
    assert np.array_equal(fly.view(np.uint32), tab.view(np.uint32))
    assert np.array_equal(fly[:1].view(np.uint32), tab1.view(np.uint32))
    ref = _oracle(u[2], dx, z, pad, method, "same", "ref32" if phase_mode == "ref32" else "fp64", **opts)
    assert rel_l2(tab[2], ref) < TOL


@pytest.mark.parametrize("shape,pad,mode,batch", [
    ((1024, 1024), 0, "same", 3),
    ((512, 512), 256, "full", 5),   # 512 input rows -> 1024 output rows, in place
    ((1000, 1000), 12, "same", 2),
    ((1022, 1022), 1, "same", 2),   # rows not a multiple of 4: row-major intermediates
    ((1536, 1536), 0, "same", 2),
    ((1024, 1536), 0, "same", 5),
])
def test_tiled_intermediates_are_bit_identical(shape, pad, mode, batch):
    """Intermediate planes in strips of 4 columns (where the pass shapes allow
    it) and the looping row pass give the same bits as row-major
    intermediates and one row group per workgroup, batch planes included
    (their workspace planes must not overlap when input and output rows
    differ); so does the fused intensity output."""
    u = _phase_field(shape, seed=12, batch=(batch,))
    lib = L.lib()
    # (tiles, row loop): plain, default, looping everywhere (3 per CU)
    combos = [(1, 1), (0, 0), (0, 3), (1, 3)]
    try:
        outs, ints = [], []
        for tiles, loop in combos:
            L.check(lib.cxm_set_option(L.CXM_OPT_TILES, tiles))
            L.check(lib.cxm_set_option(L.CXM_OPT_ROW_LOOP, loop))
            outs.append(_run(u, 0.3, 60.0, pad, "asm", mode, "ref32"))
            ints.append(cf.propagate_intensity(_mono(u, 0.3), 60.0, NM, pad, method="asm",
                                               mode=mode).cpu().numpy())
    finally:
        L.check(lib.cxm_set_option(L.CXM_OPT_TILES, 0))
        L.check(lib.cxm_set_option(L.CXM_OPT_ROW_LOOP, 0))
    for o, i in zip(outs[1:], ints[1:]):
        assert np.array_equal(outs[0].view(np.uint32), o.view(np.uint32))
        assert np.array_equal(ints[0].view(np.uint32), i.view(np.uint32))
    for b in (0, batch - 1):
        ref = _oracle(u[b], 0.3, 60.0, pad, "asm", mode, "ref32")
        assert rel_l2(outs[1][b], ref) < TOL


def test_errors_are_raised():
    f = _mono(_phase_field((6000, 8)), 0.3)
    with pytest.raises(L.CxmError):
        cf.asm_propagate(f, 1.0, NM, 0)  # 6000 + 6000 - 1 exceeds the chirp-z transform
    with pytest.raises(ValueError):
        cf.asm_propagate(_mono(_phase_field((8, 8), batch=(3,)), 0.3), [1.0, 2.0], NM, 0)


@pytest.mark.parametrize("method", ["asm", "transfer"])
@pytest.mark.parametrize("mode", ["same", "full"])
def test_autograd_adjoint_identity(method, mode):
    """<P u, c> = <u, P^H c> with P^H from the autograd rule (the same
    kernels with -z, pad and crop exchanged), checked against the FP64
    oracle adjoint."""
    rng = np.random.default_rng(11)
    u = (rng.standard_normal((2, 96, 80)) + 1j * rng.standard_normal((2, 96, 80))).astype(np.complex64)
    ut = torch.as_tensor(u).cuda().requires_grad_(True)
    f = cx.ScalarField(ut, (0.25, 0.3), 0.0, 0.532)
    fn = cf.asm_propagate if method == "asm" else cf.transfer_propagate
    out = fn(f, 40.0, 1.33, (16, 24), mode=mode, phase_mode="accurate")
    c = (rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)).astype(np.complex64)
    loss = torch.sum(torch.real(torch.conj(torch.as_tensor(c).cuda()) * out.u))
    loss.backward()
    g = ut.grad.cpu().numpy()
    # FP64: P = crop_c(ifft2(H fft2(pad_p(u)))), P^H c = crop_p(ifft2(conj(H) fft2(pad_c(c))))
    py, px = 16, 24
    cy, cx_ = (py, px) if mode == "same" else (0, 0)
    Ny, Nx = 96 + 2 * py, 80 + 2 * px
    H = (op.asm_propagator if method == "asm" else op.transfer_propagator)((Ny, Nx), (0.25, 0.3), 0.532, 40.0, 1.33)
    cp = np.pad(c.astype(np.complex128), [(0, 0), (cy, cy), (cx_, cx_)])
    adj = np.fft.ifft2(np.conj(H) * np.fft.fft2(cp), axes=(-2, -1))[..., py: Ny - py, px: Nx - px]
    # the real-valued loss Re<c, P u> has gradient P^H c
    assert rel_l2(g, adj) < 1e-5
    lhs = np.vdot(c, out.u.detach().cpu().numpy())
    rhs = np.vdot(g, u)
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


@pytest.mark.parametrize("shape", [(96, 192), (384, 48), (1536, 24), (3072, 12)])
def test_three_times_power_of_two_grids(shape):
    """Lengths 3 * 2^a (radix-3 first stage) in rows and columns, ASM, accurate."""
    rng = np.random.default_rng(sum(shape))
    u = np.exp(2j * np.pi * rng.random(shape)).astype(np.complex64)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), (0.3, 0.25), 0.0, 0.532)
    out = cf.asm_propagate(f, 30.0, 1.33, 0, phase_mode="accurate").u.cpu().numpy()
    ref = op.propagate(u, (0.3, 0.25), 0.532, 30.0, 1.33, 0)
    assert rel_l2(out, ref) < 1e-5
    out_t = cf.asm_propagate(cx.ScalarField(torch.as_tensor(u.T.copy()).cuda(), (0.25, 0.3), 0.0, 0.532),
                             30.0, 1.33, 0, phase_mode="accurate").u.cpu().numpy()
    assert rel_l2(out_t, ref.T) < 1e-5


@pytest.mark.parametrize("phase_mode", ["ref32", "accurate"])
@pytest.mark.parametrize("method,shape,pad,batch,opts", [
    ("asm", (960, 200), (32, 28), 5, dict(bandlimit=True, shift_yx=(20.0, 10.0))),  # 1024 x 256
    ("transfer", (512, 80), 0, 4, dict(kykx=(0.2, 0.1))),                          # 512 x 80
    ("asm", (2048, 40), 0, 6, {}),                                                   # 2048 x 40
    ("asm", (1024, 1024), 0, 7, {}),                                                 # tiled planes
    ("asm", (1024, 32), 0, 5, dict(dx=0.1, z=2.0)),                                  # evanescent
])
def test_persistent_column_pass_is_bit_identical(method, shape, pad, batch, opts, phase_mode):
    """A call with one parameter set runs the persistent column pass (H of a
    strip evaluated once per workgroup, the chunk's strip x plane items dealt
    in equal ranges); its bits equal those of the per-plane pass evaluating
    H on the fly, in every launch shape of the 1024-point pass."""
    opts = dict(opts)
    dx, z = opts.pop("dx", 0.3), opts.pop("z", 40.0)
    u = _phase_field(shape, seed=13, batch=(batch,))
    lib = L.lib()
    runs = [{}, {L.CXM_OPT_COL_PERSIST: 1, L.CXM_OPT_H_TABLE: 1}, {L.CXM_OPT_COL_PERSIST: 3}]
    if shape[0] + 2 * (pad if np.isscalar(pad) else pad[0]) == 1024:
        runs += [{L.CXM_OPT_COL_PERSIST_MODE: m} for m in (1, 2, 3)]
    keys = (L.CXM_OPT_COL_PERSIST, L.CXM_OPT_H_TABLE, L.CXM_OPT_COL_PERSIST_MODE)
    outs = []
    try:
        for r in runs:
            for k in keys:
                L.check(lib.cxm_set_option(k, r.get(k, 0)))
            outs.append(_run(u, dx, z, pad, method, "same", phase_mode, **opts))
    finally:
        for k in keys:
            L.check(lib.cxm_set_option(k, 0))
    for o in outs[1:]:
        assert np.array_equal(outs[0].view(np.uint32), o.view(np.uint32))
    ref = _oracle(u[batch - 1], dx, z, pad, method, "same",
                  "ref32" if phase_mode == "ref32" else "fp64", **opts)
    assert rel_l2(outs[0][batch - 1], ref) < TOL


# ---- grids of any length: chirp-z passes ------------------------------------------
@pytest.mark.parametrize("phase_mode", ["ref32", "accurate"])
@pytest.mark.parametrize("method,shape,pad,mode,z", [
    ("asm", (100, 100), 0, "same", 10.0),
    ("transfer", (100, 100), 0, "full", -7.0),
    ("asm", (90, 70), (7, 13), "same", 25.0),     # 104 x 96: one axis fused-length, one not
    ("asm", (60, 44), (3, 5), "full", 12.0),
])
def test_arbitrary_grid_lengths(method, shape, pad, mode, z, phase_mode):
    """Grid lengths without a fused transform (100, 104, 66, 54, ...) run the
    2D DFT pair as chirp-z passes; same results as the oracle's exact FFTs."""
    u = _phase_field(shape, seed=21)
    got = _run(u, 0.3, z, pad, method, mode, phase_mode)
    ref = _oracle(u, 0.3, z, pad, method, mode, "ref32" if phase_mode == "ref32" else "fp64")
    assert got.shape == ref.shape
    err = rel_l2(got, ref)
    print(f"{method} {shape} pad {pad} {mode}: {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("px,z,calc,method", [
    (512, 100.0, "exact", "asm"),        # 4604-point grid
    (512, 100.0, "transfer", "transfer"),  # 1856
    (256, 100.0, "exact", "asm"),        # 4860
    (1024, 50.0, "exact", "asm"),        # 7836
])
def test_reference_padding_calculator_grids(px, z, calc, method):
    """The grids the reference's own padding calculators produce
    (compute_padding_exact / compute_padding_transfer, propagation.py:583-656)
    for a 0.3 um, 532 nm field, propagated in "same" mode."""
    fn = cf.compute_padding_exact if calc == "exact" else cf.compute_padding_transfer
    pad = fn(px, LAM, 0.3, z)
    print(f"{px} px z {z}: pad {pad} -> grid {px + 2 * pad}", flush=True)
    assert not L.lib().cxm_fft_supported(px + 2 * pad)
    u = _phase_field((px, px), seed=22)
    got = _run(u, 0.3, z, pad, method, "same", "accurate")
    ref = op.propagate(u, (0.3, 0.3), LAM, z, NM, pad, method=method, mode="same",
                       precision="fp64", workers=16)
    err = rel_l2(got, ref)
    print(f"  rel_l2 vs FP64 = {err:.2e}")
    assert err < TOL


def test_arbitrary_grid_batches_distances_and_wavelengths():
    zs = np.array([-20.0, 5.0, 12.5])
    u = _phase_field((100, 60), seed=23, batch=(3,))
    out = cf.asm_propagate(_mono(u, 0.3), zs, NM, (9, 3), mode="same")  # zip: plane b at zs[b]
    for b in range(3):
        rb = _oracle(u[b], 0.3, zs[b], (9, 3), "asm", "same", "ref32")
        assert rel_l2(out.u[b].cpu().numpy(), rb) < TOL
    ub = _phase_field((100, 60), seed=24)
    o2 = cf.asm_propagate(_mono(ub, 0.3), zs, NM, (9, 3), mode="same")           # outer
    assert o2.shape == (3, 100, 60)
    for k, zk in enumerate(zs):
        assert rel_l2(o2.u[k].cpu().numpy(), _oracle(ub, 0.3, zk, (9, 3), "asm", "same", "ref32")) < TOL
    wl = np.array([0.45, 0.65])
    uc = np.exp(2j * np.pi * np.random.default_rng(25).random((70, 50, 2))).astype(np.complex64)
    fc = cx.ChromaticScalarField(torch.as_tensor(uc).cuda(), 0.3, 0.0, cx.Spectrum(wl))
    oc = cf.transfer_propagate(fc, 15.0, NM, 5, mode="same").u.cpu().numpy()
    for w in range(2):
        r = op.propagate(uc[..., w], (0.3, 0.3), wl[w], 15.0, NM, 5, method="transfer",
                         precision="ref32", fft_precision="fp64")
        assert rel_l2(oc[..., w], r) < TOL
    # intensity output of the same grid
    I = cf.propagate_intensity(_mono(ub, 0.3), 12.5, NM, (9, 3))
    ref = np.abs(_oracle(ub, 0.3, 12.5, (9, 3), "asm", "same", "ref32")) ** 2
    assert rel_l2(I.cpu().numpy().reshape(ref.shape), ref) < TOL



def test_arbitrary_grid_adjoint_identity():
    """<P u, c> = <u, P^H c> on a 100 x 90 grid (chirp-z passes both ways)."""
    rng = np.random.default_rng(26)
    u = torch.as_tensor(_phase_field((80, 70), seed=27)).cuda().requires_grad_(True)
    c = torch.as_tensor((rng.standard_normal((80, 70)) + 1j * rng.standard_normal((80, 70)))
                        .astype(np.complex64)).cuda()
    f = cx.ScalarField(u, 0.3, 0.0, LAM)
    out = cf.asm_propagate(f, 30.0, NM, (10, 10), mode="same", remove_evanescent=True)
    lhs = torch.vdot(out.u.reshape(-1), c.reshape(-1))
    (g,) = torch.autograd.grad((out.u * c.conj()).real.sum(), u)
    rhs = torch.vdot(u.detach().reshape(-1), g.reshape(-1))
    assert abs((lhs - rhs).item()) / abs(lhs.item()) < 1e-5


def test_lazily_conjugated_inputs_are_read_conjugated():
    """torch's conj() is a view with a conjugate bit over the same memory;
    the library reads device pointers, so the wrappers resolve it first."""
    from chromatix_amd.ops import fourier_convolution

    u = _phase_field((64, 48), seed=31)
    ut = torch.as_tensor(u).cuda()
    got = cf.asm_propagate(cx.ScalarField(ut.conj(), 0.3, 0.0, LAM), 20.0, NM, 8, mode="same")
    ref = _oracle(np.conj(u), 0.3, 20.0, 8, "asm", "same", "ref32")
    assert rel_l2(got.u.cpu().numpy(), ref) < TOL
    k = torch.as_tensor(_phase_field((5, 7), seed=32)).cuda()
    a = fourier_convolution(ut.conj(), k.conj(), axes=(0, 1))
    b = fourier_convolution(torch.conj_physical(ut), torch.conj_physical(k), axes=(0, 1))
    assert torch.equal(a, b)
