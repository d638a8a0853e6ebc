"""GPU parity of the chirp-z transform (cxm_czt_axis_c64), the CZT and
scaled-convolution (Yu et al.) output resampling of kernel_propagate /
asm_propagate, and high_na_ff_lens against the FP64 oracle.  Bar: relative
L2 <= 1e-5.  (The reference builds its chirps in complex64 and is ~5e-3
from FP64 at these sizes; this build's chirps are FP64, so the gate is
FP64.)"""
import numpy as np
import pytest
import torch

import chromatix_amd as cx
import chromatix_amd.functional as cf
from chromatix_amd.utils import czt as uc
from oracle import czt as oc
from oracle import propagation as op
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _x(shape, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("n,m", [(100, 37), (256, 256), (2048, 512), (1280, 300)])
@pytest.mark.parametrize("axis", [0, 1])
def test_czt_axis(n, m, axis):
    shape = (2, n, 24) if axis == 0 else (2, 24, n)
    x = _x(shape, n + m)
    a = np.exp(-1j * 0.9)
    w = np.exp(1j * 2 * np.pi / (1.7 * m))
    got = uc.czt_axis(torch.as_tensor(x).cuda(), m, a, w, axis).cpu().numpy()
    ref = oc.czt_direct(x.astype(np.complex128), m, a, w, axis=1 + axis)
    err = rel_l2(got, ref)
    print(f"czt n={n} m={m} axis={axis}: {err:.2e}")
    assert err < TOL


def test_czt_generic_axis_and_parameter_sets():
    x = _x((5, 40, 3), 1)
    a, w = np.exp(-0.3j), np.exp(2j * np.pi / 61)
    got = uc.czt(torch.as_tensor(x).cuda(), 17, a, w, axis=1).cpu().numpy()
    assert rel_l2(got, oc.czt_direct(x, 17, a, w, axis=1)) < TOL
    # per-plane parameter sets (plane p uses set p % 2)
    xs = _x((4, 30, 30), 2)
    aa = np.array([np.exp(-0.2j), np.exp(-0.7j)])
    ww = np.array([np.exp(0.05j), np.exp(0.11j)])
    got = uc.czt_axis(torch.as_tensor(xs).cuda(), 20, aa, ww, 1, num_param=2).cpu().numpy()
    for p in range(4):
        ref = oc.czt_direct(xs[p], 20, aa[p % 2], ww[p % 2], axis=1)
        assert rel_l2(got[p], ref) < TOL


@pytest.mark.parametrize("path", ["mfma", "bluestein"])
def test_zoomed_fft(path, monkeypatch):
    monkeypatch.setenv("CXM_ZOOM_PATH", path)
    x = _x((64, 48), 3)
    got = uc.zoomed_fft(torch.as_tensor(x).cuda(), 0.4, 2.1, (40, 30)).cpu().numpy()
    assert rel_l2(got, oc.zoomed_fft(x, 0.4, 2.1, (40, 30))) < TOL


@pytest.mark.parametrize("path", ["default", "mfma", "bluestein"])
@pytest.mark.parametrize("n,m,odx,shift", [(256, 64, 0.2, (3.0, -1.5)), (2048, 512, 0.35, (0.0, 0.0)),
                                           (2048, 64, 0.35, (1.0, -2.0))])
def test_asm_propagate_czt_output_grid(n, m, odx, shift, path, monkeypatch):
    if path != "default":
        monkeypatch.setenv("CXM_ZOOM_PATH", path)
    rng = np.random.default_rng(4)
    u = np.exp(2j * np.pi * rng.random((n, n))).astype(np.complex64)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.5, 0.0, 0.532)
    out = cf.asm_propagate(f, 80.0, 1.33, 0, output_dx=odx, output_shape=(m, m), shift_yx=shift,
                           phase_mode="accurate")
    ref, dxo, org = oc.czt_propagate(u, (0.5, 0.5), 0.532, 80.0, 1.33, odx, (m, m), shift_yx=shift)
    assert out.u.shape == (m, m)
    err = rel_l2(out.u.cpu().numpy(), ref)
    print(f"czt propagate {n}->{m}: {err:.2e}")
    assert err < TOL
    np.testing.assert_allclose(out.dx, dxo)
    np.testing.assert_allclose(out.origin, org)


@pytest.mark.parametrize("per_wv_dx", [False, True])
def test_chromatic_czt_output_grid(per_wv_dx):
    """A chromatic field on the CZT output grid: each wavelength on its own
    input and output spacing equals the monochromatic oracle of that
    wavelength."""
    rng = np.random.default_rng(6)
    wl = np.array([0.45, 0.532, 0.65])
    dx = np.array([[0.5, 0.5], [0.45, 0.5], [0.4, 0.42]]) if per_wv_dx else 0.5
    u = np.exp(2j * np.pi * rng.random((256, 256, 3))).astype(np.complex64)
    f = cx.ChromaticScalarField(torch.as_tensor(u).cuda(), dx, 0.0, cx.Spectrum(wl))
    odx = np.array([[0.2, 0.2], [0.25, 0.22], [0.3, 0.3]]) if per_wv_dx else 0.2
    out = cf.asm_propagate(f, 80.0, 1.33, 0, output_dx=odx, output_shape=(64, 48),
                           shift_yx=(2.0, -1.0), phase_mode="accurate")
    assert out.u.shape == (64, 48, 3)
    for w in range(3):
        dxw = f.dx[w]
        odxw = np.broadcast_to(np.asarray(odx, np.float64), (3, 2))[w]
        ref, dxo, org = oc.czt_propagate(u[..., w], dxw, wl[w], 80.0, 1.33, odxw, (64, 48),
                                         shift_yx=(2.0, -1.0))
        err = rel_l2(out.u[..., w].cpu().numpy(), ref)
        print(f"chromatic czt wavelength {wl[w]}: {err:.2e}")
        assert err < TOL
        np.testing.assert_allclose(out.dx[w], dxo)
        np.testing.assert_allclose(out.origin[w], org)


@pytest.mark.parametrize("shape,odx,shift", [
    ((256, 256), 0.45, (0.0, 0.0)),
    ((512, 512), (0.2, 0.35), (3.0, -1.5)),
    ((320, 384), 0.6, (-2.0, 1.0)),
    ((1024, 1024), 0.3, (0.0, 0.0)),
])
def test_kernel_propagate_scaled_convolution(shape, odx, shift):
    """kernel_propagate(..., output_dx, use_czt=False): the scaled convolution
    of Yu et al. (propagation.py:444-458) against the FP64 oracle."""
    rng = np.random.default_rng(shape[0] + shape[1])
    u = np.exp(2j * np.pi * rng.random(shape)).astype(np.complex64)
    H = op.asm_propagator(shape, (0.3, 0.3), 0.532, 40.0, 1.33)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
    out = cf.kernel_propagate(f, torch.as_tensor(H.astype(np.complex64)).cuda(), output_dx=odx,
                              shift_yx=shift, use_czt=False)
    ref = oc.yu_kernel_propagate(u, H, (0.3, 0.3), odx, shift)
    assert out.u.shape == shape
    err = rel_l2(out.u.cpu().numpy(), ref)
    print(f"scaled convolution {shape} dx_out {odx}: {err:.2e}")
    assert err < TOL
    np.testing.assert_allclose(out.dx, np.broadcast_to(np.asarray(odx, float), (2,)))
    np.testing.assert_allclose(out.origin, np.asarray(shift, float))


def test_asm_propagate_scaled_convolution_same_mode():
    """asm_propagate with output_dx and use_czt=False: pad, propagator with
    the output shift, scaled convolution, crop (propagation.py:325-352)."""
    rng = np.random.default_rng(9)
    u = np.exp(2j * np.pi * rng.random((192, 192))).astype(np.complex64)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.25, 0.0, 0.532)
    out = cf.asm_propagate(f, 30.0, 1.33, 32, output_dx=0.4, shift_yx=(1.0, 2.0), use_czt=False,
                           mode="same", phase_mode="accurate")
    up = np.pad(u.astype(np.complex128), 32)
    H = op.asm_propagator(up.shape, (0.25, 0.25), 0.532, 30.0, 1.33, shift_yx=(1.0, 2.0))
    ref = oc.yu_kernel_propagate(up, H, (0.25, 0.25), 0.4, (1.0, 2.0))[32:-32, 32:-32]
    assert out.u.shape == (192, 192)
    assert rel_l2(out.u.cpu().numpy(), ref) < TOL
    with pytest.raises(ValueError):
        cf.asm_propagate(f, 30.0, 1.33, 32, output_dx=0.4, output_shape=(128, 128), use_czt=False,
                         mode="same")


@pytest.mark.parametrize("n,out_shape,odx,f_len,NA", [
    (256, None, None, 100.0, 0.8),
    # pupil edge inside the grid (181 bins masked); NA chosen so that no bin
    # lies within rounding of the edge, where the defocus 1/sz of the
    # reference's formula is unbounded (NA 1.2 puts bin (256, 192) on it)
    (512, (200, 200), 0.15, 2000.0, 1.08),
])
@pytest.mark.parametrize("path", ["mfma", "bluestein"])
def test_high_na_ff_lens(n, out_shape, odx, f_len, NA, path, monkeypatch):
    monkeypatch.setenv("CXM_ZOOM_PATH", path)
    rng = np.random.default_rng(n)
    u = np.exp(2j * np.pi * rng.random((n, n))).astype(np.complex64)
    f = cx.ScalarField(torch.as_tensor(u).cuda(), 0.3, 0.0, 0.532)
    out = cf.high_na_ff_lens(f, f_len, 1.33, NA, output_shape=out_shape, output_dx=odx)
    ref, dxo = oc.high_na_ff_lens(u, (0.3, 0.3), 0.532, f_len, 1.33, NA, out_shape, odx)
    err = rel_l2(out.u.cpu().numpy(), ref)
    print(f"high-NA lens {n} -> {ref.shape}: {err:.2e}")
    assert out.u.shape == ref.shape
    assert err < TOL
    np.testing.assert_allclose(out.dx, dxo)


@pytest.mark.parametrize("n,m", [(1024, 512), (700, 300), (1200, 900), (2048, 512), (3000, 40), (600, 500)])
@pytest.mark.parametrize("axis", [0, 1])
def test_dedicated_chirp_passes_match_pipeline(n, m, axis):
    """The dedicated chirp-z row and column passes (CXM_OPT_CZT_PASSES 0;
    FFT lengths 1024 ... 5120, the chirp multiplies in registers) give the
    operator pipeline's bits, with two chirp parameter sets; and the Yu
    scaled-convolution pass likewise."""
    from chromatix_amd import _lib as L

    shape = (3, n, 20) if axis == 0 else (3, 20, n)
    x = torch.as_tensor(_x(shape, n + m)).cuda()
    aa = np.array([np.exp(-0.9j), np.exp(-0.4j)])
    ww = np.array([np.exp(2j * np.pi / (1.7 * m)), np.exp(2j * np.pi / (2.3 * m))])
    outs, yus = [], []
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_CZT_PASSES, opt))
            outs.append(uc.czt_axis(x, m, aa, ww, axis, num_param=2, mod_n=n // 2, scale=0.5))
            yus.append(uc._yu_raw(x, axis, 1.5, 0.01, 0.3, 0.2))
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_CZT_PASSES, 0))
    print(f"dedicated chirp passes n={n} m={m} axis={axis}: czt {bool(torch.equal(outs[0], outs[1]))}, "
          f"yu {bool(torch.equal(yus[0], yus[1]))}")
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(yus[0], yus[1])
    ref = oc.czt_direct(x.cpu().numpy()[1].astype(np.complex128), m, aa[1], ww[1], axis=axis)
    ref = ref * ww[1] ** (-(n // 2) * np.arange(m)).reshape((-1, 1) if axis == 0 else (1, -1)) * aa[1] ** (n // 2) * 0.5
    assert rel_l2(outs[0].cpu().numpy()[1], ref) < TOL


@pytest.mark.parametrize("shape,use_czt", [((1024, 1024), False), ((96, 160), False), ((96, 160), True)])
def test_fused_propagator_multiply_matches_separate(shape, use_czt):
    """kernel_propagate's output-grid branches without gradients multiply the
    shifted spectrum by fftshift(H) on the transform's column store; with
    gradients (the autograd path) in a separate pass.  Same bits."""
    ny, nx = shape
    rng = np.random.default_rng(3)
    u = torch.polar(torch.ones(2, ny, nx), torch.as_tensor(rng.uniform(0, 2 * np.pi, (2, ny, nx)),
                                                            dtype=torch.float32)).cuda()
    f = cx.ScalarField(u, 0.3, 0.0, 0.532)
    H = cf.compute_asm_propagator(f, 50.0, 1.33)
    kw = dict(output_dx=0.45, use_czt=use_czt)
    if use_czt:
        kw["output_shape"] = (64, 48)
    fused = cf.kernel_propagate(f, H, **kw).u
    ug = u.clone().requires_grad_(True)
    sep = cf.kernel_propagate(cx.ScalarField(ug, 0.3, 0.0, 0.532), H, **kw).u.detach()
    print(f"fused vs separate propagator multiply {shape} czt={use_czt}: {bool(torch.equal(fused, sep))}")
    assert torch.equal(fused, sep)
