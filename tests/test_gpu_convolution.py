"""GPU parity of fourier_convolution (cxm_fconv_c64), the sensor kernels
(cxm_sum_axis_f32, cxm_shot_noise_approx_f32) and Microscope.image against
the oracle (FP64).  Bar: relative L2 <= 1e-5; the reference tests' shapes
also meet their own criterion (allclose to scipy, atol 1e-3)."""
import numpy as np
import pytest
import scipy.signal
import torch

import chromatix_amd.functional as cf
from chromatix_amd.elements import BasicSensor
from chromatix_amd.ops import (approximate_shot_noise, fourier_convolution,
                               planewise_convolution_sum, shot_noise)
from chromatix_amd.systems import Microscope
from oracle import convolution as oc
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _arr(rng, shape, cplx=False):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
        return x.astype(np.complex64)
    return x.astype(np.float32)


def _gpu(a, k, **kw):
    return fourier_convolution(torch.as_tensor(a).cuda(), torch.as_tensor(k).cuda(), **kw).cpu().numpy()


@pytest.mark.parametrize("shapes,fast", [
    (((512, 512), (256, 256)), True),     # tests/test_convolution.py:8-16
    (((513, 513), (253, 253)), False),    # tests/test_convolution.py:30-40
    (((37, 53), (9, 12)), True),
    (((64, 64), (64, 64)), True),
    (((30, 100), (101, 7)), False),
    (((1, 40), (1, 5)), True),
])
def test_same_real(shapes, fast):
    rng = np.random.default_rng(0)
    (H, W), (h, w) = shapes
    a, k = _arr(rng, (H, W)), _arr(rng, (h, w))
    print("fconv same real", shapes, fast, flush=True)
    out = _gpu(a, k, fast_fft_shape=fast)
    ref = oc.fourier_convolution(a, k, fast_fft_shape=fast)
    assert out.dtype == np.float32 and out.shape == (H, W)
    assert rel_l2(out, ref) < TOL
    if H >= 512:
        sp = scipy.signal.convolve(a.astype(np.float64), k.astype(np.float64), method="fft", mode="same")
        assert np.allclose(sp, out, atol=1e-3)


@pytest.mark.parametrize("ci,ck", [(True, False), (False, True), (True, True)])
def test_same_complex(ci, ck):
    rng = np.random.default_rng(1)
    a, k = _arr(rng, (3, 70, 45), ci), _arr(rng, (3, 11, 8), ck)
    out = _gpu(a, k, axes=(-2, -1))
    ref = oc.fourier_convolution(a, k, axes=(-2, -1))
    assert out.dtype == np.complex64
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("ishape,kshape", [
    ((3, 40, 50), (1, 9, 9)),        # one kernel for every plane
    ((3, 40, 50), (3, 9, 9)),        # one kernel per plane
    ((2, 3, 40, 50), (1, 3, 7, 5)),  # kernel b % 3
    ((2, 3, 40, 50), (2, 1, 7, 5)),  # explicit broadcast
    ((1, 40, 50), (4, 9, 9)),        # same mode keeps kernel 0 (the reference's crop)
])
def test_batch_broadcast(ishape, kshape):
    rng = np.random.default_rng(2)
    a, k = _arr(rng, ishape), _arr(rng, kshape)
    out = _gpu(a, k, axes=(-2, -1))
    ref = oc.fourier_convolution(a, k, axes=(-2, -1))
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("cplx", [False, True])
def test_full_mode(fast, cplx):
    rng = np.random.default_rng(3)
    a, k = _arr(rng, (2, 33, 60), cplx), _arr(rng, (1, 16, 5), cplx)
    out = _gpu(a, k, axes=(1, 2), mode="full", fast_fft_shape=fast)
    ref = oc.fourier_convolution(a, k, axes=(1, 2), mode="full", fast_fft_shape=fast)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL
    # past the linear convolution: zeros (to rounding)
    assert np.abs(out[:, 33 + 16 - 1:, :]).max() < 1e-5 * np.abs(out).max()


def test_default_axes_channels_last():
    rng = np.random.default_rng(4)
    a, k = _arr(rng, (40, 50, 3)), _arr(rng, (9, 7, 3))
    out = _gpu(a, k)
    ref = oc.fourier_convolution(a, k)
    assert out.shape == (40, 50, 3)
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("ishape,kshape,mode,cplx", [
    ((5, 512, 512), (3, 256, 256), "same", False),  # tests/test_convolution.py:19-27
    ((5, 40, 50), (3, 7, 9), "same", True),
    ((4, 30, 20), (6, 5, 5), "full", False),
    ((2, 5, 32, 24), (1, 3, 8, 8), "same", False),
    ((2, 5, 32, 24), (2, 3, 8, 8), "same", True),
])
def test_three_axes(ishape, kshape, mode, cplx):
    rng = np.random.default_rng(5)
    a, k = _arr(rng, ishape, cplx), _arr(rng, kshape, cplx)
    ax = (-3, -2, -1)
    print("fconv 3d", ishape, kshape, mode, flush=True)
    out = _gpu(a, k, axes=ax, mode=mode)
    ref = oc.fourier_convolution(a, k, axes=ax, mode=mode)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL
    if ishape == (5, 512, 512):
        sp = scipy.signal.convolve(a.astype(np.float64), k.astype(np.float64), method="fft", mode="same")
        assert np.allclose(sp, out, atol=1e-3)


def test_empty_batch_and_errors():
    z = fourier_convolution(torch.zeros(0, 8, 8, device="cuda"), torch.ones(1, 3, 3, device="cuda"),
                            axes=(-2, -1))
    assert tuple(z.shape) == (0, 8, 8)
    x = torch.ones(8, 8, device="cuda")
    with pytest.raises(ValueError):
        fourier_convolution(x, torch.ones(1, 3, 3, device="cuda"))
    with pytest.raises(ValueError):
        fourier_convolution(torch.ones(2, 8, 8, device="cuda"), torch.ones(2, 3, 3, device="cuda"),
                            axes=(0, 2))
    with pytest.raises(ValueError):
        fourier_convolution(x, x, mode="valid")
    with pytest.raises(RuntimeError):
        fourier_convolution(torch.ones(8, 8), torch.ones(3, 3))


def test_sum_axis_and_approximate_noise():
    rng = np.random.default_rng(6)
    x = (rng.random((4, 7, 33, 20)) * 50).astype(np.float32)
    xt = torch.as_tensor(x).cuda()
    for ax in (0, 1, -1):
        s = cf.basic_sensor(xt, reduce_axis=ax).cpu().numpy()
        assert rel_l2(s, x.astype(np.float64).sum(axis=ax)) < 1e-6
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    out = approximate_shot_noise(g, xt).cpu().numpy()
    g.manual_seed(11)
    nrm = torch.randn(xt.shape, generator=g, device="cuda").cpu().numpy()
    ref = oc.approximate_shot_noise(x, nrm)
    assert rel_l2(out, ref) < 1e-6
    assert (out >= 0).all()
    # same seed, same samples through basic_sensor
    s = cf.basic_sensor(xt, "approximate", reduce_axis=1, noise_key=5).cpu().numpy()
    g.manual_seed(5)
    nrm = torch.randn((4, 33, 20), generator=g, device="cuda").cpu().numpy()
    ref = oc.basic_sensor(x, "approximate", reduce_axis=1, normal=nrm)
    assert rel_l2(s, ref) < 1e-6


def test_poisson_noise_statistics():
    lam = torch.full((512, 512), 100.0, device="cuda")
    y = shot_noise(3, lam)
    assert y.dtype == torch.float32
    m, v = float(y.mean()), float(y.var())
    # 2^18 samples: standard errors of mean and variance ~0.02 and ~0.3
    assert abs(m - 100.0) < 0.2 and abs(v - 100.0) < 3.0
    assert torch.equal(y, y.round())


def test_microscope_image_planewise():
    rng = np.random.default_rng(7)
    D, H, W = 4, 96, 80
    sample = rng.random((D, H, W)).astype(np.float32)
    psf = rng.random((D, 31, 31)).astype(np.float32)
    sensor = BasicSensor((H, W), 0.5, reduce_axis=0)
    m = Microscope(lambda mic: psf, sensor, f=5.0, n=1.33, NA=0.8, spectrum=0.532)
    img = m.image(torch.as_tensor(sample).cuda(), torch.as_tensor(psf).cuda()).cpu().numpy()
    ref = oc.microscope_image(sample, psf, reduce_axis=0)
    assert img.shape == (H, W)
    assert rel_l2(img, ref) < TOL


@pytest.mark.parametrize("shapes", [((5, 96, 80), (5, 17, 33)), ((3, 7, 4), (1, 3, 3)),
                                    ((24, 1000, 1000), (24, 41, 41))])
def test_planewise_convolution_sum(shapes):
    # the plane sum taken inside the convolution (one inverse row transform
    # for all planes) against the oracle's planewise convolutions summed in
    # FP64; 24 planes of 1280 x 1280 transforms run as two chunks (12 + 12
    # planes), the second added to the first's sum
    rng = np.random.default_rng(12)
    a = rng.random(shapes[0]).astype(np.float32)
    k = rng.random(shapes[1]).astype(np.float32)
    out = planewise_convolution_sum(torch.as_tensor(a).cuda(), torch.as_tensor(k).cuda())
    ref = oc.fourier_convolution(a, k, axes=(-2, -1)).sum(0)
    assert out is not None and tuple(out.shape) == shapes[0][1:]
    assert rel_l2(out.cpu().numpy(), ref) < TOL
    # the same sum through the unfused convolution and the sensor's sum
    two_step = fourier_convolution(torch.as_tensor(a).cuda(), torch.as_tensor(k).cuda(),
                                   axes=(-2, -1)).sum(0)
    assert rel_l2(out.cpu().numpy(), two_step.cpu().numpy()) < TOL


def test_planewise_convolution_sum_declines():
    # complex values and odd row transform lengths are left to the unfused path
    c = torch.ones((2, 8, 8), dtype=torch.complex64, device="cuda")
    assert planewise_convolution_sum(c, c) is None
    odd = torch.ones((2, 5, 3), device="cuda")  # 3 + 3 - 1 = 5 columns
    assert planewise_convolution_sum(odd, torch.ones((2, 2, 3), device="cuda")) is None


def test_microscope_call_crops_padding():
    rng = np.random.default_rng(8)
    H, W = 64, 64
    sample = rng.random((H, W)).astype(np.float32)
    padded = rng.random((80, 80)).astype(np.float32)  # padding_ratio 0.25: 64 -> 80

    def system_psf(mic):
        return torch.as_tensor(padded).cuda()

    sensor = BasicSensor((H, W), 0.5, resampling_method=None)
    m = Microscope(system_psf, sensor, f=5.0, n=1.33, NA=0.8, spectrum=0.532, padding_ratio=0.25)
    img = m(torch.as_tensor(sample).cuda()).cpu().numpy()
    ref = oc.fourier_convolution(sample, padded[8:72, 8:72], axes=(-2, -1))
    assert rel_l2(img, ref) < TOL


def test_microscope_with_4f_psf_end_to_end():
    # Microscope(Optical4FSystemPSF)(volume, z) against the oracle's 4f PSF
    # intensity convolved plane by plane and summed (microscopes.py:190-297)
    from chromatix_amd.systems import Optical4FSystemPSF
    from oracle import psf as op

    rng = np.random.default_rng(9)
    n, D, pitch, f_tube = 96, 4, 0.3, 100.0
    z = np.linspace(-4, 4, D)
    phase = rng.uniform(-np.pi, np.pi, (n, n)).astype(np.float32)
    volume = rng.random((D, n, n)).astype(np.float32)
    sensor = BasicSensor((n, n), pitch, resampling_method=None, reduce_axis=0)
    m = Microscope(Optical4FSystemPSF((n, n), pitch, f_tube, torch.as_tensor(phase).cuda()), sensor,
                   f=100.0, n=1.33, NA=0.8, spectrum=0.532)
    img = m(torch.as_tensor(volume).cuda(), z).cpu().numpy()
    dx = op.required_spacing(n, pitch, f_tube, 1.33, 0.532)
    psf = op.psf4f(phase.astype(np.float64), (dx, dx), 0.532, z, 100.0, 1.33, 0.8, f_tube)
    ref = oc.microscope_image(volume, psf, reduce_axis=0)
    assert img.shape == (n, n)
    assert rel_l2(img, ref) < TOL


@pytest.mark.parametrize("mode", ["same", "full"])
@pytest.mark.parametrize("cplx", [False, True])
def test_one_axis(mode, cplx):
    rng = np.random.default_rng(10)
    a, k = _arr(rng, (3, 200), cplx), _arr(rng, (3, 31), cplx)
    out = _gpu(a, k, axes=(1,), mode=mode)
    ref = oc.fourier_convolution(a, k, axes=(1,), mode=mode)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL
    # the default axes (0, 1) of a 2D array against axes (-1,) of a 3D one
    b = _arr(rng, (2, 5, 64), cplx)
    kk = _arr(rng, (1, 1, 9), cplx)
    out = _gpu(b, kk, axes=(-1,), mode=mode)
    ref = oc.fourier_convolution(b, kk, axes=(-1,), mode=mode)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("scale_k", [1.0, 1e-7, 3e5])
@pytest.mark.parametrize("shapes", [((3, 96, 80), (3, 17, 33)), ((2, 64, 64), (1, 64, 64)),
                                    ((1, 5, 3), (1, 2, 3)), ((2, 7, 4), (1, 3, 3)),
                                    ((1, 2600, 4), (1, 2521, 3)), ((2, 700, 40), (2, 61, 9)),
                                    ((1, 40, 700), (1, 9, 61))])
def test_paired_real_path_matches_separate_spectra(scale_k, shapes):
    # real image and kernel transformed as one complex plane (the kernel
    # scaled by a power of two to the image's norm) against the path with
    # separate kernel spectra (CXM_OPT_FCONV_PAIR 1), and both against FP64,
    # with kernels many orders of magnitude smaller or larger than the image.
    # Transform lengths (rows x columns): 128 x 128, 128 x 128, 6 x 5 (odd row
    # length: separate spectra only), 10 x 6 (a half-filled last strip of
    # column pairs, row counts not multiples of 4), 5120 x 6 (two columns per
    # column-pass workgroup), 768 x 48 (a 3 * 2^a column length, the column
    # pass at half the propagation's values per thread), 48 x 768 (the row
    # passes at half the values per thread)
    from chromatix_amd import _lib as L

    rng = np.random.default_rng(11)
    a = _arr(rng, shapes[0])
    k = (_arr(rng, shapes[1]) * scale_k).astype(np.float32)
    ref = oc.fourier_convolution(a, k, axes=(-2, -1))
    outs = []
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FCONV_PAIR, opt))
            outs.append(_gpu(a, k, axes=(-2, -1)))
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_FCONV_PAIR, 0))
    for o in outs:
        assert rel_l2(o, ref) < TOL


def _band_limited_psf(n, radius):
    """|ifft2(circular pupil)|^2, centred: a PSF whose spectrum (the pupil's
    autocorrelation) vanishes beyond 2 * radius bins."""
    f = np.fft.fftfreq(n) * n
    pupil = (f[:, None] ** 2 + f[None, :] ** 2) <= radius**2
    return np.fft.fftshift(np.abs(np.fft.ifft2(pupil)) ** 2)


@pytest.mark.parametrize("scale_k", [1.0, 3e5])
def test_paired_real_path_zero_and_sparse_planes(scale_k):
    # planes the shared transform of the paired path handles worst: an
    # all-zero image plane, an all-zero kernel plane (the pair's product is
    # exactly 0), and a sparse image (five point emitters) under a
    # band-limited PSF; each plane against FP64 on its own, both paths
    from chromatix_amd import _lib as L

    rng = np.random.default_rng(13)
    a = np.zeros((4, 96, 80), np.float32)
    a[0] = rng.random((96, 80))
    a[2] = rng.random((96, 80))
    for y, x in rng.integers(5, 75, size=(5, 2)):
        a[3, y, x] = rng.uniform(0.5, 2.0)
    k = np.zeros((4, 33, 33), np.float32)
    k[1] = rng.random((33, 33))
    k[2] = rng.random((33, 33))
    k[3] = _band_limited_psf(33, 6)
    k = (k * scale_k).astype(np.float32)
    a[1] = 0.0  # image plane 1 empty (kernel plane 1 is not)
    ref = oc.fourier_convolution(a, k, axes=(-2, -1))
    try:
        for opt in (0, 1):
            L.check(L.lib().cxm_set_option(L.CXM_OPT_FCONV_PAIR, opt))
            out = _gpu(a, k, axes=(-2, -1))
            assert np.all(out[0] == 0) and np.all(out[1] == 0), f"pair path {opt}: zero planes"
            for p in (2, 3):
                assert rel_l2(out[p], ref[p]) < TOL, (opt, p)
            if opt == 0:  # the summed convolution runs on the paired path only
                s = planewise_convolution_sum(torch.as_tensor(a).cuda(), torch.as_tensor(k).cuda())
                assert rel_l2(s.cpu().numpy(), ref.sum(0)) < TOL
    finally:
        L.check(L.lib().cxm_set_option(L.CXM_OPT_FCONV_PAIR, 0))
