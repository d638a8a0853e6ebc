"""GPU parity of the sensor's resampling (cxm_resample_f32 via
ops.resample.init_plane_resample), BasicSensor with resampling and the
Microscope's PSF taper, against the FP64 oracle (oracle/resample.py,
pinned against Pillow).  Bar: relative L2 <= 1e-5."""
import numpy as np
import pytest
import torch

import chromatix_amd.functional as cf
from chromatix_amd.elements import BasicSensor
from chromatix_amd.ops.resample import init_plane_resample
from chromatix_amd.systems import Microscope
from oracle import convolution as oc
from oracle import resample as ors
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("method", ["linear", "cubic", "lanczos3", "lanczos5"])
@pytest.mark.parametrize("in_hw,in_sp,out_hw,out_sp", [
    ((96, 80), 0.3, (48, 40), 0.6),       # same field of view, downsampled by 2
    ((100, 90), 0.25, (31, 27), 0.7),     # odd factors, fields of view differ
    ((40, 36), 0.6, (96, 90), 0.25),      # upsampling
    ((64, 64), (0.3, 0.2), (40, 50), 0.45)])  # non-square input pixels
def test_interpolating_resampler(method, in_hw, in_sp, out_hw, out_sp):
    rng = np.random.default_rng(0)
    x = rng.random((2, 3) + in_hw).astype(np.float32)
    r = init_plane_resample(out_hw, out_sp, method)
    got = r(torch.as_tensor(x).cuda(), in_sp).cpu().numpy()
    ref = ors.interpolating_resample(x.astype(np.float64), out_hw, out_sp, in_sp, method)
    assert got.shape == (2, 3) + out_hw
    err = rel_l2(got, ref)
    print(f"resample {method} {in_hw}->{out_hw}: {err:.2e}")
    assert err < TOL


def test_pooling_resampler():
    x = np.random.default_rng(1).random((4, 60, 90)).astype(np.float32)
    got = init_plane_resample((20, 30), 0.9, "pool")(torch.as_tensor(x).cuda(), 0.3).cpu().numpy()
    assert rel_l2(got, ors.pooling_resample(x.astype(np.float64), (20, 30))) < 1e-6
    with pytest.raises(ValueError):
        init_plane_resample((7, 30), 0.9, "pool")(torch.as_tensor(x).cuda(), 0.3)


@pytest.mark.parametrize("method", ["linear", "cubic", "pool"])
def test_resampler_gradient_is_the_adjoint(method):
    # <R x, y> = <x, R^T y>: autograd's rule against the FP64 weight matrices
    rng = np.random.default_rng(2)
    x = torch.as_tensor(rng.random((2, 60, 48)).astype(np.float32)).cuda().requires_grad_(True)
    out_hw = (20, 16)
    r = init_plane_resample(out_hw, 0.9, method)
    y = torch.as_tensor(rng.standard_normal((2,) + out_hw).astype(np.float32)).cuda()
    (r(x, 0.3) * y).sum().backward()
    if method == "pool":
        wy = np.kron(np.eye(20), np.ones((3, 1)))
        wx = np.kron(np.eye(16), np.ones((3, 1)))
    else:
        wy = ors.weight_matrix(60, 20, 1 / 3, 0.0, method) * 3
        wx = ors.weight_matrix(48, 16, 1 / 3, 0.0, method) * 3
    ref = np.einsum("yi,xj,bij->byx", wy, wx, y.cpu().numpy().astype(np.float64))
    assert rel_l2(x.grad.cpu().numpy(), ref) < TOL


def test_basic_sensor_resamples_a_field():
    # the reference's test_sensors.py case: a 512^2 3-plane point source seen
    # by a 256^2 sensor of 0.6 um pixels (cubic), and by pooling summed over z
    field = cf.objective_point_source((512, 512), 0.3, 0.532, np.linspace(-5, 5, 3), f=100.0, n=1.0,
                                      NA=0.8)
    I = field.intensity.cpu().numpy().astype(np.float64).reshape(3, 512, 512)
    s1 = BasicSensor((256, 256), 0.6, resampling_method="cubic")
    img = s1(field).cpu().numpy().reshape(3, 256, 256)
    ref = ors.interpolating_resample(I, (256, 256), 0.6, 0.3, "cubic")
    assert rel_l2(img, ref) < TOL
    s2 = BasicSensor((256, 256), 0.6, resampling_method="pool", reduce_axis=0)
    img2 = s2(field).cpu().numpy().squeeze()
    assert img2.shape == (256, 256)
    assert rel_l2(img2, ors.pooling_resample(I, (256, 256)).sum(0)) < 1e-6
    # the same from the intensity array with its spacing
    img3 = s2(field.intensity, input_spacing=0.3).cpu().numpy().squeeze()
    assert np.array_equal(img2, img3)


def test_microscope_tapers_and_resamples_the_psf():
    # padded PSF -> crop -> sigmoid taper -> resampling to the sensor's pixels
    # -> planewise convolution summed over z (microscopes.py:229-297)
    rng = np.random.default_rng(3)
    D, ps, n_s, pitch = 3, 96, 48, 0.5
    padded = rng.random((D, 120, 120)).astype(np.float32)  # padding_ratio 0.25: 96 -> 120
    volume = rng.random((D, n_s, n_s)).astype(np.float32)
    sensor = BasicSensor((n_s, n_s), pitch, resampling_method="linear", reduce_axis=0)

    def system_psf(mic):
        return torch.as_tensor(padded).cuda()

    m = Microscope(system_psf, sensor, f=5.0, n=1.33, NA=0.8, spectrum=0.532, padding_ratio=0.25,
                   taper_width=3.0)
    img = m(torch.as_tensor(volume).cuda()).cpu().numpy()
    psf = padded[:, 12:108, 12:108].astype(np.float64) * ors.sigmoid_taper((ps, ps), 3.0)
    # an intensity PSF's spacing: the sensor pitch times sensor pixels / PSF pixels
    psf = ors.interpolating_resample(psf, (n_s, n_s), pitch, pitch * n_s / ps, "linear")
    ref = oc.microscope_image(volume.astype(np.float64), psf, reduce_axis=0)
    assert img.shape == (n_s, n_s)
    assert rel_l2(img, ref) < TOL


def test_taper_without_resampling_and_its_gradient():
    rng = np.random.default_rng(4)
    n = 64
    sample = torch.as_tensor(rng.random((n, n)).astype(np.float32)).cuda()
    base = torch.as_tensor(rng.random((n, n)).astype(np.float32)).cuda().requires_grad_(True)
    sensor = BasicSensor((n, n), 0.5, resampling_method=None)
    m = Microscope(lambda mic: base, sensor, f=5.0, n=1.33, NA=0.8, spectrum=0.532, taper_width=2.0)
    img = m(sample)
    t = ors.sigmoid_taper((n, n), 2.0)
    ref = oc.fourier_convolution(sample.cpu().numpy().astype(np.float64),
                                 base.detach().cpu().numpy().astype(np.float64) * t, axes=(-2, -1))
    assert rel_l2(img.detach().cpu().numpy(), ref) < TOL
    w = torch.as_tensor(rng.standard_normal((n, n)).astype(np.float32)).cuda()
    (img * w).sum().backward()
    # d/dbase <conv(s, base t), w> = t * corr(w, s)
    s64 = sample.cpu().numpy().astype(np.float64)
    eps = 1e-3
    for (i, j) in [(10, 20), (31, 33), (1, 62)]:
        e = np.zeros((n, n))
        e[i, j] = eps
        fd = np.sum(w.cpu().numpy() * (oc.fourier_convolution(s64, (base.detach().cpu().numpy() + e) * t,
                                                              axes=(-2, -1)) - ref)) / eps
        assert abs(base.grad[i, j].item() - fd) <= 1e-3 * max(1.0, abs(fd))
