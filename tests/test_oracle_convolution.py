"""The convolution / sensor oracle pinned on the CPU: the defining double
sum (brute force) on small cases, scipy.signal.convolve (the reference's
own test oracle, tests/test_convolution.py) on the reference tests' shapes,
closed forms of the shot-noise clip; and the cross-process sum of the
product's basic_sensor over gloo (two processes, host tensors)."""
import os
import socket

import numpy as np
import pytest
import scipy.signal
import torch

from oracle import convolution as oc
from tests.conftest import rel_l2


def _rng_pair(rng, shape, cplx):
    x = rng.standard_normal(shape)
    return x + 1j * rng.standard_normal(shape) if cplx else x


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("shapes", [((9, 11), (3, 4)), ((8, 6), (5, 5)), ((5, 7), (9, 2)), ((1, 6), (1, 1))])
def test_oracle_matches_direct_sum(cplx, fast, shapes):
    rng = np.random.default_rng(1)
    (H, W), (h, w) = shapes
    a = _rng_pair(rng, (2, H, W), cplx)
    k = _rng_pair(rng, (2, h, w), cplx)
    full = oc.direct_convolution(a, k)
    same = oc.fourier_convolution(a, k, axes=(-2, -1), fast_fft_shape=fast)
    ref = full[:, (h - 1) // 2: (h - 1) // 2 + H, (w - 1) // 2: (w - 1) // 2 + W]
    assert same.shape == (2, H, W)
    assert rel_l2(same, ref if cplx else ref.real) < 1e-12
    out = oc.fourier_convolution(a, k, axes=(-2, -1), fast_fft_shape=fast, mode="full")
    sh = tuple(oc.next_order(n + m) if fast else n + m for n, m in ((H, h), (W, w)))
    assert out.shape == (2,) + sh
    pad = np.zeros((2,) + sh, np.complex128)
    pad[:, : H + h - 1, : W + w - 1] = full
    assert np.abs(out - (pad if cplx else pad.real)).max() < 1e-12 * np.abs(full).max()


def test_same_mode_crops_broadcast_dims_to_the_image():
    # the reference slices every non-convolved axis to the image's length:
    # a (1, ...) image against (3, ...) kernels keeps kernel 0 only
    rng = np.random.default_rng(2)
    a = rng.standard_normal((1, 6, 7))
    k = rng.standard_normal((3, 3, 3))
    out = oc.fourier_convolution(a, k, axes=(1, 2))
    assert out.shape == (1, 6, 7)
    ref = oc.fourier_convolution(a, k[:1], axes=(1, 2))
    assert np.allclose(out, ref, atol=1e-13)


def test_default_axes_are_the_first_two():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((10, 12, 2))
    k = rng.standard_normal((3, 5, 2))
    out = oc.fourier_convolution(a, k)
    for c in range(2):
        ref = scipy.signal.convolve(a[..., c], k[..., c], method="direct", mode="same")
        assert rel_l2(out[..., c], ref) < 1e-12


def test_reference_shapes_against_scipy():
    # tests/test_convolution.py:8-41 (the reference's own criterion: atol 1e-3)
    rng = np.random.default_rng(4)
    A = rng.normal(size=(512, 512))
    B = rng.normal(size=(256, 256))
    ref = scipy.signal.convolve(A, B, method="fft", mode="same")
    assert np.allclose(ref, oc.fourier_convolution(A, B), atol=1e-3)
    A = rng.normal(size=(513, 513))
    B = rng.normal(size=(253, 253))
    ref = scipy.signal.convolve(A, B, method="fft", mode="same")
    assert np.allclose(ref, oc.fourier_convolution(A, B, fast_fft_shape=False), atol=1e-3)


@pytest.mark.parametrize("fast", [True, False])
def test_three_axes_against_scipy(fast):
    rng = np.random.default_rng(5)
    A = rng.normal(size=(5, 40, 36))
    B = rng.normal(size=(3, 9, 12))
    ref = scipy.signal.convolve(A, B, method="direct", mode="same")
    out = oc.fourier_convolution(A, B, axes=(0, 1, 2), fast_fft_shape=fast)
    assert rel_l2(out, ref) < 1e-12


def test_shot_noise_and_sum_closed_forms():
    img = np.array([[0.0, 4.0, 4.0, 9.0]])
    nrm = np.array([[2.0, -3.0, 1.0, -1.0]])
    # max(x + sqrt(x) n, 0): 0 stays 0, 4 - 6 clips to 0, 4 + 2, 9 - 3
    assert np.array_equal(oc.approximate_shot_noise(img, nrm), [[0.0, 0.0, 6.0, 6.0]])
    x = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    s = oc.basic_sensor(x, reduce_axis=1)
    assert np.array_equal(s, [[12, 15, 18, 21], [48, 51, 54, 57]])


def test_microscope_oracle_is_conv_then_sum():
    rng = np.random.default_rng(6)
    sample = rng.random((3, 12, 10))
    psf = rng.random((3, 5, 5))
    img = oc.microscope_image(sample, psf, reduce_axis=0)
    ref = sum(scipy.signal.convolve(sample[z], psf[z], method="direct", mode="same") for z in range(3))
    assert rel_l2(img, ref) < 1e-12


def _free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def _psum_worker(rank, world, port, q):
    import torch.distributed as dist

    from chromatix_amd.functional.sensors import basic_sensor

    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        x = torch.full((3, 4), float(rank + 1))
        out = basic_sensor(x, reduce_parallel_axis_name="planes")
        q.put((rank, out.numpy().tolist(), x.numpy().tolist()))
    finally:
        dist.destroy_process_group()


def test_basic_sensor_cross_process_sum_gloo():
    ctx = torch.multiprocessing.get_context("spawn")
    q = ctx.Queue()
    port = _free_port()
    procs = [ctx.Process(target=_psum_worker, args=(r, 2, port, q)) for r in range(2)]
    for p in procs:
        p.start()
    res = [q.get(timeout=120) for _ in range(2)]
    for p in procs:
        p.join(timeout=60)
        assert p.exitcode == 0
    for rank, out, x in res:
        assert np.array_equal(out, np.full((3, 4), 3.0))  # 1 + 2
        assert np.array_equal(x, np.full((3, 4), rank + 1.0))  # input untouched
