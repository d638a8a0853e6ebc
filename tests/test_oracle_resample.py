"""Pins of the resampling oracle (oracle/resample.py) on the CPU.

jax.image.scale_and_translate cannot run here; its restatement is checked
against Pillow's resize of float images, an independent implementation of
the same antialiased resampling (triangle, Keys cubic a = -0.5, Lanczos 3;
output sample j at input coordinate box0 + (j + 1/2) w_box / n_out, the
kernel widened by the downscale factor and normalised per sample), and
against what the mathematics fixes: the identity at scale 1, a linear ramp
reproduced by linear interpolation, sums preserved by pooling, and the
taper's distance transform against a brute-force distance."""
import numpy as np
import pytest

from oracle import resample as rs

PIL = pytest.importorskip("PIL.Image")
METHODS = {"linear": PIL.BILINEAR, "cubic": PIL.BICUBIC, "lanczos3": PIL.LANCZOS}


def _pil(img, out_hw, method, box=None):
    im = PIL.fromarray(img.astype(np.float32), mode="F")
    r = im.resize((out_hw[1], out_hw[0]), METHODS[method], box=box)
    return np.asarray(r, dtype=np.float64)


@pytest.mark.parametrize("method", ["linear", "cubic", "lanczos3"])
@pytest.mark.parametrize("in_hw,out_hw", [((64, 48), (32, 24)), ((60, 45), (20, 15)), ((40, 32), (96, 80)),
                                          ((75, 50), (31, 22))])
def test_scale_and_translate_matches_pillow(method, in_hw, out_hw):
    rng = np.random.default_rng(0)
    img = rng.random(in_hw)
    scale = (out_hw[0] / in_hw[0], out_hw[1] / in_hw[1])
    got = rs.scale_and_translate(img, out_hw, scale, (0.0, 0.0), method)
    ref = _pil(img, out_hw, method)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 2e-6  # Pillow rounds its float output to float32


@pytest.mark.parametrize("method", ["linear", "cubic", "lanczos3"])
def test_translation_matches_pillow_box(method):
    # a sub-window of the input: translation t and scale s map output sample
    # j to input coordinate (j + 1/2 - t) / s, Pillow's box (b0, b0 + n / s)
    rng = np.random.default_rng(1)
    img = rng.random((80, 96))
    out_hw, s = (30, 36), (0.5, 0.5)
    t = (-5.0, -7.5)  # box origin (10, 15) in input pixels
    got = rs.scale_and_translate(img, out_hw, s, t, method)
    box = (-t[1] / s[1], -t[0] / s[0], -t[1] / s[1] + out_hw[1] / s[1], -t[0] / s[0] + out_hw[0] / s[0])
    ref = _pil(img, out_hw, method, box=box)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 2e-6


@pytest.mark.parametrize("method", list(rs.KERNELS))
def test_identity_at_unit_scale(method):
    img = np.random.default_rng(2).random((17, 23))
    got = rs.scale_and_translate(img, img.shape, (1.0, 1.0), (0.0, 0.0), method)
    np.testing.assert_allclose(got, img, atol=1e-12)


def test_linear_ramp_reproduced():
    # upsampling a linear ramp by linear interpolation reproduces it inside
    y, x = np.mgrid[0:20, 0:30].astype(np.float64)
    img = 0.3 * y - 1.7 * x + 2.0
    out_hw = (40, 60)
    got = rs.scale_and_translate(img, out_hw, (2.0, 2.0), (0.0, 0.0), "linear")
    oy, ox = np.mgrid[0:40, 0:60].astype(np.float64)
    sy, sx = (oy + 0.5) / 2 - 0.5, (ox + 0.5) / 2 - 0.5
    want = 0.3 * sy - 1.7 * sx + 2.0
    inner = (sy >= 0) & (sy <= 19) & (sx >= 0) & (sx <= 29)
    np.testing.assert_allclose(got[inner], want[inner], atol=1e-12)


def test_weights_are_normalised_and_antialiased():
    w = rs.weight_matrix(60, 20, 1 / 3, 0.0, "linear")
    np.testing.assert_allclose(w.sum(axis=0), 1.0, atol=1e-12)
    # downsampling by 3 widens the triangle to 3 input samples per side
    assert np.count_nonzero(w[:, 10]) == 5


def test_interpolating_resample_centres_the_fields():
    # same field of view: no translation; the division by the scale product
    # keeps sums (integrated intensity) for a downsampling by pooling-like weights
    img = np.random.default_rng(3).random((64, 64))
    out = rs.interpolating_resample(img, (32, 32), 0.6, 0.3, "linear")
    direct = rs.scale_and_translate(img, (32, 32), (0.5, 0.5), (0.0, 0.0), "linear") / 0.25
    np.testing.assert_allclose(out, direct, atol=1e-12)
    assert abs(out.sum() / img.sum() - 1.0) < 1e-2


def test_pooling():
    img = np.random.default_rng(4).random((2, 12, 18))
    out = rs.pooling_resample(img, (4, 6))
    np.testing.assert_allclose(out.sum(axis=(-2, -1)), img.sum(axis=(-2, -1)), rtol=1e-12)
    np.testing.assert_allclose(out[1, 2, 3], img[1, 6:9, 9:12].sum(), rtol=1e-12)


def test_sigmoid_taper_distance():
    shape, width = (9, 12), 1.5
    t = rs.sigmoid_taper(shape, width)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    border = (yy == 0) | (yy == shape[0] - 1) | (xx == 0) | (xx == shape[1] - 1)
    by, bx = yy[border], xx[border]
    d = np.sqrt((yy[..., None] - by) ** 2 + (xx[..., None] - bx) ** 2).min(axis=-1)
    np.testing.assert_allclose(t, 2 * (1 / (1 + np.exp(-d / width)) - 0.5), atol=1e-12)
    assert np.all(t[border] == 0.0)
