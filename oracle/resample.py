"""Oracle: plane resampling and the PSF edge taper (test infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  PoolingPlaneDownsampler        src/chromatix/ops/resample.py:10-27
  InterpolatingPlaneResampler    src/chromatix/ops/resample.py:29-66
  init_plane_resample            src/chromatix/ops/resample.py:69-128
  sigmoid_taper                  src/chromatix/utils/utils.py:76-81
  Microscope._process_psf        src/chromatix/systems/microscopes.py:229-269

The interpolating resampler calls ``jax.image.scale_and_translate`` (jax,
``jax>=0.4.1`` in the reference's pyproject.toml; absent here), whose
published algorithm is restated below: per axis a weight matrix
``compute_weight_mat`` (kernel evaluated at the distance between each input
sample and the output sample's position mapped into input coordinates,
widened by 1/scale when downsampling -- the antialiasing -- normalised per
output sample, zeroed where the sample falls outside the input), applied as
two contractions.  Kernels: triangle ("linear"), Keys cubic with a = -0.5
("cubic"), Lanczos with radius 3 and 5.  The same resampling with the same
normalised, widened kernels is what Pillow's resize does for float images
(BILINEAR, BICUBIC, LANCZOS), which pins the restatement
(tests/test_oracle_resample.py).
"""
from __future__ import annotations

import numpy as np


def triangle(x):
    return np.maximum(0.0, 1.0 - np.abs(x))


def keys_cubic(x):
    """Keys' cubic convolution kernel with a = -0.5."""
    x = np.abs(x)
    out = ((1.5 * x - 2.5) * x) * x + 1.0
    out = np.where(x >= 1.0, ((-0.5 * x + 2.5) * x - 4.0) * x + 2.0, out)
    return np.where(x >= 2.0, 0.0, out)


def lanczos(radius):
    def k(x):
        x = np.abs(x)
        y = radius * np.sin(np.pi * x) * np.sin(np.pi * x / radius)
        with np.errstate(divide="ignore", invalid="ignore"):
            out = np.where(x > 1e-3, y / np.where(x != 0, np.pi**2 * x**2, 1.0), 1.0)
        return np.where(x > radius, 0.0, out)
    return k


KERNELS = {"linear": triangle, "bilinear": triangle, "triangle": triangle,
           "cubic": keys_cubic, "bicubic": keys_cubic, "keyscubic": keys_cubic,
           "lanczos3": lanczos(3.0), "lanczos5": lanczos(5.0)}


def weight_matrix(in_size: int, out_size: int, scale: float, translation: float, method: str,
                  antialias: bool = True) -> np.ndarray:
    """(in_size, out_size) weights of scale_and_translate along one axis:
    output sample j sits at input coordinate (j + 0.5 - translation) / scale
    - 0.5; the kernel is widened by 1/scale when downsampling."""
    kernel = KERNELS[method]
    inv = 1.0 / scale
    kscale = max(inv, 1.0) if antialias else 1.0
    sample_f = (np.arange(out_size) + 0.5) * inv - translation * inv - 0.5
    x = np.abs(sample_f[None, :] - np.arange(in_size)[:, None]) / kscale
    w = kernel(x)
    tot = np.sum(w, axis=0, keepdims=True)
    w = np.where(np.abs(tot) > 1000.0 * float(np.finfo(np.float32).eps),
                 w / np.where(tot != 0, tot, 1.0), 0.0)
    inside = (sample_f >= -0.5) & (sample_f <= in_size - 0.5)
    return np.where(inside[None, :], w, 0.0)


def scale_and_translate(img, out_shape, scale, translation, method="linear"):
    """Resample the last two axes of ``img``."""
    wy = weight_matrix(img.shape[-2], out_shape[0], scale[0], translation[0], method)
    wx = weight_matrix(img.shape[-1], out_shape[1], scale[1], translation[1], method)
    return np.einsum("...yx,yi,xj->...ij", img, wy, wx)


def interpolating_resample(img, out_shape, out_spacing, in_spacing, method="linear"):
    """InterpolatingPlaneResampler.__call__ (resample.py:49-66) on the last
    two axes: scale = in_spacing / out_spacing, translation centring the
    fields of view, result divided by the scale product."""
    in_sp = np.broadcast_to(np.asarray(in_spacing, np.float64), (2,))
    out_sp = np.broadcast_to(np.asarray(out_spacing, np.float64), (2,))
    scale = in_sp / out_sp
    translation = -0.5 * (np.asarray(img.shape[-2:], np.float64) * scale - np.asarray(out_shape))
    return scale_and_translate(img, out_shape, scale, translation, method) / np.prod(scale)


def pooling_resample(img, out_shape):
    """PoolingPlaneDownsampler.__call__ (resample.py:21-27): sum over
    (h hf) (w wf) -> h w on the last two axes."""
    H, W = img.shape[-2:]
    h, w = out_shape
    fh, fw = H // h, W // w
    return img.reshape(img.shape[:-2] + (h, fh, w, fw)).sum(axis=(-3, -1))


def sigmoid_taper(shape, width):
    """utils.py:76-81: 2 (sigmoid(d / width) - 1/2), d the Euclidean distance
    transform of a one-pixel zero border around ones."""
    from scipy.ndimage import distance_transform_edt

    dist = distance_transform_edt(np.pad(np.ones((shape[0] - 2, shape[1] - 2)), 1))
    return 2.0 * (1.0 / (1.0 + np.exp(-dist / width)) - 0.5)
