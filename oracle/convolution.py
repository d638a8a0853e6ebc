"""Oracle: Fourier convolution and the basic sensor (test infrastructure, FP64).

Follows (paths relative to the chromatix repository root):
  next_order                     src/chromatix/utils/utils.py:11-12
  fourier_convolution            src/chromatix/ops/ops.py:9-86
  approximate_shot_noise         src/chromatix/ops/noise.py:9-16
  basic_sensor (no resampling)   src/chromatix/functional/sensors.py:14-93
  Microscope.image               src/chromatix/systems/microscopes.py:271-297

``fourier_convolution`` restates the reference's steps (zero-padded
fftn / rfftn at the padded or power-of-two shape, product, inverse, "same"
crop) in float64.  It is pinned by ``direct_convolution`` (the defining
double sum, brute force) on small cases and by scipy.signal.convolve, which
is the reference's own test oracle (tests/test_convolution.py).
"""
from __future__ import annotations

import numpy as np


def next_order(val: int) -> int:
    """Smallest power of two >= val (utils.py:11-12)."""
    return int(2 ** np.ceil(np.log2(val)))


def fourier_convolution(image, kernel, axes=(0, 1), fast_fft_shape=True, mode="same"):
    """ops.py:9-86 in FP64: transform shape n + k (rounded up to a power of
    two when fast_fft_shape) over ``axes``; real inputs use rfftn/irfftn."""
    image = np.asarray(image)
    kernel = np.asarray(kernel)
    assert image.ndim == kernel.ndim
    axes = [a + image.ndim if a < 0 else a for a in axes]
    for i in range(len(axes) - 1):
        assert axes[i + 1] == axes[i] + 1, "axes must be contiguous"
    padded = [image.shape[a] + kernel.shape[a] for a in axes]
    shape = [next_order(k) for k in padded] if fast_fft_shape else padded
    cplx = np.iscomplexobj(image) or np.iscomplexobj(kernel)
    if cplx:
        conv = np.fft.ifftn(np.fft.fftn(image.astype(np.complex128), s=shape, axes=axes) *
                            np.fft.fftn(kernel.astype(np.complex128), s=shape, axes=axes),
                            s=shape, axes=axes)
    else:
        conv = np.fft.irfftn(np.fft.rfftn(image.astype(np.float64), s=shape, axes=axes) *
                             np.fft.rfftn(kernel.astype(np.float64), s=shape, axes=axes),
                             s=shape, axes=axes)
    if mode == "same":
        sl = tuple(slice((k - 1) // 2, (k - 1) // 2 + i) if idx in axes else slice(i)
                   for idx, (i, k) in enumerate(zip(image.shape, kernel.shape)))
        conv = conv[sl]
    return conv


def direct_convolution(image, kernel):
    """Full linear convolution over the last two axes by the defining sum
    out[y, x] = sum_{j, i} image[y - j, x - i] kernel[j, i] (brute force)."""
    image = np.asarray(image, np.complex128)
    kernel = np.asarray(kernel, np.complex128)
    H, W = image.shape[-2:]
    Hk, Wk = kernel.shape[-2:]
    batch = np.broadcast_shapes(image.shape[:-2], kernel.shape[:-2])
    out = np.zeros(batch + (H + Hk - 1, W + Wk - 1), np.complex128)
    for j in range(Hk):
        for i in range(Wk):
            out[..., j: j + H, i: i + W] += image * kernel[..., j: j + 1, i: i + 1]
    return out


def approximate_shot_noise(image, normal):
    """max(image + sqrt(image) * normal, 0) with given standard normals
    (noise.py:9-16)."""
    image = np.asarray(image, np.float64)
    return np.maximum(image + np.sqrt(image) * np.asarray(normal, np.float64), 0.0)


def basic_sensor(intensity, shot_noise_mode=None, reduce_axis=None, normal=None):
    """sensors.py:14-93 for an intensity array without resampling or a
    cross-device sum; the approximate noise takes its normals as input."""
    image = np.asarray(intensity, np.float64)
    if reduce_axis is not None:
        image = image.sum(axis=reduce_axis)
    if shot_noise_mode == "approximate":
        image = approximate_shot_noise(image, normal)
    elif shot_noise_mode is not None:
        raise ValueError("the oracle draws no random samples")
    return image


def microscope_image(sample, psf, convolution_axes=(-2, -1), fast_fft_shape=True,
                     shot_noise_mode=None, reduce_axis=None, normal=None):
    """Microscope.image (microscopes.py:271-297): Fourier convolution of the
    sample with the PSF, then the sensor without resampling."""
    image = fourier_convolution(sample, psf, axes=convolution_axes, fast_fft_shape=fast_fft_shape)
    return basic_sensor(image, shot_noise_mode, reduce_axis, normal)
