"""Microscope (src/chromatix/systems/microscopes.py:20-297) and
Optical4FSystemPSF (microscopes.py:300-365) on libcxm.

``Microscope.image`` convolves the sample with the PSF intensity
(``ops.fourier_convolution``, one library call) and hands the image to the
sensor without resampling; when a BasicSensor sums the planes of a (D, H, W)
sample, the sum is taken inside the convolution
(``ops.planewise_convolution_sum``).  ``Microscope.__call__`` computes the PSF,
crops its padding, tapers its edges (``utils.sigmoid_taper``, a mask applied
inside the resampling call) and resamples it to the sensor's pixels.

``Optical4FSystemPSF`` takes any object with the objective's ``f``, ``n``,
``NA``, ``spectrum`` and ``padding_ratio`` (a ``Microscope`` has them).
"""
from __future__ import annotations

import numpy as np
import torch

from ..core.field import ScalarField
from ..core.spectrum import MonoSpectrum, Spectrum
from ..elements.sensors import BasicSensor, apply_mask
from ..functional.psf import psf_4f
from ..functional.sensors import basic_sensor
from ..ops import fourier_convolution, planewise_convolution_sum


_TAPERS: dict = {}


def _taper(shape, width, device) -> torch.Tensor:
    """sigmoid_taper (src/chromatix/utils/utils.py:76-81): 2 (sigmoid(d /
    width) - 1/2), d the Euclidean distance to a one-pixel zero border
    (scipy's exact distance transform, FP64 on the host), as a float32
    device mask; cached per shape, width and device."""
    key = (shape, width, str(device))
    t = _TAPERS.get(key)
    if t is None:
        from scipy.ndimage import distance_transform_edt

        dist = distance_transform_edt(np.pad(np.ones((shape[0] - 2, shape[1] - 2)), 1))
        mask = 2.0 * (1.0 / (1.0 + np.exp(-dist / width)) - 0.5)
        t = torch.as_tensor(mask.astype(np.float32), device=device)
        _TAPERS[key] = t
    return t


class Microscope:
    """Microscope with a planewise spatially invariant PSF: images are the
    sample convolved with the PSF intensity, seen by ``sensor``.

    system_psf(microscope, *args, **kwargs) returns a Field or an intensity
    tensor; ``convolution_axes`` (-2, -1) convolves plane by plane, (-3, -2,
    -1) in 3D; ``fast_fft_shape`` is accepted for the reference's interface
    (the transform lengths here are the implemented FFT lengths at or above
    the linear-convolution size either way)."""

    def __init__(self, system_psf, sensor, f, n, NA, spectrum, padding_ratio=0, taper_width=0,
                 convolution_axes=(-2, -1), fast_fft_shape=True):
        self.system_psf = system_psf
        self.sensor = sensor
        self.f = f
        self.n = n
        self.NA = NA
        self.spectrum = Spectrum.build(spectrum)
        self.padding_ratio = padding_ratio
        self.taper_width = taper_width
        self.convolution_axes = tuple(convolution_axes)
        self.fast_fft_shape = fast_fft_shape

    def __call__(self, sample, *args, key=None, **kwargs):
        """PSF for ``args`` / ``kwargs``, then ``image(sample, psf)``."""
        if isinstance(self.system_psf, Optical4FSystemPSF):
            # the fused intensity kernels: |field|^2 without storing the field
            # (the field's spacing is the requested output spacing)
            psf = self.system_psf.intensity(self, *args, **kwargs)
            psf = self._process_psf(psf, spacing=self.system_psf.spacing)
        else:
            psf = self._process_psf(self.psf(*args, **kwargs))
        return self.image(sample, psf, key=key)

    def psf(self, *args, **kwargs):
        """PSF of the system (microscopes.py:225-227)."""
        return self.system_psf(self, *args, **kwargs)

    def _process_psf(self, system_psf, spacing=None):
        """Intensity, padding crop and resampling to the sensor
        (microscopes.py:229-269); ``spacing``: that of an intensity
        ``system_psf`` when known."""
        if isinstance(system_psf, ScalarField):
            psf = system_psf.intensity
            spacing = system_psf.central_rectangular_dx
        else:
            psf = torch.as_tensor(system_psf)
        shape = tuple(psf.shape[-2:])
        if self.padding_ratio:
            unpadded = tuple(int(s / (1.0 + self.padding_ratio)) for s in shape)
            padding = tuple(s - u for s, u in zip(shape, unpadded))
        else:
            unpadded, padding = shape, (0, 0)
        if spacing is None:
            spacing = np.asarray(self.sensor.spacing, np.float64) * np.array(
                [self.sensor.shape[0] / (shape[0] - padding[0]),
                 self.sensor.shape[1] / (shape[1] - padding[1])])
        if self.padding_ratio and self.padding_ratio > 0:
            # center_crop by padding // 2 on each side (utils.py:26-41)
            cy, cx = padding[0] // 2, padding[1] // 2
            psf = psf[..., cy: shape[0] - cy, cx: shape[1] - cx]
        taper = None
        if self.taper_width > 0:
            taper = _taper(tuple(psf.shape[-2:]), float(self.taper_width), psf.device)
        if isinstance(self.sensor, BasicSensor):
            return self.sensor.resample(psf, spacing, mask=taper)
        if taper is not None:
            psf = apply_mask(psf, taper)
        return self.sensor.resample(psf, spacing)

    def image(self, sample, psf, key=None):
        """Image of ``sample`` through ``psf`` (both already at the sensor's
        pixels): Fourier convolution over ``convolution_axes``, then the
        sensor without resampling (microscopes.py:271-297)."""
        summed = self._summed_image(sample, psf)
        if summed is not None:
            return basic_sensor(summed, self.sensor.shot_noise_mode, None, None,
                                self.sensor.reduce_parallel_axis_name,
                                input_spacing=self.sensor.spacing, noise_key=key)
        image = fourier_convolution(sample, psf, axes=self.convolution_axes,
                                    fast_fft_shape=self.fast_fft_shape)
        return self.sensor(image, self.sensor.spacing, resample=False, key=key)

    def _summed_image(self, sample, psf):
        """The planewise convolution of a (D, H, W) sample already summed over
        the planes (``planewise_convolution_sum``) when a BasicSensor sums
        axis 0 of it; None otherwise."""
        sample = torch.as_tensor(sample)
        if torch.is_grad_enabled() and (sample.requires_grad or torch.as_tensor(psf).requires_grad):
            return None  # the differentiable convolution and sensor sum instead
        axes = tuple(a % sample.ndim for a in self.convolution_axes) if sample.ndim else ()
        if (not isinstance(self.sensor, BasicSensor) or sample.ndim != 3 or axes != (1, 2)
                or not isinstance(self.sensor.reduce_axis, int) or self.sensor.reduce_axis % 3 != 0
                or not sample.is_cuda):
            return None
        return planewise_convolution_sum(sample, psf)


class Optical4FSystemPSF:
    """PSF of a 4f system with a phase mask in the pupil plane.

    shape: (height, width) of the simulation; spacing: output spacing (e.g.
    the camera pitch); f_tube: tube-lens focal length; phase: (height,
    width) phase mask in radians (a torch tensor; set requires_grad for
    gradients through ``intensity``)."""

    def __init__(self, shape, spacing, f_tube, phase):
        self.shape = tuple(int(s) for s in shape)
        self.spacing = float(spacing)
        self.f_tube = float(f_tube)
        self.phase = phase

    @staticmethod
    def compute_required_spacing(height, output_spacing, f, n, spectrum):
        """Pupil-plane spacing giving `output_spacing` after the tube lens:
        f lam / (n height output_spacing) (microscopes.py:357-365)."""
        return f * spectrum / (n * height * output_spacing)

    def _setup(self, microscope):
        padding = tuple(int(s * getattr(microscope, "padding_ratio", 0)) for s in self.shape)
        padded = tuple(s + p for s, p in zip(self.shape, padding))
        spectrum = Spectrum.build(microscope.spectrum)
        if not isinstance(spectrum, MonoSpectrum):
            raise NotImplementedError("Optical4FSystemPSF: monochromatic spectra only")
        lam = float(spectrum.wavelength[0])
        dx = self.compute_required_spacing(padded[0], self.spacing, self.f_tube, microscope.n, lam)
        phase = torch.as_tensor(self.phase)
        if tuple(phase.shape) != padded:
            raise ValueError(f"phase must have the simulation shape {padded}")
        return phase, dx, lam

    def __call__(self, microscope, z) -> ScalarField:
        phase, dx, lam = self._setup(microscope)
        return psf_4f(phase, dx, lam, z, microscope.f, microscope.n, microscope.NA, self.f_tube,
                      output="field")

    def intensity(self, microscope, z) -> torch.Tensor:
        """|field|^2 of __call__, differentiable with respect to ``phase``."""
        phase, dx, lam = self._setup(microscope)
        return psf_4f(phase, dx, lam, z, microscope.f, microscope.n, microscope.NA, self.f_tube)
