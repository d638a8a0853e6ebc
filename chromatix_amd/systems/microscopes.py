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

from ..core.field import ChromaticScalarField, ScalarField
from ..core.spectrum import MonoSpectrum, Spectrum
from ..elements.sensors import BasicSensor, apply_mask
from ..functional.psf import psf_4f
from ..functional.sensors import basic_sensor, sum_axis
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
        phase = torch.as_tensor(self.phase)
        if tuple(phase.shape) != padded:
            raise ValueError(f"phase must have the simulation shape {padded}")
        return phase, padded, spectrum

    def _wavelengths(self, microscope, padded, spectrum, with_density):
        """Per wavelength k of the spectrum: (lam_k, dx_k, amplitude_k, m_k)
        for one monochromatic PSF call each.  dx_k = f_tube lam_k / (n H
        spacing) keeps every wavelength's output spacing at ``spacing``; the
        point sources of all wavelengths are normalised as one chromatic
        field (sources.py:206-208, field.py:729-733: u_k *= sqrt(P / (dx_k^2
        S)) with S = sum_k w_k M_k / L_k^4, M_k the samples inside the pupil,
        L_k^2 = lam_k f / n), so the call takes the amplitude 1 * sqrt(P /
        (dx_k^2 S)) without its own normalisation; ``with_density`` folds
        sqrt(w_k) in as well, which makes the intensity call return w_k
        |out_k|^2 (field.py:735-740); m_k = lam_0 / lam_k the phase
        modulation (phase_masks.py:105-124)."""
        lam = np.asarray(spectrum.wavelength, np.float64)
        w = np.asarray(spectrum.density, np.float64)
        f, n, NA = float(microscope.f), float(microscope.n), float(microscope.NA)
        dxs = [self.compute_required_spacing(padded[0], self.spacing, self.f_tube, n, lk) for lk in lam]
        S = sum(wk * _pupil_samples(padded, dk, f, NA, n) / (lk * f / n) ** 2
                for wk, dk, lk in zip(w, dxs, lam))
        out = []
        for lk, dk, wk, mk in zip(lam, dxs, w, spectrum.spectral_modulation):
            amp = float(np.sqrt(1.0 / (dk * dk * S)))  # power 1 (objective_point_source's default)
            if with_density:
                amp *= float(np.sqrt(wk))
            out.append((float(lk), dk, amp, float(mk)))
        return out

    def __call__(self, microscope, z):
        phase, padded, spectrum = self._setup(microscope)
        f, n, NA = microscope.f, microscope.n, microscope.NA
        if isinstance(spectrum, MonoSpectrum):
            lam = float(spectrum.wavelength[0])
            dx = self.compute_required_spacing(padded[0], self.spacing, self.f_tube, n, lam)
            return psf_4f(phase, dx, lam, z, f, n, NA, self.f_tube, output="field")
        fields = [psf_4f(phase, dk, lk, z, f, n, NA, self.f_tube, power=None, amplitude=ak,
                         spectral_modulation=mk, output="field")
                  for lk, dk, ak, mk in self._wavelengths(microscope, padded, spectrum, False)]
        u = torch.stack([fl.u for fl in fields], dim=-1)  # (..., H, W, wv)
        du = np.stack([np.asarray(fl.dx, np.float64).reshape(2) for fl in fields])
        return ChromaticScalarField(u, du, 0.0, spectrum)

    def intensity(self, microscope, z) -> torch.Tensor:
        """|field|^2 of __call__ (for a chromatic spectrum, sum_k w_k
        |out_k|^2), differentiable with respect to ``phase``."""
        phase, padded, spectrum = self._setup(microscope)
        f, n, NA = microscope.f, microscope.n, microscope.NA
        if isinstance(spectrum, MonoSpectrum):
            lam = float(spectrum.wavelength[0])
            dx = self.compute_required_spacing(padded[0], self.spacing, self.f_tube, n, lam)
            return psf_4f(phase, dx, lam, z, f, n, NA, self.f_tube)
        planes = [psf_4f(phase, dk, lk, z, f, n, NA, self.f_tube, power=None, amplitude=ak,
                         spectral_modulation=mk)
                  for lk, dk, ak, mk in self._wavelengths(microscope, padded, spectrum, True)]
        return sum_axis(torch.stack(planes, dim=0), 0)


def _pupil_samples(shape, dx, f, NA, n) -> int:
    """Samples of a (H, W) grid of spacing dx inside the circular pupil of
    diameter 2 f NA / n (pupils.py:17-23), in the grid's float32 arithmetic
    (the count the point-source kernel's pupil test gives)."""
    H, W = shape
    d = np.float32(dx)
    gy = d * (np.arange(H, dtype=np.float32) - np.float32(0.5 * H))
    gx = d * (np.arange(W, dtype=np.float32) - np.float32(0.5 * W))
    r = np.float32(2.0 * f * NA / n) / np.float32(2.0)
    l2 = gy[:, None] * gy[:, None] + gx[None, :] * gx[None, :]
    return int(np.count_nonzero(l2 <= r * r))
