"""BasicSensor (src/chromatix/elements/sensors.py:14-132) over
``functional.basic_sensor``, with the resampler of
``ops.resample.init_plane_resample`` built from its shape, pitch and
``resampling_method``."""
from __future__ import annotations

from ..functional.sensors import basic_sensor
from ..ops.resample import init_plane_resample

__all__ = ["BasicSensor"]


class BasicSensor:
    """Intensity image of an incoming field or intensity with shot noise.

    shape: (height, width) in pixels; spacing: pixel pitch;
    shot_noise_mode: None, "approximate" or "poisson"; resampling_method:
    "linear", "cubic", "lanczos3", "lanczos5", "pool" or None;
    reduce_axis: axis summed over (e.g. depth planes reaching the sensor);
    reduce_parallel_axis_name: sum across processes as well (a
    torch.distributed ProcessGroup, or any name for the default group)."""

    def __init__(self, shape, spacing, shot_noise_mode=None, resampling_method="linear",
                 reduce_axis=None, reduce_parallel_axis_name=None):
        self.shape = tuple(int(s) for s in shape)
        self.spacing = spacing
        self.shot_noise_mode = shot_noise_mode
        self.resampling_method = resampling_method
        self.resampler = (init_plane_resample(self.shape, self.spacing, resampling_method)
                          if resampling_method is not None else None)
        self.reduce_axis = reduce_axis
        self.reduce_parallel_axis_name = reduce_parallel_axis_name

    def __call__(self, sensor_input, input_spacing=None, resample: bool = True, key=None):
        """Resample to the sensor's pixels (unless ``resample`` is False),
        reduce, add shot noise; ``key``: torch.Generator or integer seed."""
        resampler = self.resampler if (resample and self.resampling_method is not None) else None
        if self.shot_noise_mode is not None and key is None:
            raise ValueError("Must provide a PRNGKey if shot_noise_mode is not None")
        return basic_sensor(sensor_input, self.shot_noise_mode, resampler, self.reduce_axis,
                            self.reduce_parallel_axis_name, input_spacing=input_spacing,
                            noise_key=key)

    def resample(self, resample_input, input_spacing, mask=None):
        """The input's last two axes resampled to the sensor's pixels (the
        input unchanged without a resampling method); ``mask`` multiplies
        the input first (the Microscope's PSF taper)."""
        if self.resampling_method is not None:
            return self.resampler(resample_input, input_spacing, mask=mask)
        if mask is not None:
            return apply_mask(resample_input, mask)
        return resample_input


def apply_mask(x, mask):
    """x * mask over the last two axes (the resampling call with one-tap
    identity tables), differentiable in x."""
    from ..ops.resample import identity_tables, resample_planes

    H, W = x.shape[-2:]
    return resample_planes(x, identity_tables(H), identity_tables(W), (H, W), mask)
