"""BasicSensor (src/chromatix/elements/sensors.py:14-132) over
``functional.basic_sensor``.  Resampling to the sensor's pixels is not
implemented: a sensor built with a ``resampling_method`` works wherever it
is called with ``resample=False`` (as ``Microscope.image`` does) or on an
input already at the sensor's shape and pitch."""
from __future__ import annotations

import numpy as np
import torch

from ..functional.sensors import basic_sensor

__all__ = ["BasicSensor"]


class BasicSensor:
    """Intensity image of an incoming field or intensity with shot noise.

    shape: (height, width) in pixels; spacing: pixel pitch;
    shot_noise_mode: None, "approximate" or "poisson"; reduce_axis: axis
    summed over (e.g. depth planes reaching the sensor);
    reduce_parallel_axis_name: sum across processes as well (a
    torch.distributed ProcessGroup, or any name for the default group)."""

    def __init__(self, shape, spacing, shot_noise_mode=None, resampling_method="linear",
                 reduce_axis=None, reduce_parallel_axis_name=None):
        self.shape = tuple(int(s) for s in shape)
        self.spacing = spacing
        self.shot_noise_mode = shot_noise_mode
        self.resampling_method = resampling_method
        self.reduce_axis = reduce_axis
        self.reduce_parallel_axis_name = reduce_parallel_axis_name

    def _needs_resampling(self, x, input_spacing) -> bool:
        shape = tuple(x.shape[-2:])
        if input_spacing is None:
            return shape != self.shape
        same = np.allclose(np.broadcast_to(np.asarray(input_spacing, np.float64), (2,)),
                           np.broadcast_to(np.asarray(self.spacing, np.float64), (2,)),
                           rtol=1e-9, atol=0.0)
        return not (same and shape == self.shape)

    def __call__(self, sensor_input, input_spacing=None, resample: bool = True, key=None):
        """``key``: torch.Generator or integer seed for the shot noise."""
        if resample and self.resampling_method is not None:
            x = sensor_input.intensity if hasattr(sensor_input, "intensity") else sensor_input
            sp = input_spacing
            if sp is None and hasattr(sensor_input, "central_rectangular_dx"):
                sp = sensor_input.central_rectangular_dx
            if self._needs_resampling(x, sp):
                raise NotImplementedError("BasicSensor: resampling to the sensor's pixels is not "
                                          "implemented")
        if self.shot_noise_mode is not None and key is None:
            raise ValueError("Must provide a PRNGKey if shot_noise_mode is not None")
        return basic_sensor(sensor_input, self.shot_noise_mode, None, self.reduce_axis,
                            self.reduce_parallel_axis_name, input_spacing=input_spacing,
                            noise_key=key)

    def resample(self, resample_input: torch.Tensor, input_spacing) -> torch.Tensor:
        """Identity when the input is already at the sensor's shape and pitch
        (or no resampling method is set); otherwise not implemented."""
        if self.resampling_method is None:
            return resample_input
        if self._needs_resampling(resample_input, input_spacing):
            raise NotImplementedError("BasicSensor: resampling to the sensor's pixels is not "
                                      "implemented")
        return resample_input
