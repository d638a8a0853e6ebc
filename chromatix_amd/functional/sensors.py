"""The basic sensor on libcxm (mirrors src/chromatix/functional/sensors.py).

``basic_sensor`` (sensors.py:14-93): intensity of a field (or a given
intensity), resampled to the sensor's pixels by a resampler of
``ops.resample.init_plane_resample`` (cxm_resample_f32), summed over
``reduce_axis`` (cxm_sum_axis_f32, in index order) and across processes
(the reference's ``jax.lax.psum`` over a pmapped axis: here an all-reduce
over a ``torch.distributed`` process group, one process per GPU), then
shot noise.
"""
from __future__ import annotations

import torch

from .. import _lib as L
from ..core.field import Field
from ..ops import approximate_shot_noise, shot_noise

__all__ = ["basic_sensor", "sum_axis"]


class _SumAxisFn(torch.autograd.Function):
    """sum_axis with its reverse-mode rule (the cotangent broadcast along
    the summed axis)."""

    @staticmethod
    def forward(ctx, x, ax):
        ctx.shape, ctx.ax = tuple(x.shape), ax
        return _sum_axis(x.detach(), ax)

    @staticmethod
    def backward(ctx, g):
        return g.unsqueeze(ctx.ax).expand(ctx.shape).resolve_conj().contiguous(), None


def sum_axis(x: torch.Tensor, axis: int) -> torch.Tensor:
    """Sum of a float32 device tensor over ``axis`` (terms added in index
    order, one thread per output element); differentiable."""
    x = torch.as_tensor(x)
    L.ensure_device(x)
    x = x.to(torch.float32).resolve_conj().contiguous()
    nd = x.ndim
    ax = axis + nd if axis < 0 else axis
    if not 0 <= ax < nd:
        raise ValueError(f"axis {axis} out of range for {nd} dimensions")
    if torch.is_grad_enabled() and x.requires_grad:
        return _SumAxisFn.apply(x, ax)
    return _sum_axis(x, ax)


def _sum_axis(x: torch.Tensor, ax: int) -> torch.Tensor:
    shape = tuple(x.shape)
    outer = 1
    for s in shape[:ax]:
        outer *= s
    inner = 1
    for s in shape[ax + 1:]:
        inner *= s
    out = torch.empty(shape[:ax] + shape[ax + 1:], dtype=torch.float32, device=x.device)
    L.check(L.lib().cxm_sum_axis_f32(L.stream_ptr(x.device), x.data_ptr(), out.data_ptr(),
                                     outer, shape[ax], inner))
    return out


def _all_reduce(image: torch.Tensor, group) -> torch.Tensor:
    """Sum across the processes of ``group`` (a ProcessGroup, or any name for
    the default group)."""
    import torch.distributed as dist

    if not dist.is_available() or not dist.is_initialized():
        raise RuntimeError("reduce_parallel_axis_name needs an initialised torch.distributed "
                           "process group (one process per device)")
    pg = group if isinstance(group, dist.ProcessGroup) else None
    if dist.get_backend(pg) == "gloo" and image.is_cuda:
        # gloo reduces host tensors: copy out and back synchronously
        host = image.to("cpu")
        dist.all_reduce(host, op=dist.ReduceOp.SUM, group=pg)
        return host.to(image.device)
    image = image.clone()
    dist.all_reduce(image, op=dist.ReduceOp.SUM, group=pg)
    return image


def basic_sensor(sensor_input, shot_noise_mode=None, resampler=None, reduce_axis=None,
                 reduce_parallel_axis_name=None, input_spacing=None, noise_key=None) -> torch.Tensor:
    """Intensity image of ``sensor_input`` (a Field or an intensity tensor),
    summed over ``reduce_axis`` and across the processes named by
    ``reduce_parallel_axis_name``, with shot noise ("approximate" or
    "poisson"; ``noise_key`` a torch.Generator or an integer seed)."""
    if isinstance(sensor_input, Field):
        intensity = sensor_input.intensity
        # the spacing of the first wavelength (sensors.py:67-70)
        input_spacing = sensor_input.central_rectangular_dx
    else:
        intensity = torch.as_tensor(sensor_input)
    if resampler is not None:
        if input_spacing is None:
            raise ValueError("Must provide input_spacing for intensity array")
        image = resampler(intensity, input_spacing)
    else:
        image = intensity
    if reduce_axis is not None:
        image = sum_axis(image, reduce_axis)
    if reduce_parallel_axis_name is not None:
        image = _all_reduce(image, reduce_parallel_axis_name)
    if shot_noise_mode is not None and noise_key is None:
        raise ValueError("Must provide a PRNG key to generate noise")
    if shot_noise_mode == "approximate":
        image = approximate_shot_noise(noise_key, image)
    elif shot_noise_mode == "poisson":
        image = shot_noise(noise_key, image)
    elif shot_noise_mode is not None:
        raise ValueError(f"unknown shot_noise_mode {shot_noise_mode!r}")
    return image
