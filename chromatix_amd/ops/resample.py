"""Plane resampling on libcxm (mirrors src/chromatix/ops/resample.py).

``init_plane_resample`` (resample.py:69-128) returns a resampler of the
last two axes of a float32 tensor: sum pooling ("pool",
PoolingPlaneDownsampler, :10-27) or interpolation by
``jax.image.scale_and_translate`` ("linear", "cubic", "lanczos3",
"lanczos5"; InterpolatingPlaneResampler, :29-66) divided by the scale
product.  The reference's dependency computes, per axis, a weight matrix
(the kernel at the distance between each input sample and the output
sample's position in input coordinates, widened by 1/scale when
downsampling, normalised per output sample, zero where the sample falls
outside the input) and contracts with it; here those matrices are built on
the host in FP64 (they depend only on shapes, spacings and the method, and
are cached), reduced to per-output tap tables, and applied by
``cxm_resample_f32`` in two passes.  The resampler is linear; its
reverse-mode rule is the same call with the transposed tables.
"""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from .. import _lib as L

__all__ = ["init_plane_resample", "PoolingPlaneDownsampler", "InterpolatingPlaneResampler",
           "resample_planes", "identity_tables"]


# ---- the weight matrices of jax.image.scale_and_translate (host, FP64) -----
def _triangle(x):
    return np.maximum(0.0, 1.0 - np.abs(x))


def _keys_cubic(x):
    # Keys' cubic convolution kernel, a = -0.5
    x = np.abs(x)
    out = ((1.5 * x - 2.5) * x) * x + 1.0
    out = np.where(x >= 1.0, ((-0.5 * x + 2.5) * x - 4.0) * x + 2.0, out)
    return np.where(x >= 2.0, 0.0, out)


def _lanczos(radius):
    def k(x):
        x = np.abs(x)
        y = radius * np.sin(np.pi * x) * np.sin(np.pi * x / radius)
        with np.errstate(divide="ignore", invalid="ignore"):
            out = np.where(x > 1e-3, y / np.where(x != 0, np.pi**2 * x**2, 1.0), 1.0)
        return np.where(x > radius, 0.0, out)
    return k


_KERNELS = {"linear": _triangle, "bilinear": _triangle, "triangle": _triangle, "cubic": _keys_cubic,
            "bicubic": _keys_cubic, "keyscubic": _keys_cubic, "lanczos3": _lanczos(3.0),
            "lanczos5": _lanczos(5.0)}


def _weights(in_size, out_size, scale, translation, method):
    """(in_size, out_size) weights along one axis (antialiased)."""
    kernel = _KERNELS[method]
    inv = 1.0 / scale
    kscale = max(inv, 1.0)
    sample_f = (np.arange(out_size) + 0.5) * inv - translation * inv - 0.5
    w = kernel(np.abs(sample_f[None, :] - np.arange(in_size)[:, None]) / kscale)
    tot = np.sum(w, axis=0, keepdims=True)
    w = np.where(np.abs(tot) > 1000.0 * float(np.finfo(np.float32).eps),
                 w / np.where(tot != 0, tot, 1.0), 0.0)
    inside = (sample_f >= -0.5) & (sample_f <= in_size - 0.5)
    return np.where(inside[None, :], w, 0.0)


def _taps(W):
    """Tap table of the contraction out[j] = sum_i W[i, j] in[i]: first
    input index per output (int32) and the weights from it (float32, zero
    past each output's last nonzero)."""
    n_in, n_out = W.shape
    nz = W != 0
    has = nz.any(axis=0)
    first = np.where(has, nz.argmax(axis=0), 0)
    last = np.where(has, n_in - 1 - nz[::-1].argmax(axis=0), 0)
    k = int(max(1, (last - first + 1)[has].max() if has.any() else 1))
    start = np.minimum(first, n_in - k).astype(np.int32)
    tab = np.zeros((n_out, k), np.float64)
    for j in np.nonzero(has)[0]:
        s = start[j]
        tab[j] = W[s: s + k, j]
    return start, tab.astype(np.float32)


@functools.lru_cache(maxsize=64)
def _axis_tables(kind, in_size, out_size, scale, translation, method):
    """(forward, adjoint) tap tables of one axis; the forward contraction
    folds in 1/scale (the division by the scale product)."""
    if kind == "pool":
        f = in_size // out_size
        W = np.zeros((in_size, out_size))
        for j in range(out_size):
            W[j * f:(j + 1) * f, j] = 1.0
    else:
        W = _weights(in_size, out_size, scale, translation, method) / scale
    return _taps(W), _taps(W.T)


@functools.lru_cache(maxsize=64)
def identity_tables(n):
    """(forward, adjoint) one-tap identity tables of an axis of length n."""
    t = (np.arange(n, dtype=np.int32), np.ones((n, 1), np.float32))
    return t, t


_DEV_CACHE: dict = {}


def _dev(tables, device):
    key = (id(tables), str(device))
    hit = _DEV_CACHE.get(key)
    if hit is None or hit[0] is not tables:
        (s, w) = tables
        hit = (tables, torch.as_tensor(s, device=device), torch.as_tensor(w, device=device))
        _DEV_CACHE[key] = hit
    return hit[1], hit[2]


def _run(x, ty, tx, out_hw, mask, mask_on_output):
    """cxm_resample_f32 over the planes of x (..., H, W) -> (..., h, w)."""
    H, W = x.shape[-2:]
    lead = tuple(x.shape[:-2])
    batch = int(np.prod(lead)) if lead else 1
    out = torch.empty(lead + tuple(out_hw), dtype=torch.float32, device=x.device)
    ys, yw = _dev(ty, x.device)
    xs, xw = _dev(tx, x.device)
    d = L.ResampleDesc()
    d.batch, d.in_h, d.in_w, d.out_h, d.out_w = batch, H, W, out_hw[0], out_hw[1]
    d.ky, d.kx = yw.shape[1], xw.shape[1]
    d.mask_on_output = 1 if mask_on_output else 0
    d.y_start, d.wy, d.x_start, d.wx = ys.data_ptr(), yw.data_ptr(), xs.data_ptr(), xw.data_ptr()
    d.mask = mask.data_ptr() if mask is not None else None
    ws = torch.empty(max(1, L.lib().cxm_resample_workspace_size(ctypes.byref(d))), dtype=torch.uint8,
                     device=x.device)
    L.check(L.lib().cxm_resample_f32(L.stream_ptr(x.device), x.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                     ws.numel(), ctypes.byref(d)))
    return out


class _ResampleFn(torch.autograd.Function):
    """Resampling with its reverse-mode rule: the transposed tap tables with
    the mask on the output."""

    @staticmethod
    def forward(ctx, x, tables, out_hw, mask):
        ctx.tables, ctx.in_hw, ctx.mask = tables, tuple(x.shape[-2:]), mask
        (fy, _), (fx, _) = tables
        return _run(x.detach(), fy, fx, out_hw, mask, False)

    @staticmethod
    def backward(ctx, g):
        (_, ay), (_, ax) = ctx.tables
        g = g.to(torch.float32).resolve_conj().contiguous()
        return _run(g, ay, ax, ctx.in_hw, ctx.mask, True), None, None, None


def resample_planes(x, y_tables, x_tables, out_hw, mask=None):
    """Resample the last two axes of the float32 device tensor ``x`` by the
    per-axis (forward, adjoint) tap tables; ``mask`` (H, W) multiplies the
    input first.  Differentiable in ``x``."""
    x = torch.as_tensor(x)
    L.ensure_device(x)
    x = x.to(torch.float32).resolve_conj().contiguous()
    if mask is not None:
        mask = torch.as_tensor(mask, dtype=torch.float32, device=x.device).contiguous()
        if tuple(mask.shape) != tuple(x.shape[-2:]):
            raise ValueError("the mask must have the planes' shape")
    tables = (y_tables, x_tables)
    if torch.is_grad_enabled() and x.requires_grad:
        return _ResampleFn.apply(x, tables, tuple(out_hw), mask)
    return _run(x, y_tables[0], x_tables[0], tuple(out_hw), mask, False)


def _spacing2(s):
    return tuple(float(v) for v in np.broadcast_to(np.asarray(s, np.float64).reshape(-1), (2,)))


class PoolingPlaneDownsampler:
    """Sum pooling (h hf) (w wf) -> h w of the last two axes (resample.py:10-27)."""

    def __init__(self, out_shape, out_spacing):
        self.out_shape = tuple(int(s) for s in out_shape)
        self.out_spacing = out_spacing

    def __call__(self, resample_input, in_spacing=None, mask=None):
        H, W = resample_input.shape[-2:]
        h, w = self.out_shape
        if H % h or W % w:
            raise ValueError(f"pooling needs the input shape {(H, W)} to be a multiple of {(h, w)}")
        ty = _axis_tables("pool", H, h, 1.0, 0.0, "")
        tx = _axis_tables("pool", W, w, 1.0, 0.0, "")
        return resample_planes(resample_input, ty, tx, self.out_shape, mask)


class InterpolatingPlaneResampler:
    """scale_and_translate of the last two axes to ``out_shape`` at
    ``out_spacing``, the fields of view centred, divided by the scale
    product (resample.py:29-66)."""

    def __init__(self, out_shape, out_spacing, resampling_method="linear"):
        if resampling_method not in _KERNELS:
            raise ValueError(f"unknown resampling method {resampling_method!r}")
        self.out_shape = tuple(int(s) for s in out_shape)
        self.out_spacing = out_spacing
        self.resampling_method = resampling_method

    def __call__(self, resample_input, in_spacing, mask=None):
        in_sp = _spacing2(in_spacing)
        out_sp = _spacing2(self.out_spacing)
        H, W = resample_input.shape[-2:]
        scale = (in_sp[0] / out_sp[0], in_sp[1] / out_sp[1])
        if (H, W) == self.out_shape and scale == (1.0, 1.0):
            # the weights are the identity (every kernel is 1 at 0 and 0 at
            # the other integers) and the scale product 1
            if mask is None:
                return torch.as_tensor(resample_input)
            return resample_planes(resample_input, identity_tables(H), identity_tables(W), (H, W), mask)
        trans = (-0.5 * (H * scale[0] - self.out_shape[0]), -0.5 * (W * scale[1] - self.out_shape[1]))
        ty = _axis_tables("interp", H, self.out_shape[0], scale[0], trans[0], self.resampling_method)
        tx = _axis_tables("interp", W, self.out_shape[1], scale[1], trans[1], self.resampling_method)
        return resample_planes(resample_input, ty, tx, self.out_shape, mask)


def init_plane_resample(out_shape, out_spacing, resampling_method="linear"):
    """A resampler of the last two axes to ``out_shape`` / ``out_spacing``:
    "pool" for sum pooling, else "linear", "cubic", "lanczos3" or
    "lanczos5" interpolation (resample.py:69-128)."""
    if len(tuple(out_shape)) != 2:
        raise ValueError("Shape must be tuple of form (height width)")
    if np.atleast_1d(np.asarray(out_spacing).squeeze()).size > 2:
        raise ValueError("Spacing is either a float or array of shape (2,) for non-square pixels")
    if resampling_method == "pool":
        return PoolingPlaneDownsampler(out_shape, out_spacing)
    return InterpolatingPlaneResampler(out_shape, out_spacing, resampling_method)
