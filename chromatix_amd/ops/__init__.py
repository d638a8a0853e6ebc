"""Fourier convolution and shot noise on libcxm (mirrors chromatix.ops).

``fourier_convolution`` (src/chromatix/ops/ops.py:9-86) runs as one library
call (cxm_fconv_c64): the kernel spectra once, then zero-padded forward
transforms, the spectral product and the inverse transforms on the
propagation passes.  The transform lengths are the implemented FFT lengths
at or above n + k - 1 per axis (the linear convolution); the reference's
n + k or power-of-two shapes give the same linear convolution, so results
agree to float32 rounding, and the "full" output keeps the reference's
shape (zeros past the linear convolution).

``approximate_shot_noise`` / ``shot_noise`` (src/chromatix/ops/noise.py:
9-45) take a ``torch.Generator`` or an integer seed in place of a JAX
PRNG key: the distributions are the reference's, the samples are torch's.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .. import _lib as L

__all__ = ["fourier_convolution", "planewise_convolution_sum", "approximate_shot_noise",
           "shot_noise", "next_order"]


def next_order(val: int) -> int:
    """Smallest power of two >= val (src/chromatix/utils/utils.py:11-12)."""
    return int(2 ** math.ceil(math.log2(val)))


def _kernel_planes(kb: tuple[int, ...], ob: tuple[int, ...]) -> int | None:
    """Kernel volumes when output volume b can use kernel volume b % k:
    kernel batch dims equal the output's over a trailing run and are 1
    before it.  None when the kernel must be broadcast explicitly."""
    kb = (1,) * (len(ob) - len(kb)) + tuple(kb)
    for s in range(len(ob) + 1):  # s: first dim of the trailing run
        if all(d == 1 for d in kb[:s]) and tuple(kb[s:]) == tuple(ob[s:]):
            return int(np.prod(ob[s:], dtype=np.int64))
    return None


def fourier_convolution(image, kernel, *, axes=(0, 1), fast_fft_shape: bool = True,
                        mode: str = "same") -> torch.Tensor:
    """Convolution of ``image`` with ``kernel`` over one, two or three
    contiguous ``axes`` by the Fourier convolution theorem, zero-padded against
    circular wrap.  ``mode`` "same" crops to the image's shape (window
    starting at (k - 1) // 2 per axis, as the reference), "full" returns the
    reference's transform shape (n + k per axis, rounded up to a power of two
    when ``fast_fft_shape``).  Real inputs give a float32 result (the
    reference's rfftn / irfftn), complex ones complex64.  Dimensions outside
    ``axes`` broadcast; in "same" mode they are cropped to the image's, as
    the reference's slices do."""
    image = torch.as_tensor(image)
    L.ensure_device(image)
    kernel = torch.as_tensor(kernel, device=image.device)
    if image.ndim != kernel.ndim:
        raise ValueError(f"Input ({image.ndim}D) and kernel ({kernel.ndim}D) must have same "
                         "number of dimensions")
    if mode not in ("same", "full"):
        raise ValueError(f"mode must be 'same' or 'full', got {mode!r}")
    nd = image.ndim
    axes = [a + nd if a < 0 else a for a in axes]
    for i in range(len(axes) - 1):
        if axes[i + 1] != axes[i] + 1:
            raise ValueError("Axes to convolve over must be contiguous")
    na = len(axes)
    if na == 1:
        # one axis: the two-axis convolution with a unit axis before it
        ax = axes[0]
        out = fourier_convolution(image.unsqueeze(ax), kernel.unsqueeze(ax), axes=(ax, ax + 1),
                                  fast_fft_shape=fast_fft_shape, mode=mode)
        if mode == "full":
            out = out.narrow(ax, 0, 1)
        return out.squeeze(ax)
    if na not in (2, 3):
        raise NotImplementedError("fourier_convolution: convolution over one, two or three axes")
    cplx = image.is_complex() or kernel.is_complex()
    dt = torch.complex64 if cplx else torch.float32
    perm = [i for i in range(nd) if i not in axes] + axes
    img = image.permute(perm).to(dt)
    ker = kernel.permute(perm).to(dt)
    ib, kb = tuple(img.shape[:-na]), tuple(ker.shape[:-na])
    isp, ksp = tuple(img.shape[-na:]), tuple(ker.shape[-na:])
    bshape = tuple(torch.broadcast_shapes(ib, kb))
    if mode == "same":
        # the reference slices every non-convolved dim to the image's length
        ob = ib
        ker = ker[tuple(slice(0, i) if k > 1 else slice(None) for i, k in zip(ib, kb))]
        kb = tuple(ker.shape[:-na])
        out_sp = isp
    else:
        ob = bshape
        padded = [i + k for i, k in zip(isp, ksp)]
        out_sp = tuple(next_order(p) for p in padded) if fast_fft_shape else tuple(padded)
    if ib != ob:
        img = img.expand(ob + isp)
    k_planes = _kernel_planes(kb, ob)
    if k_planes is None:
        ker = ker.expand(ob + ksp)
        k_planes = int(np.prod(ob, dtype=np.int64))
    img = img.resolve_conj().contiguous()
    ker = ker.resolve_conj().contiguous()
    meta = (na, mode, ob, isp, ksp, out_sp, k_planes, cplx)
    if torch.is_grad_enabled() and (img.requires_grad or ker.requires_grad):
        out = _FconvFn.apply(img, ker, meta)
    else:
        out = _fconv_core(img, ker, meta)
    inv = [0] * nd
    for i, p in enumerate(perm):
        inv[p] = i
    return out.permute(inv)


def _fconv_core(img, ker, meta) -> torch.Tensor:
    """One cxm_fconv_c64 call on contiguous (batch..., spatial...) planes;
    kernel plane b % k_planes serves output plane b."""
    na, mode, ob, isp, ksp, out_sp, k_planes, cplx = meta
    dt = torch.complex64 if cplx else torch.float32
    img = img.detach().to(dt).resolve_conj().contiguous()
    ker = ker.detach().to(dt).resolve_conj().contiguous()
    batch = int(np.prod(ob, dtype=np.int64))
    d = L.FconvDesc()
    d.batch = batch
    (d.depth, d.height, d.width) = (1,) * (3 - na) + tuple(isp)
    d.k_planes = k_planes
    (d.k_depth, d.k_height, d.k_width) = (1,) * (3 - na) + tuple(ksp)
    d.real_input = 0 if cplx else 1
    d.mode = 0 if mode == "same" else 1
    (d.out_depth, d.out_height, d.out_width) = (1,) * (3 - na) + tuple(out_sp)
    lib = L.lib()
    ws_bytes = lib.cxm_fconv_workspace_size(d)
    if ws_bytes == 0:
        L.check(lib.cxm_fconv_c64(None, None, None, None, None, 0, d))
    if mode == "full":
        # entries past the transform used here are not written
        out = torch.zeros(tuple(ob) + tuple(out_sp), dtype=dt, device=img.device)
    else:
        out = torch.empty(tuple(ob) + tuple(out_sp), dtype=dt, device=img.device)
    if batch > 0:
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=img.device)
        L.check(lib.cxm_fconv_c64(L.stream_ptr(img.device), img.data_ptr(), ker.data_ptr(),
                                  out.data_ptr(), ws.data_ptr(), ws_bytes, d))
    return out


class _FconvFn(torch.autograd.Function):
    """fourier_convolution with its reverse-mode rule.  With the output
    window starting at s per axis (k - 1) // 2 ("same") or 0 ("full"), and
    a flipped, conjugated copy of the other operand:
      grad_image[j]  = (g * conj(flip(kernel)))_full[j + K - 1 - s]
      grad_kernel[i] = (g * conj(flip(image)))_full[i + N - 1 - s]
    both full linear convolutions by the same library call, the kernel's
    summed over the output planes that share it (the reference gets these
    from JAX's FFT transpose rules)."""

    @staticmethod
    def forward(ctx, img, ker, meta):
        ctx.meta = meta
        ctx.save_for_backward(img, ker)
        return _fconv_core(img, ker, meta)

    @staticmethod
    def backward(ctx, g):
        img, ker = ctx.saved_tensors
        na, mode, ob, isp, ksp, out_sp, k_planes, cplx = ctx.meta
        sp = tuple(range(len(ob), len(ob) + na))
        starts = [(k - 1) // 2 if mode == "same" else 0 for k in ksp]
        gsp = tuple(g.shape[len(ob):])
        dt = torch.complex64 if cplx else torch.float32
        g = g.to(dt).resolve_conj().contiguous()
        gi = gk = None
        if ctx.needs_input_grad[0]:
            kf = torch.flip(ker.detach(), sp)
            kf = torch.conj_physical(kf) if cplx else kf
            full = _fconv_core(g, kf, (na, "full", ob, gsp, ksp,
                                       tuple(a + b for a, b in zip(gsp, ksp)), k_planes, cplx))
            for d, (n, k, s) in enumerate(zip(isp, ksp, starts)):
                full = full.narrow(len(ob) + d, k - 1 - s, n)
            gi = full.resolve_conj().contiguous()
        if ctx.needs_input_grad[1]:
            af = torch.flip(img.detach(), sp)
            af = torch.conj_physical(af) if cplx else af
            batch = int(np.prod(ob, dtype=np.int64))
            full = _fconv_core(g, af, (na, "full", ob, gsp, isp,
                                       tuple(a + b for a, b in zip(gsp, isp)), batch, cplx))
            for d, (n, k, s) in enumerate(zip(isp, ksp, starts)):
                full = full.narrow(len(ob) + d, n - 1 - s, k)
            # kernel plane q serves output planes q, q + k_planes, ...
            full = full.reshape((batch // k_planes, k_planes) + tuple(ksp)).sum(0)
            gk = full.reshape(ker.shape)
        return gi, gk, None


def _generator(key, device) -> torch.Generator:
    if isinstance(key, torch.Generator):
        return key
    if key is None:
        raise ValueError("Must provide a PRNG key to generate noise")
    g = torch.Generator(device=device)
    g.manual_seed(int(key))
    return g


def planewise_convolution_sum(image, kernel) -> torch.Tensor | None:
    """``fourier_convolution(image, kernel, axes=(-2, -1)).sum(0)`` for a real
    (D, H, W) ``image`` and a real (D, h, w) or (1, h, w) ``kernel`` in one
    library call (cxm_fconv_c64 with sum_batch): the sum over the planes is
    taken on the column transforms' results, so one inverse row transform
    serves every plane (a Microscope's planewise convolution summed by its
    sensor, src/chromatix/systems/microscopes.py:271-297 and
    src/chromatix/functional/sensors.py:79-80).  None where the library's
    summed path does not cover the input (complex values, an odd row
    transform length); the (H, W) result otherwise."""
    image = torch.as_tensor(image)
    L.refuse_grad("planewise_convolution_sum", image, kernel)
    L.ensure_device(image)
    kernel = torch.as_tensor(kernel, device=image.device)
    if (image.ndim != 3 or kernel.ndim != 3 or image.is_complex() or kernel.is_complex()
            or kernel.shape[0] not in (1, image.shape[0]) or image.shape[0] == 0):
        return None
    img = image.to(torch.float32).resolve_conj().contiguous()
    ker = kernel.to(torch.float32).resolve_conj().contiguous()
    d = L.FconvDesc()
    d.batch = img.shape[0]
    d.depth, d.height, d.width = 1, img.shape[1], img.shape[2]
    d.k_planes = ker.shape[0]
    d.k_depth, d.k_height, d.k_width = 1, ker.shape[1], ker.shape[2]
    d.real_input = 1
    d.sum_batch = 1
    lib = L.lib()
    ws_bytes = lib.cxm_fconv_workspace_size(d)
    if ws_bytes == 0:
        return None
    out = torch.empty(tuple(img.shape[1:]), dtype=torch.float32, device=img.device)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=img.device)
    L.check(lib.cxm_fconv_c64(L.stream_ptr(img.device), img.data_ptr(), ker.data_ptr(),
                              out.data_ptr(), ws.data_ptr(), ws_bytes, d))
    return out

def approximate_shot_noise(key, image) -> torch.Tensor:
    """max(image + sqrt(image) * N(0, 1), 0) (noise.py:9-16); ``key`` a
    torch.Generator on the image's device or an integer seed."""
    L.refuse_grad("approximate_shot_noise", image)
    x = torch.as_tensor(image)
    L.ensure_device(x)
    x = x.to(torch.float32).resolve_conj().contiguous()
    normal = torch.randn(x.shape, generator=_generator(key, x.device), device=x.device,
                         dtype=torch.float32)
    out = torch.empty_like(x)
    L.check(L.lib().cxm_shot_noise_approx_f32(L.stream_ptr(x.device), x.data_ptr(),
                                              normal.data_ptr(), out.data_ptr(), x.numel()))
    return out


def shot_noise(key, image) -> torch.Tensor:
    """Poisson samples with mean ``image``, as float32 (noise.py:38-45)."""
    L.refuse_grad("shot_noise", image)
    x = torch.as_tensor(image)
    L.ensure_device(x)
    x = x.to(torch.float32)
    return torch.poisson(x, generator=_generator(key, x.device))
