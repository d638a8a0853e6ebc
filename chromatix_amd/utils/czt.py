"""Chirp-z transforms on libcxm (mirrors src/chromatix/utils/czt.py).

``czt`` runs one Bluestein pass along an axis (cxm_czt_axis_c64): chirps in
FP64, FFTs of the next supported length >= n + m - 1, complex64 data.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib as L

__all__ = ["czt", "cztn", "zoomed_fft", "czt_axis", "scaled_conv_axis"]


def _pairs(v, count: int):
    a = np.atleast_1d(np.asarray(v, dtype=np.complex128)).reshape(-1)
    if a.size == 1:
        a = np.repeat(a, count)
    if a.size != count:
        raise ValueError("a / w must be scalars or one value per parameter set")
    out = np.empty(2 * count)
    out[0::2], out[1::2] = a.real, a.imag
    return out


def _czt_desc(shape, m, axis, num_param, mod_n, scale, ab, wb):
    H, W = shape[-2:]
    d = L.CztDesc()
    d.batch = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    d.height, d.width = H, W
    d.axis = int(axis)
    d.m = int(m)
    d.num_param = int(num_param)
    d.mod_n = int(mod_n)
    d.scale = float(scale)
    d.a = ctypes.cast(ab, ctypes.POINTER(ctypes.c_double))
    d.w = ctypes.cast(wb, ctypes.POINTER(ctypes.c_double))
    return d


def _cgemm(dev, A, B, C, M, N, K, lda, ldb, ldc, sa, sb, sc, batch):
    """C[z] = A[z] B[z] on the matrix cores (pointers as integers)."""
    L.check(L.lib().cxm_cgemm_mfma_c64(L.stream_ptr(dev), A, B, C, int(M), int(N), int(K), int(lda),
                                       int(ldb), int(ldc), int(sa), int(sb), int(sc), int(batch)))


def _czt_matrix(d, param, rows, cols, ld_r, ld_c, dev):
    """The conjugated matrix of parameter set ``param`` of a CZT pass,
    element (k, j) at [k * ld_r + j * ld_c] of a (rows, cols) tensor."""
    out = torch.empty((rows, cols), dtype=torch.complex64, device=dev)
    L.check(L.lib().cxm_czt_matrix_c64(L.stream_ptr(dev), ctypes.byref(d), int(param), out.data_ptr(),
                                       int(ld_r), int(ld_c), 1))
    return out


def _axis_adjoint(gy, x_shape, axis, n, m, mats):
    """x_p = A_q^H y_p along ``axis`` of the planes of gy, q = p % len(mats);
    mats[q] is the conjugate transpose (axis 0, n x m) or the conjugate
    (axis 1, m x n) of the pass's matrix, row-major."""
    dev = gy.device
    H, W = x_shape[-2:]
    B = int(np.prod(x_shape[:-2])) if len(x_shape) > 2 else 1
    gx = torch.empty(tuple(x_shape), dtype=torch.complex64, device=dev)
    P = len(mats)
    g_plane = (m * W) if axis == 0 else (H * m)
    x_plane = H * W
    for q, A in enumerate(mats):
        cnt = (B - q + P - 1) // P
        if cnt <= 0:
            continue
        if axis == 0:  # (n x m) (m x W)
            _cgemm(dev, A.data_ptr(), gy.data_ptr() + 8 * q * g_plane, gx.data_ptr() + 8 * q * x_plane,
                   n, W, m, m, W, W, 0, P * g_plane, P * x_plane, cnt)
        elif P == 1:  # all rows of all planes at once: (B H x m) (m x n)
            _cgemm(dev, gy.data_ptr(), A.data_ptr(), gx.data_ptr(), B * H, n, m, m, n, n, 0, 0, 0, 1)
        else:  # (H x m) (m x n) per plane
            _cgemm(dev, gy.data_ptr() + 8 * q * g_plane, A.data_ptr(), gx.data_ptr() + 8 * q * x_plane,
                   H, n, m, m, n, n, P * g_plane, 0, P * x_plane, cnt)
    return gx


def _czt_axis_raw(x, m, axis, num_param, mod_n, scale, ab, wb):
    L.ensure_device(x)
    x = x.to(torch.complex64).resolve_conj().contiguous()
    H, W = x.shape[-2:]
    d = _czt_desc(x.shape, m, axis, num_param, mod_n, scale, ab, wb)
    out_hw = (int(m), W) if axis == 0 else (H, int(m))
    out = torch.empty(tuple(x.shape[:-2]) + out_hw, dtype=torch.complex64, device=x.device)
    nbytes = L.lib().cxm_czt_axis_workspace_size(ctypes.byref(d))
    if nbytes == 0:
        L.check(L.lib().cxm_czt_axis_c64(None, None, None, None, 0, ctypes.byref(d)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    L.check(L.lib().cxm_czt_axis_c64(L.stream_ptr(x.device), x.data_ptr(), out.data_ptr(),
                                     ws.data_ptr(), nbytes, ctypes.byref(d)))
    return out


class _CztAxisFn(torch.autograd.Function):
    """cxm_czt_axis_c64 with its reverse-mode rule: the product with the
    conjugate transpose of the pass's matrix (cxm_czt_matrix_c64 +
    cxm_cgemm_mfma_c64), the rule JAX derives for czt (czt.py:7-52)."""

    @staticmethod
    def forward(ctx, x, m, axis, num_param, mod_n, scale, ab, wb):
        ctx.args = (m, axis, num_param, mod_n, scale, ab, wb)
        ctx.x_shape = tuple(x.shape)
        return _czt_axis_raw(x, m, axis, num_param, mod_n, scale, ab, wb)

    @staticmethod
    def backward(ctx, gy):
        m, axis, P, mod_n, scale, ab, wb = ctx.args
        gy = gy.to(torch.complex64).resolve_conj().contiguous()
        d = _czt_desc(ctx.x_shape, m, axis, P, mod_n, scale, ab, wb)
        n = ctx.x_shape[-2] if axis == 0 else ctx.x_shape[-1]
        if axis == 0:  # conjugate transpose, n x m
            mats = [_czt_matrix(d, q, n, m, 1, m, gy.device) for q in range(P)]
        else:  # conjugate, m x n
            mats = [_czt_matrix(d, q, m, n, n, 1, gy.device) for q in range(P)]
        return _axis_adjoint(gy, ctx.x_shape, axis, n, m, mats), None, None, None, None, None, None, None


def czt_axis(x: torch.Tensor, m: int, a, w, axis: int, *, num_param: int = 1, mod_n: int = -1,
             scale: float = 1.0) -> torch.Tensor:
    """CZT of the planes of ``x`` (..., H, W) along axis 0 (H) or 1 (W) of
    the last two; plane p uses parameter set p % num_param; with mod_n = N
    the output is also multiplied by w^(-N k) a^N (the modulation of the CZT
    branch of kernel_propagate) and by ``scale``.  Differentiable in x."""
    x = torch.as_tensor(x)
    ab, wb = L.doubles(_pairs(a, num_param)), L.doubles(_pairs(w, num_param))
    return _CztAxisFn.apply(x, int(m), int(axis), int(num_param), int(mod_n), float(scale), ab, wb)


def czt(x, m: int, a, w, axis: int = -1) -> torch.Tensor:
    """Chirp-z transform along ``axis``: y_k = sum_j x_j a^-j w^(j k),
    k < m (czt.py:7-52, Bluestein)."""
    x = torch.as_tensor(x)
    ax = axis % x.ndim
    if x.ndim >= 2 and ax == x.ndim - 2:
        # the column transform of the planes of x: no transpose
        H, W = x.shape[-2:]
        y = czt_axis(x.reshape((-1, H, W)), m, a, w, 0)
        return y.reshape(tuple(x.shape[:-2]) + (int(m), W))
    xm = x.movedim(ax, -1)
    lead = xm.shape[:-1]
    y = czt_axis(xm.reshape(1, -1, xm.shape[-1]), m, a, w, 1)
    return y.reshape(tuple(lead) + (int(m),)).movedim(-1, ax)


# Output windows up to this length per axis take the matrix-DFT path (two
# complex GEMMs on the matrix cores) instead of two Bluestein passes.
# Measured at n = 2048: 16 planes to 64^2 0.34 ms against 1.06 ms, 128^2
# 0.66 against 1.13, 192^2 1.03 against 1.15, 256^2 1.31 against 1.21
# (Bluestein faster from ~230 on); one plane to 64^2 0.059 against 0.147 ms.
ZOOM_MFMA_MAX = 192


def use_zoom_mfma(m) -> bool:
    """Whether an (m_y, m_x) window goes to the matrix cores.  The
    environment variable CXM_ZOOM_PATH = "mfma" / "bluestein" forces a path
    (tests and the crossover measurement)."""
    import os

    force = os.environ.get("CXM_ZOOM_PATH", "")
    if force == "mfma":
        return True
    if force == "bluestein":
        return False
    return max(int(m[0]), int(m[1])) <= ZOOM_MFMA_MAX


def _zoom_raw(x, m, a, w, mod_n, scale):
    L.ensure_device(x)
    x = x.to(torch.complex64).resolve_conj().contiguous()
    H, W = x.shape[-2:]
    d = L.ZoomDesc()
    d.batch = int(np.prod(x.shape[:-2])) if x.ndim > 2 else 1
    d.height, d.width = H, W
    d.m_y, d.m_x = int(m[0]), int(m[1])
    d.mod_n_y, d.mod_n_x = int(mod_n[0]), int(mod_n[1])
    for name, v in (("a_y", a[0]), ("w_y", w[0]), ("a_x", a[1]), ("w_x", w[1])):
        c = complex(v)
        getattr(d, name)[0], getattr(d, name)[1] = c.real, c.imag
    d.scale_y, d.scale_x = float(scale[0]), float(scale[1])
    out = torch.empty(tuple(x.shape[:-2]) + (d.m_y, d.m_x), dtype=torch.complex64, device=x.device)
    nbytes = L.lib().cxm_zoom_dft_workspace_size(ctypes.byref(d))
    if nbytes == 0:
        L.check(L.lib().cxm_zoom_dft_mfma_c64(None, None, None, None, 0, ctypes.byref(d)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    L.check(L.lib().cxm_zoom_dft_mfma_c64(L.stream_ptr(x.device), x.data_ptr(), out.data_ptr(),
                                          ws.data_ptr(), nbytes, ctypes.byref(d)))
    return out


class _ZoomFn(torch.autograd.Function):
    """cxm_zoom_dft_mfma_c64 with its reverse-mode rule X = A_y^H Y conj(A_x)
    (two complex GEMMs on the matrix cores)."""

    @staticmethod
    def forward(ctx, x, m, a, w, mod_n, scale):
        ctx.args = (m, a, w, mod_n, scale)
        ctx.x_shape = tuple(x.shape)
        return _zoom_raw(x, m, a, w, mod_n, scale)

    @staticmethod
    def backward(ctx, gy):
        m, a, w, mod_n, scale = ctx.args
        gy = gy.to(torch.complex64).resolve_conj().contiguous()
        dev = gy.device
        H, W = ctx.x_shape[-2:]
        B = int(np.prod(ctx.x_shape[:-2])) if len(ctx.x_shape) > 2 else 1
        my, mx = int(m[0]), int(m[1])
        keep = []

        def desc(axis, mm, av, wv, mn, sc):
            ab, wb = L.doubles(_pairs(av, 1)), L.doubles(_pairs(wv, 1))
            keep.append((ab, wb))
            return _czt_desc((H, W), mm, axis, 1, mn, sc, ab, wb)

        ayh = _czt_matrix(desc(0, my, a[0], w[0], mod_n[0], scale[0]), 0, H, my, 1, my, dev)
        axc = _czt_matrix(desc(1, mx, a[1], w[1], mod_n[1], scale[1]), 0, mx, W, W, 1, dev)
        t = torch.empty((B, my, W), dtype=torch.complex64, device=dev)
        # T = Y conj(A_x): all planes' rows at once, (B m_y x m_x) (m_x x W)
        _cgemm(dev, gy.data_ptr(), axc.data_ptr(), t.data_ptr(), B * my, W, mx, mx, W, W, 0, 0, 0, 1)
        gx = torch.empty(ctx.x_shape, dtype=torch.complex64, device=dev)
        # X_p = A_y^H T_p, (H x m_y) (m_y x W)
        _cgemm(dev, ayh.data_ptr(), t.data_ptr(), gx.data_ptr(), H, W, my, my, W, W, 0, my * W, H * W, B)
        return gx, None, None, None, None, None


def zoom_dft(x, m, a, w, mod_n=(-1, -1), scale=(1.0, 1.0)) -> torch.Tensor:
    """CZT of the planes of ``x`` (..., H, W) along both of its last two axes
    as two complex GEMMs on the matrix cores (cxm_zoom_dft_mfma_c64):
    y[ky, kx] = sum x[jy, jx] A_y[ky, jy] A_x[kx, jx], A_d[k, j] = a_d^-j
    w_d^(j k) (times w_d^(-N k) a_d^N scale_d with mod_n_d = N >= 0), the
    values of ``czt_axis`` along each axis.  ``m``, ``a``, ``w``, ``mod_n``
    and ``scale`` are (y, x) pairs.  Differentiable in x."""
    x = torch.as_tensor(x)
    m = (int(m[0]), int(m[1]))
    a = (complex(a[0]), complex(a[1]))
    w = (complex(w[0]), complex(w[1]))
    return _ZoomFn.apply(x, m, a, w, (int(mod_n[0]), int(mod_n[1])), (float(scale[0]), float(scale[1])))


def cztn(x, m, a, w, axes=(-2, -1)) -> torch.Tensor:
    """CZT along several axes in turn (czt.py:55-79); over the last two axes
    of a window up to ZOOM_MFMA_MAX per axis, as two complex GEMMs on the
    matrix cores (zoom_dft)."""
    x = torch.as_tensor(x)
    if len(axes) == 2 and x.ndim >= 2:
        ax = tuple(a_ % x.ndim for a_ in axes)
        if ax == (x.ndim - 2, x.ndim - 1) and use_zoom_mfma(m):
            return zoom_dft(x, m, a, w)
    for d, ax in enumerate(axes):
        x = czt(x, m[d], a[d], w[d], ax)
    return x


def zoomed_fft(x, k_start: float, k_end: float, output_shape, axes=(-2, -1),
               include_end: bool = True) -> torch.Tensor:
    """Zoomed FFT between angular frequencies k_start and k_end
    (czt.py:82-124)."""
    renorm = tuple(m - 1 for m in output_shape) if include_end else tuple(output_shape)
    w = tuple(np.exp(1j * (k_end - k_start) / n) for n in renorm)
    a = np.exp(-1j * k_start)
    return cztn(x, tuple(output_shape), tuple(a for _ in output_shape), w, axes)


def _yu_desc(shape, axis, alpha, df, out_dx, x0):
    H, W = shape[-2:]
    d = L.YuDesc()
    d.batch = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    d.height, d.width = H, W
    d.axis = int(axis)
    d.alpha, d.df, d.out_dx, d.x0 = float(alpha), float(df), float(out_dx), float(x0)
    return d


def _yu_raw(x, axis, alpha, df, out_dx, x0):
    L.ensure_device(x)
    x = x.to(torch.complex64).resolve_conj().contiguous()
    d = _yu_desc(x.shape, axis, alpha, df, out_dx, x0)
    out = torch.empty_like(x)
    nbytes = L.lib().cxm_yu_axis_workspace_size(ctypes.byref(d))
    if nbytes == 0:
        L.check(L.lib().cxm_yu_axis_c64(None, None, None, None, 0, ctypes.byref(d)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    L.check(L.lib().cxm_yu_axis_c64(L.stream_ptr(x.device), x.data_ptr(), out.data_ptr(),
                                    ws.data_ptr(), nbytes, ctypes.byref(d)))
    return out


class _YuAxisFn(torch.autograd.Function):
    """cxm_yu_axis_c64 with its reverse-mode rule: the product with the
    conjugate transpose of the pass's matrix (cxm_yu_matrix_c64 +
    cxm_cgemm_mfma_c64)."""

    @staticmethod
    def forward(ctx, x, axis, alpha, df, out_dx, x0):
        ctx.args = (axis, alpha, df, out_dx, x0)
        ctx.x_shape = tuple(x.shape)
        return _yu_raw(x, axis, alpha, df, out_dx, x0)

    @staticmethod
    def backward(ctx, gy):
        axis = ctx.args[0]
        gy = gy.to(torch.complex64).resolve_conj().contiguous()
        d = _yu_desc(ctx.x_shape, *ctx.args)
        n = ctx.x_shape[-2] if axis == 0 else ctx.x_shape[-1]
        mat = torch.empty((n, n), dtype=torch.complex64, device=gy.device)
        ld_r, ld_c = (1, n) if axis == 0 else (n, 1)  # conjugate transpose / conjugate
        L.check(L.lib().cxm_yu_matrix_c64(L.stream_ptr(gy.device), ctypes.byref(d), mat.data_ptr(),
                                          ld_r, ld_c, 1))
        return _axis_adjoint(gy, ctx.x_shape, axis, n, n, [mat]), None, None, None, None, None


def scaled_conv_axis(x: torch.Tensor, axis: int, alpha: float, df: float, out_dx: float,
                     x0: float) -> torch.Tensor:
    """One axis of kernel_propagate's scaled convolution (Yu et al. 2012,
    propagation.py:444-458) on the planes of ``x`` (..., H, W) along axis 0
    (H) or 1 (W) of the last two (cxm_yu_axis_c64):
    y_j = out_dx e^{i pi x_j^2/alpha} sum_k x_k (1/alpha) e^{i pi alpha f_k^2}
    kern[j + (n-1)//2 - k], kern[t] = e^{-i pi alpha f_t^2} on [0, n),
    f_k = (k - n//2) df, x_j = x0 + j out_dx.  Differentiable in x."""
    x = torch.as_tensor(x)
    return _YuAxisFn.apply(x, int(axis), float(alpha), float(df), float(out_dx), float(x0))
