"""GPU parity of the matrix-DFT zoom on MFMA (cxm_zoom_dft_mfma_c64,
cxm_cgemm_mfma_c64) against FP64: the complex GEMM against numpy's product
in complex128, the zoom against the CZT's defining sum (oracle.czt_direct)
and against the Bluestein passes (cxm_czt_axis_c64).  Bar: relative L2
<= 1e-5."""
import ctypes

import numpy as np
import pytest
import torch

from chromatix_amd import _lib as L
from chromatix_amd.utils.czt import czt_axis, zoom_dft
from oracle import czt as oz
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("M,N,K,batch,shared_b", [
    (16, 16, 4, 1, False), (64, 64, 16, 1, False), (37, 50, 23, 1, False),
    (130, 70, 300, 3, True), (1, 200, 1, 2, False), (257, 65, 1031, 2, False),
])
def test_cgemm_matches_numpy(M, N, K, batch, shared_b):
    rng = np.random.default_rng(M + N + K)
    a = _c(rng, (batch, M, K))
    b = _c(rng, (1 if shared_b else batch, K, N))
    ta, tb = torch.as_tensor(a).cuda(), torch.as_tensor(b).cuda()
    L.ensure_device(ta)
    tc = torch.empty((batch, M, N), dtype=torch.complex64, device="cuda")
    print(f"cgemm {M}x{N}x{K} batch {batch}", flush=True)
    L.check(L.lib().cxm_cgemm_mfma_c64(L.stream_ptr(), ta.data_ptr(), tb.data_ptr(), tc.data_ptr(),
                                       M, N, K, K, N, N, M * K, 0 if shared_b else K * N, M * N,
                                       batch))
    ref = np.matmul(a.astype(np.complex128), b.astype(np.complex128))
    assert rel_l2(tc.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape,m,kr", [
    ((256, 192), (40, 24), (-1.3, 0.9)),
    ((100, 64), (64, 100), (-0.5, 2.5)),
    ((2048, 2048), (64, 64), (-0.8, 0.8)),
])
def test_zoomed_fft_window_matches_defining_sum(shape, m, kr):
    """zoomed_fft parameters (czt.py:82-124, include_end): a = e^{-i k0},
    w = e^{i (k1 - k0) / (m - 1)} on both axes."""
    rng = np.random.default_rng(sum(shape))
    x = _c(rng, (2,) + shape)
    a = np.exp(-1j * kr[0])
    w = tuple(np.exp(1j * (kr[1] - kr[0]) / (mm - 1)) for mm in m)
    print(f"zoom {shape} -> {m}", flush=True)
    got = zoom_dft(torch.as_tensor(x).cuda(), m, (a, a), w).cpu().numpy()
    ref = oz.czt_direct(oz.czt_direct(x, m[0], a, w[0], axis=-2), m[1], a, w[1], axis=-1)
    assert got.shape == (2,) + tuple(m)
    assert rel_l2(got, ref) < TOL


def test_zoom_with_modulation_matches_bluestein_passes():
    """The CZT branch's per-axis modulation w^{-N k} a^N and scale folded
    into the matrices equals two cxm_czt_axis_c64 passes with mod_n."""
    rng = np.random.default_rng(3)
    x = torch.as_tensor(_c(rng, (3, 320, 256))).cuda()
    m = (48, 33)
    a = (np.exp(-0.37j), np.exp(0.21j))
    w = (np.exp(0.011j), np.exp(-0.017j))
    nmod = ((m[0] - 1) // 2, (m[1] - 1) // 2)
    scale = (0.7, 1.9)
    got = zoom_dft(x, m, a, w, mod_n=nmod, scale=scale).cpu().numpy()
    blue = czt_axis(czt_axis(x, m[0], a[0], w[0], 0, mod_n=nmod[0], scale=scale[0]),
                    m[1], a[1], w[1], 1, mod_n=nmod[1], scale=scale[1]).cpu().numpy()
    xn = x.cpu().numpy()
    ref = oz.czt_direct(xn, m[0], a[0], w[0], axis=-2)
    ref = ref * (w[0] ** (-nmod[0] * np.arange(m[0])) * a[0] ** nmod[0] * scale[0])[:, None]
    ref = oz.czt_direct(ref, m[1], a[1], w[1], axis=-1)
    ref = ref * (w[1] ** (-nmod[1] * np.arange(m[1])) * a[1] ** nmod[1] * scale[1])[None, :]
    assert rel_l2(got, ref) < TOL
    assert rel_l2(blue, ref) < TOL
    assert rel_l2(got, blue) < 2 * TOL


def test_zoom_validation():
    d = L.ZoomDesc()
    d.batch, d.height, d.width, d.m_y, d.m_x = 1, 8, 8, 0, 4
    assert L.lib().cxm_zoom_dft_mfma_c64(None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_INVALID
