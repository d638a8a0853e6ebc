"""The allocation-free propagation entry point (cxm_prop2d_c64_noalloc)
against cxm_prop2d_c64 with a caller workspace, through the C ABI.

Two calls in a row on one stream, the second with a smaller input: the
static scratch array keeps the first call's contents, and the second call
must not depend on them.  Bar: relative L2 <= 1e-6 against the workspace
entry point (the same kernels; chunks of fewer planes may pick another
column-pass launch shape, whose rounding differs at the 1e-7 level), and
<= 1e-5 against the oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

from chromatix_amd import _lib as L
from oracle import propagation as op
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
LAM, NM = 0.532, 1.33


def _desc(batch, h, w, z, pad=0, out_kind=L.CXM_OUT_COMPLEX):
    d = L.Prop2dDesc()
    d.batch, d.height, d.width = batch, h, w
    d.pad_y = d.pad_x = pad
    d.crop_y = d.crop_x = pad
    d.num_z = 1
    d.kind = L.CXM_H_ASM
    d.phase_mode = L.CXM_PHASE_REF32
    d.out_kind = out_kind
    d.dy = d.dx = 0.3
    d.wavelength = LAM
    d.n = NM
    zs = L.doubles([z])
    d.z = ctypes.cast(zs, ctypes.POINTER(ctypes.c_double))
    return d, zs


def _inputs(batch, h, w, seed):
    rng = np.random.default_rng(seed)
    u = np.exp(1j * rng.uniform(0, 2 * np.pi, (batch, h, w))).astype(np.complex64)
    return u, torch.as_tensor(u).cuda()


def _with_workspace(ut, d):
    lib = L.lib()
    out = torch.empty_like(ut)
    ws = torch.empty(lib.cxm_prop2d_workspace_size(ctypes.byref(d)), dtype=torch.uint8, device="cuda")
    L.check(lib.cxm_prop2d_c64(L.stream_ptr(), L.dptr(ut), L.dptr(out), L.dptr(ws), ws.numel(),
                               ctypes.byref(d)))
    return out


def test_noalloc_matches_workspace_entry_point():
    lib = L.lib()
    assert lib.cxm_noalloc_scratch_bytes() == 64 << 20
    # first call: 12 planes of 1024^2 (two chunks of the 64 MiB scratch),
    # second call: 3 planes of 512^2 with padding (a smaller input)
    for batch, h, pad, z, seed in ((12, 1024, 0, 100.0, 0), (3, 512, 64, -35.0, 1)):
        print(f"noalloc: batch {batch} {h}^2 pad {pad} z {z}", flush=True)
        u, ut = _inputs(batch, h, h, seed)
        L.ensure_device(ut)
        d, _zs = _desc(batch, h, h, z, pad)
        out = torch.empty_like(ut)
        L.check(lib.cxm_prop2d_c64_noalloc(L.stream_ptr(), L.dptr(ut), L.dptr(out), ctypes.byref(d)))
        ref_ws = _with_workspace(ut, d)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        err_ws = rel_l2(got, ref_ws.cpu().numpy())
        ref = op.propagate(u[:2], (0.3, 0.3), LAM, z, NM, pad, precision="ref32", fft_precision="fp64")
        err_or = rel_l2(got[:2], ref)
        print(f"  vs workspace entry {err_ws:.2e}, vs oracle {err_or:.2e}", flush=True)
        assert err_ws <= 1e-6
        assert err_or <= 1e-5


def test_noalloc_rejects_grids_beyond_the_scratch():
    lib = L.lib()
    u = torch.zeros((1, 8, 8), dtype=torch.complex64, device="cuda")
    L.ensure_device(u)
    d, _zs = _desc(1, 4096, 4096, 10.0)
    rc = lib.cxm_prop2d_c64_noalloc(L.stream_ptr(), L.dptr(u), L.dptr(u), ctypes.byref(d))
    assert rc == L.CXM_ERR_UNSUPPORTED
    assert b"static scratch" in lib.cxm_last_error()
