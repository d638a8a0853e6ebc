"""CPU build of the device phase arithmetic vs the oracle.

cxm_hfunc.h (on-the-fly transfer function), cxm_ops.h (pointwise operators)
and cxm_hparams.cpp (host parameter builder) are compiled for the host with
g++ (tests/cpu_check shims) and evaluated over whole grids.  ref32 must
reproduce the oracle's reference-order float32 phase (difference = the
final sin/cos rounding only); accurate must match FP64.
"""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from chromatix_amd import _lib as L
from oracle import propagation as op

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="session")
def chk(tmp_path_factory):
    out = tmp_path_factory.mktemp("cpu_check") / "libchk.so"
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-DCXM_CPU_CHECK", "-Wno-unknown-pragmas",
           "-I", str(ROOT / "tests" / "cpu_check"), "-I", str(ROOT / "chromatix_amd" / "csrc"),
           "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpu_check" / "hfunc_check.cpp"),
           str(ROOT / "chromatix_amd" / "csrc" / "cxm_hparams.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True)
    lib = ctypes.CDLL(str(out))
    lib.chk_eval_h.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.chk_build_hparams.argtypes = [ctypes.POINTER(L.Prop2dDesc), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_int]
    return lib


def _desc(kind, lam, n, z, dy, dx, kykx=(0.0, 0.0), shift=(0.0, 0.0), bandlimit=0, rem=0):
    d = L.Prop2dDesc()
    d.kind, d.wavelength, d.n, d.dy, d.dx = kind, lam, n, dy, dx
    d.ky, d.kx = kykx
    d.shift_y, d.shift_x = shift
    d.bandlimit, d.remove_evanescent = bandlimit, rem
    d.num_wv, d.num_z = 1, 1
    zz = (ctypes.c_double * 1)(z)
    d.z = ctypes.cast(zz, ctypes.POINTER(ctypes.c_double))
    return d, zz


def device_H(chk, kind, shape, dx, lam, z, n, mode, **kw):
    ny, nx = shape
    d, keep = _desc(kind, lam, n, z, dx[0], dx[1], kw.get("kykx", (0.0, 0.0)),
                    kw.get("shift_yx", (0.0, 0.0)), int(kw.get("bandlimit", False)),
                    int(kw.get("remove_evanescent", False)))
    hp = ctypes.create_string_buffer(chk.chk_hparam_size())
    assert chk.chk_build_hparams(ctypes.byref(d), ny, nx, hp, 1) == 0
    out = np.empty((ny, nx, 2), np.float32)
    chk.chk_eval_h(hp, ny, nx, kind, mode, int(kw.get("remove_evanescent", False)),
                   int(kw.get("bandlimit", False)), out.ctypes.data)
    return out[..., 0] + 1j * out[..., 1]


CASES = [
    (L.CXM_H_ASM, (1024, 1024), (0.3, 0.3), 0.532, 100.0, 1.33, {}),  # C1 grid
    (L.CXM_H_ASM, (160, 128), (0.15, 0.15), 0.532, 800.0, 1.33, {"kykx": (0.4, -0.7)}),
    (L.CXM_H_ASM, (160, 128), (0.15, 0.15), 0.532, 800.0, 1.33, {"shift_yx": (3.0, -5.5)}),
    (L.CXM_H_ASM, (160, 128), (0.15, 0.15), 0.532, 800.0, 1.33, {"bandlimit": True}),
    (L.CXM_H_ASM, (160, 128), (0.15, 0.15), 0.532, 800.0, 1.33, {"remove_evanescent": True}),
    (L.CXM_H_ASM, (128, 128), (0.3, 0.3), 0.532, -40.0, 1.33, {}),
    (L.CXM_H_ASM, (64, 32), (0.1, 0.1), 0.532, 2.0, 1.33, {}),  # evanescent
    (L.CXM_H_ASM, (1280, 1280), (40 / 256, 40 / 256), 0.532, 100.0, 1.33, {}),
    (L.CXM_H_TRANSFER, (1024, 1024), (0.3, 0.3), 0.532, 100.0, 1.33, {}),
    (L.CXM_H_TRANSFER, (160, 128), (0.15, 0.2), 0.532, -250.0, 1.33, {"kykx": (0.2, 0.1)}),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1][0]}x{c[1][1]}-{i}" for i, c in enumerate(CASES)])
def test_device_H_matches_oracle(chk, case):
    kind, shape, dx, lam, z, n, kw = case
    fn = op.asm_propagator if kind == L.CXM_H_ASM else op.transfer_propagator
    kw_o = dict(kw) if kind == L.CXM_H_ASM else {k: v for k, v in kw.items() if k == "kykx"}
    # bins on the evanescent boundary (|kernel| < 1e-9) are ill-conditioned:
    # the sign of the kernel there is rounding noise for any implementation
    fy = np.fft.fftfreq(shape[0], dx[0])[:, None] - kw.get("kykx", (0, 0))[0]
    fx = np.fft.fftfreq(shape[1], dx[1])[None, :] - kw.get("kykx", (0, 0))[1]
    well = np.abs(1 - (lam / n) ** 2 * (fy**2 + fx**2)) > 1e-9
    for mode, prec, tol in ((L.CXM_PHASE_REF32, "ref32", 3e-7), (L.CXM_PHASE_ACCURATE, "fp64", 1e-6)):
        Hd = device_H(chk, kind, shape, dx, lam, z, n, mode, **kw)
        Ho = fn(shape, dx, lam, z, n, precision=prec, **kw_o)
        err = np.max(np.abs(Hd - Ho)[well])
        assert err < tol, (mode, err)
        assert np.linalg.norm(Hd - Ho) / np.linalg.norm(Ho) < 1e-5
