"""The C ABI (CPU): libcxm.so loads, exports every function include/cxm.h
declares, its descriptor layout matches the ctypes mirror, and argument
validation fails cleanly before any device call."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from chromatix_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cxm.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cxm_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(L.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    L.lib()
    for n in names:
        assert getattr(L.lib(), n).restype is not None or n == "cxm_timing_enable"


@pytest.mark.parametrize("cls,cname", [(L.Prop2dDesc, "cxm_prop2d_desc"),
                                       (L.OptFftDesc, "cxm_optfft_desc"),
                                       (L.Psf4fDesc, "cxm_psf4f_desc"),
                                       (L.MultisliceDesc, "cxm_multislice_desc"),
                                       (L.CztDesc, "cxm_czt_desc"),
                                       (L.YuDesc, "cxm_yu_desc"),
                                       (L.FluorDesc, "cxm_fluor_desc"),
                                       (L.MbsDesc, "cxm_mbs_desc"),
                                       (L.FconvDesc, "cxm_fconv_desc"),
                                       (L.ZoomDesc, "cxm_zoom_desc")])
def test_descriptor_layout_matches_header(tmp_path, cls, cname):
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof({cname}, {f}));' for f in fields)
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"cxm.h\"\nint main(void){\n"
        f"{body}\n  printf(\"SIZE %zu\\n\", sizeof({cname}));\n  return 0;\n}}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_off = dict(line.split() for line in out if line.strip())
    for f in fields:
        assert int(c_off[f]) == getattr(cls, f).offset, f
    assert int(c_off["SIZE"]) == ctypes.sizeof(cls)


def _desc(**kw):
    d = L.Prop2dDesc()
    d.batch, d.height, d.width = 1, 64, 64
    d.num_z = 1
    d.kind = L.CXM_H_ASM
    d.num_wv = 1
    d.dy = d.dx = 0.3
    d.wavelength, d.n = 0.532, 1.33
    z = (ctypes.c_double * 1)(10.0)
    d.z = ctypes.cast(z, ctypes.POINTER(ctypes.c_double))
    for k, v in kw.items():
        setattr(d, k, v)
    return d, z


def test_validation_errors_without_device():
    lib = L.lib()
    d, keep = _desc()
    assert lib.cxm_prop2d_workspace_size(ctypes.byref(d)) > 0
    # a length without a fused transform runs as chirp-z passes
    gen, keep2 = _desc(height=100)
    assert lib.cxm_prop2d_workspace_size(ctypes.byref(gen)) > 0
    # ... up to n + N - 1 per axis within the largest transform (10240)
    bad, keep3 = _desc(height=6000)
    assert lib.cxm_prop2d_workspace_size(ctypes.byref(bad)) == 0
    rc = lib.cxm_prop2d_c64(None, None, None, None, 0, ctypes.byref(bad))
    assert rc == L.CXM_ERR_UNSUPPORTED
    assert b"chirp-z passes" in lib.cxm_last_error()
    # and not without a workspace
    rc = lib.cxm_prop2d_c64_noalloc(None, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.byref(gen))
    assert rc == L.CXM_ERR_UNSUPPORTED
    for kw in (dict(num_z=0), dict(kind=9), dict(phase_mode=5), dict(n=-1.0), dict(pad_y=-1),
               dict(crop_x=40), dict(num_wv=8, num_z=8), dict(z_zip=3)):
        dd, k3 = _desc(**kw)
        assert lib.cxm_prop2d_c64(None, None, None, None, 0, ctypes.byref(dd)) == L.CXM_ERR_INVALID, kw
    ok, k4 = _desc()
    rc = lib.cxm_prop2d_c64(None, ctypes.c_void_p(16), ctypes.c_void_p(16), None, 0, ctypes.byref(ok))
    assert rc == L.CXM_ERR_WORKSPACE
    assert lib.cxm_intensity_c64(None, None, None, 4, 0, None) == L.CXM_ERR_INVALID
    assert lib.cxm_plane_sumsq_c64(None, None, None, 70000, 1, 1, None, None, 0) == L.CXM_ERR_INVALID


def test_fft_length_support():
    lib = L.lib()
    for n in [2, 4, 16, 64, 256, 1024, 4096, 8192, 5, 160, 1280, 2560, 5120, 10240, 3, 96, 1536,
              3072]:
        assert lib.cxm_fft_supported(n) == 1, n
    for n in [7, 9, 100, 1000, 2559, 6144, 20480, 16384]:
        assert lib.cxm_fft_supported(n) == 0, n


def test_cpu_tensors_are_refused():
    import torch

    import chromatix_amd as cx
    import chromatix_amd.functional as cf

    f = cx.ScalarField(torch.ones(8, 8, dtype=torch.complex64), 0.3, 0.0, 0.532)
    with pytest.raises(RuntimeError, match="GPU only"):
        cf.asm_propagate(f, 1.0, 1.33, 0)
    with pytest.raises(RuntimeError, match="GPU only"):
        _ = f.intensity
    _ = np


def test_validation_errors_of_the_imaging_and_born_entries():
    lib = L.lib()
    d = L.FconvDesc()
    d.batch, d.depth, d.height, d.width = 2, 1, 40, 50
    d.k_planes, d.k_depth, d.k_height, d.k_width = 1, 1, 9, 9
    d.real_input = 1
    assert lib.cxm_fconv_workspace_size(ctypes.byref(d)) > 0
    d.mode = 2
    assert lib.cxm_fconv_workspace_size(ctypes.byref(d)) == 0
    assert lib.cxm_fconv_c64(None, None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_INVALID
    d.mode, d.height = 0, 12000  # beyond the implemented transform lengths
    assert lib.cxm_fconv_c64(None, None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_UNSUPPORTED
    d.height, d.mode = 40, 1  # full mode needs its output shape
    assert lib.cxm_fconv_c64(None, None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_INVALID
    d.mode, d.sum_batch = 0, 2
    assert lib.cxm_fconv_workspace_size(ctypes.byref(d)) == 0
    assert lib.cxm_fconv_c64(None, None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_INVALID
    d.sum_batch = 1  # the batch sum: real two-axis input only
    assert lib.cxm_fconv_workspace_size(ctypes.byref(d)) > 0
    d.real_input = 0
    assert lib.cxm_fconv_workspace_size(ctypes.byref(d)) == 0
    assert lib.cxm_fconv_c64(None, None, None, None, None, 0, ctypes.byref(d)) == L.CXM_ERR_UNSUPPORTED
    d.real_input, d.sum_batch = 1, 0
    assert lib.cxm_sum_axis_f32(None, None, None, -1, 1, 1) == L.CXM_ERR_INVALID
    assert lib.cxm_sum_axis_f32(None, None, None, 0, 4, 4) == L.CXM_OK  # nothing to do
    assert lib.cxm_shot_noise_approx_f32(None, None, None, None, -1) == L.CXM_ERR_INVALID
    assert lib.cxm_shot_noise_approx_f32(None, None, None, None, 0) == L.CXM_OK
    out = (ctypes.c_double * 3)()
    assert lib.cxm_enclosing_circle_c64(None, None, 0, out, None, 0) == L.CXM_ERR_INVALID
    assert lib.cxm_enclosing_circle_workspace_size(0) == 0
    m = L.MbsDesc()
    m.nx = m.ny = m.nz = 32
    m.spacing, m.k0 = 0.133, 11.8
    assert lib.cxm_mbs_bicgstab_state_size(ctypes.byref(m)) > 0
    m.eps_tensor = 3
    assert lib.cxm_mbs_bicgstab_state_size(ctypes.byref(m)) == 0
    assert lib.cxm_mbs_workspace_size(ctypes.byref(m)) == 0
