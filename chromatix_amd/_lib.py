"""ctypes binding of libcxm.so (the C ABI declared in include/cxm.h).

torch is imported first so that libcxm binds to the HIP runtime torch has
already loaded: device buffers and streams are shared with torch.  There is
no CPU fallback: if the library is missing or no GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch  # noqa: F401  (must precede the CDLL load)

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libcxm.so"

CXM_OK = 0
CXM_ERR_INVALID = 1
CXM_ERR_UNSUPPORTED = 2
CXM_ERR_HIP = 3
CXM_ERR_WORKSPACE = 4

CXM_H_NONE = 0
CXM_H_TRANSFER = 1
CXM_H_ASM = 2
CXM_H_ARRAY = 3

CXM_PHASE_REF32 = 0
CXM_PHASE_ACCURATE = 1

CXM_OUT_COMPLEX = 0
CXM_OUT_INTENSITY = 1

# cxm_set_option keys (launch-shape tuning and the shared H table)
CXM_OPT_COL_VARIANT = 1
CXM_OPT_CHUNK_PLANES = 2
CXM_OPT_ROW_VARIANT = 3
CXM_OPT_PIPE_ROW_VARIANT = 4
CXM_OPT_H_TABLE = 5
CXM_OPT_COL_ORDER = 6
CXM_OPT_TILES = 7
CXM_OPT_ROW_LOOP = 8
CXM_OPT_MBS_XGREEN = 9
CXM_OPT_NT_IO = 10
CXM_OPT_FCONV_PAIR = 11
CXM_OPT_COL_PERSIST = 12
CXM_OPT_COL_PERSIST_E = 13
CXM_OPT_COL_PERSIST_MODE = 14
CXM_OPT_GRAD_GROUPS = 15
CXM_OPT_MBS_ZFUSE = 16
CXM_OPT_MS_PASSES = 17
CXM_OPT_MS_ROW_E = 18
CXM_OPT_PSF_PASSES = 19
CXM_OPT_PSF_GRAD_WAVES = 20
CXM_OPT_CZT_PASSES = 21
CXM_OPT_FL_BWD_CHUNK = 22
CXM_OPT_FL_BWD_BATCH = 23
CXM_OPT_FL_ROW_E = 24
CXM_OPT_FL_FACTOR_BLOCK = 25
CXM_H_TABLE_MIN_PLANES = 4

PHASE_MODES = {"ref32": CXM_PHASE_REF32, "accurate": CXM_PHASE_ACCURATE}


class CxmError(RuntimeError):
    """Raised when a libcxm entry point returns a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libcxm error {code}: {message}")
        self.code = code


class Prop2dDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("pad_y", ctypes.c_int32),
        ("pad_x", ctypes.c_int32),
        ("crop_y", ctypes.c_int32),
        ("crop_x", ctypes.c_int32),
        ("num_z", ctypes.c_int32),
        ("z_zip", ctypes.c_int32),
        ("kind", ctypes.c_int32),
        ("phase_mode", ctypes.c_int32),
        ("out_kind", ctypes.c_int32),
        ("remove_evanescent", ctypes.c_int32),
        ("bandlimit", ctypes.c_int32),
        ("num_wv", ctypes.c_int32),
        ("dy", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("n", ctypes.c_double),
        ("ky", ctypes.c_double),
        ("kx", ctypes.c_double),
        ("shift_y", ctypes.c_double),
        ("shift_x", ctypes.c_double),
        ("z", ctypes.POINTER(ctypes.c_double)),
        ("wavelengths", ctypes.POINTER(ctypes.c_double)),
        ("spacing", ctypes.POINTER(ctypes.c_double)),
        ("propagator", ctypes.c_void_p),
    ]


CXM_OFFT_OPTICAL = 0
CXM_OFFT_TRANSFORM = 1
CXM_OFFT_RAW_FWD = 2
CXM_OFFT_RAW_INV = 3


class OptFftDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("pad_y", ctypes.c_int32),
        ("pad_x", ctypes.c_int32),
        ("crop_y", ctypes.c_int32),
        ("crop_x", ctypes.c_int32),
        ("num_wv", ctypes.c_int32),
        ("phase_mode", ctypes.c_int32),
        ("out_kind", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("shift", ctypes.c_int32),
        ("skip_initial_phase", ctypes.c_int32),
        ("skip_final_phase", ctypes.c_int32),
        ("pupil", ctypes.c_int32),
        ("adjoint", ctypes.c_int32),
        ("z", ctypes.c_double),
        ("n", ctypes.c_double),
        ("pupil_d", ctypes.c_double),
        ("central_wavelength", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("dy", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("oy", ctypes.c_double),
        ("ox", ctypes.c_double),
        ("wavelengths", ctypes.POINTER(ctypes.c_double)),
        ("spacing", ctypes.POINTER(ctypes.c_double)),
        ("origin", ctypes.POINTER(ctypes.c_double)),
        ("spec_mul", ctypes.c_void_p),
    ]


CXM_GRAD_UPSTREAM = 0
CXM_GRAD_MSE = 1


class Psf4fDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("num_z", ctypes.c_int32),
        ("phase_mode", ctypes.c_int32),
        ("out_kind", ctypes.c_int32),
        ("normalize", ctypes.c_int32),
        ("grad_kind", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dy", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("n", ctypes.c_double),
        ("f", ctypes.c_double),
        ("NA", ctypes.c_double),
        ("f_tube", ctypes.c_double),
        ("power", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("offset_y", ctypes.c_double),
        ("offset_x", ctypes.c_double),
        ("z", ctypes.POINTER(ctypes.c_double)),
        ("phase_scale", ctypes.c_double),
    ]


CXM_GRAD_COMPLEX = 2


class ResampleDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int64),
        ("in_h", ctypes.c_int32),
        ("in_w", ctypes.c_int32),
        ("out_h", ctypes.c_int32),
        ("out_w", ctypes.c_int32),
        ("ky", ctypes.c_int32),
        ("kx", ctypes.c_int32),
        ("mask_on_output", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("y_start", ctypes.c_void_p),
        ("wy", ctypes.c_void_p),
        ("x_start", ctypes.c_void_p),
        ("wx", ctypes.c_void_p),
        ("mask", ctypes.c_void_p),
    ]


class MbsDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("eps_tensor", ctypes.c_int32),
        ("spacing", ctypes.c_double),
        ("k0", ctypes.c_double),
        ("bias_re", ctypes.c_double),
        ("bias_im", ctypes.c_double),
        ("scale_re", ctypes.c_double),
        ("scale_im", ctypes.c_double),
    ]


class CztDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("axis", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("num_param", ctypes.c_int32),
        ("mod_n", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("a", ctypes.POINTER(ctypes.c_double)),
        ("w", ctypes.POINTER(ctypes.c_double)),
        ("scale", ctypes.c_double),
    ]


class ZoomDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("m_y", ctypes.c_int32),
        ("m_x", ctypes.c_int32),
        ("mod_n_y", ctypes.c_int32),
        ("mod_n_x", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("a_y", ctypes.c_double * 2),
        ("w_y", ctypes.c_double * 2),
        ("a_x", ctypes.c_double * 2),
        ("w_x", ctypes.c_double * 2),
        ("scale_y", ctypes.c_double),
        ("scale_x", ctypes.c_double),
    ]


class YuDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("axis", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("df", ctypes.c_double),
        ("out_dx", ctypes.c_double),
        ("x0", ctypes.c_double),
    ]


class FluorDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("pad_y", ctypes.c_int32),
        ("pad_x", ctypes.c_int32),
        ("phase_mode", ctypes.c_int32),
        ("remove_evanescent", ctypes.c_int32),
        ("bandlimit", ctypes.c_int32),
        ("dy", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("n", ctypes.c_double),
        ("thickness", ctypes.c_double),
        ("ky", ctypes.c_double),
        ("kx", ctypes.c_double),
        ("intensity_scale", ctypes.c_double),
        ("prop_fwd", ctypes.c_void_p),
        ("prop_bwd", ctypes.c_void_p),
    ]


class FconvDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("k_planes", ctypes.c_int32),
        ("k_depth", ctypes.c_int32),
        ("k_height", ctypes.c_int32),
        ("k_width", ctypes.c_int32),
        ("real_input", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("out_depth", ctypes.c_int32),
        ("out_height", ctypes.c_int32),
        ("out_width", ctypes.c_int32),
        ("sum_batch", ctypes.c_int32),
    ]


class MultisliceDesc(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("depth", ctypes.c_int32),
        ("pad_y", ctypes.c_int32),
        ("pad_x", ctypes.c_int32),
        ("phase_mode", ctypes.c_int32),
        ("remove_evanescent", ctypes.c_int32),
        ("bandlimit", ctypes.c_int32),
        ("use_na", ctypes.c_int32),
        ("return_stack", ctypes.c_int32),
        ("out_kind", ctypes.c_int32),
        ("checkpoint_every", ctypes.c_int32),
        ("grad_kind", ctypes.c_int32),
        ("dy", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("wavelength", ctypes.c_double),
        ("n", ctypes.c_double),
        ("thickness", ctypes.c_double),
        ("NA", ctypes.c_double),
        ("ky", ctypes.c_double),
        ("kx", ctypes.c_double),
        ("reverse_distance", ctypes.c_double),
    ]


_lib = None
_lock = threading.Lock()


def _declare(lib):
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    dp = ctypes.POINTER(ctypes.c_double)
    sigs = {
        "cxm_init": (i32, [i32]),
        "cxm_last_error": (ctypes.c_char_p, []),
        "cxm_version": (i32, []),
        "cxm_fft_supported": (i32, [ctypes.c_int64]),
        "cxm_prop2d_workspace_size": (sz, [ctypes.POINTER(Prop2dDesc)]),
        "cxm_prop2d_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(Prop2dDesc)]),
        "cxm_propagator_c64": (i32, [vp, vp, ctypes.POINTER(Prop2dDesc)]),
        "cxm_zoom_dft_workspace_size": (sz, [ctypes.POINTER(ZoomDesc)]),
        "cxm_zoom_dft_mfma_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(ZoomDesc)]),
        "cxm_cgemm_mfma_c64": (i32, [vp, vp, vp, vp, i32, i32, i32, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, i32]),
        "cxm_modulation_vjp_c64": (i32, [vp, i32, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64, i32,
                                         i32, i32, dp, ctypes.c_double]),
        "cxm_noalloc_scratch_bytes": (sz, []),
        "cxm_prop2d_c64_noalloc": (i32, [vp, vp, vp, ctypes.POINTER(Prop2dDesc)]),
        "cxm_prop2d_param_bytes": (sz, []),
        "cxm_prop2d_build_params": (i32, [ctypes.POINTER(Prop2dDesc), dp, ctypes.c_int64, vp]),
        "cxm_prop2d_c64_params": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(Prop2dDesc), vp]),
        "cxm_set_option": (i32, [i32, i32]),
        "cxm_optical_fft_workspace_size": (sz, [ctypes.POINTER(OptFftDesc)]),
        "cxm_optical_fft_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(OptFftDesc)]),
        "cxm_sas_workspace_size": (sz, [ctypes.POINTER(OptFftDesc)]),
        "cxm_sas_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(OptFftDesc)]),
        "cxm_sas_precompensation_c64": (i32, [vp, vp, ctypes.POINTER(OptFftDesc)]),
        "cxm_timing_enable": (i32, [i32]),
        "cxm_timing_collect": (i32, []),
        "cxm_timing_get": (i32, [i32, ctypes.c_char_p, i32, dp, ctypes.POINTER(ctypes.c_int64)]),
        "cxm_intensity_c64": (i32, [vp, vp, vp, ctypes.c_int64, ctypes.c_int32, dp]),
        "cxm_plane_sumsq_workspace_size": (sz, [ctypes.c_int64]),
        "cxm_copy_c64": (i32, [vp, vp, vp, ctypes.c_int64]),
        "cxm_psf4f_workspace_size": (sz, [ctypes.POINTER(Psf4fDesc)]),
        "cxm_psf4f_fwd": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(Psf4fDesc)]),
        "cxm_psf4f_bwd": (i32, [vp, vp, vp, vp, vp, sz, ctypes.POINTER(Psf4fDesc)]),
        "cxm_psf4f_value_and_grad": (i32, [vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(Psf4fDesc)]),
        "cxm_point_source_c64": (i32, [vp, vp, ctypes.POINTER(Psf4fDesc)]),
        "cxm_phase_change_c64": (i32, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, dp]),
        "cxm_thin_sample_c64": (i32, [vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, dp, ctypes.c_double]),
        "cxm_intensity_vjp_c64": (i32, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int32, dp]),
        "cxm_cmul_c64": (i32, [vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_int32]),
        "cxm_czt_axis_workspace_size": (sz, [ctypes.POINTER(CztDesc)]),
        "cxm_resample_workspace_size": (sz, [ctypes.POINTER(ResampleDesc)]),
        "cxm_resample_f32": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(ResampleDesc)]),
        "cxm_mbs_workspace_size": (sz, [ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_precondition": (i32, [vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_iterate": (i32, [vp, vp, vp, vp, ctypes.c_int32, vp, vp, sz, ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_iterate_stop": (i32, [vp, vp, vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_double, vp, vp, sz, ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_output": (i32, [vp, vp, vp, vp, ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_bicgstab_state_size": (sz, [ctypes.POINTER(MbsDesc)]),
        "cxm_enclosing_circle_workspace_size": (sz, [ctypes.c_int64]),
        "cxm_enclosing_circle_c64": (i32, [vp, vp, ctypes.c_int64, dp, vp, sz]),
        "cxm_mbs_bicgstab_init": (i32, [vp, vp, vp, vp, vp, sz, ctypes.c_double, ctypes.c_double, vp,
                                        vp, sz, ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_bicgstab_iterate": (i32, [vp, vp, vp, vp, sz, ctypes.c_int32, vp, vp, sz,
                                           ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_vjp_workspace_size": (sz, [ctypes.POINTER(MbsDesc)]),
        "cxm_mbs_vjp": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(MbsDesc)]),
        "cxm_czt_axis_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(CztDesc)]),
        "cxm_fluor_multislice_workspace_size": (sz, [ctypes.POINTER(FluorDesc)]),
        "cxm_fluor_multislice_sample_c64": (i32, [vp, vp, vp, vp, vp, vp, vp, sz,
                                                  ctypes.POINTER(FluorDesc)]),
        "cxm_yu_axis_workspace_size": (sz, [ctypes.POINTER(YuDesc)]),
        "cxm_yu_axis_c64": (i32, [vp, vp, vp, vp, sz, ctypes.POINTER(YuDesc)]),
        "cxm_czt_matrix_c64": (i32, [vp, ctypes.POINTER(CztDesc), ctypes.c_int32, vp, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_int32]),
        "cxm_yu_matrix_c64": (i32, [vp, ctypes.POINTER(YuDesc), vp, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int32]),
        "cxm_high_na_defocus_c64": (i32, [vp, vp, vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "cxm_fconv_workspace_size": (sz, [ctypes.POINTER(FconvDesc)]),
        "cxm_fconv_c64": (i32, [vp, vp, vp, vp, vp, sz, ctypes.POINTER(FconvDesc)]),
        "cxm_sum_axis_f32": (i32, [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
        "cxm_shot_noise_approx_f32": (i32, [vp, vp, vp, vp, ctypes.c_int64]),
        "cxm_multislice_workspace_size": (sz, [ctypes.POINTER(MultisliceDesc)]),
        "cxm_multislice_checkpoint_size": (sz, [ctypes.POINTER(MultisliceDesc)]),
        "cxm_multislice_fwd": (i32, [vp, vp, vp, vp, vp, vp, vp, sz, ctypes.POINTER(MultisliceDesc)]),
        "cxm_multislice_bwd_workspace_size": (sz, [ctypes.POINTER(MultisliceDesc)]),
        "cxm_multislice_bwd": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, sz,
                                     ctypes.POINTER(MultisliceDesc)]),
        "cxm_plane_sumsq_c64": (i32, [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, dp, vp, sz]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The loaded libcxm (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise RuntimeError(
                        f"{LIB_PATH} is missing; build it with `python -m chromatix_amd.build`"
                    )
                _lib = _declare(ctypes.CDLL(str(LIB_PATH)))
    return _lib


def check(code: int) -> None:
    if code != CXM_OK:
        msg = lib().cxm_last_error()
        raise CxmError(code, msg.decode() if msg else "")


_initialised: set[int] = set()


def ensure_device(t: torch.Tensor) -> None:
    """Require a HIP device tensor and initialise libcxm on its device."""
    if not t.is_cuda:
        raise RuntimeError(
            "chromatix_amd runs on the GPU only: move the field to a HIP device "
            "(no CPU fallback exists)"
        )
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    if dev not in _initialised:
        with torch.cuda.device(dev):
            check(lib().cxm_init(dev))
        _initialised.add(dev)


def refuse_grad(what: str, *tensors) -> None:
    """Raise where an entry point without an adjoint would receive an input
    that requires grad: its output would carry no grad_fn and the gradient
    would be lost without an error."""
    if not torch.is_grad_enabled():
        return
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.requires_grad:
            raise NotImplementedError(
                f"{what}: gradients are not implemented for this entry point; "
                "call it under torch.no_grad() or detach the input")


def stream_ptr(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def dptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def doubles(values) -> ctypes.Array:
    vals = [float(v) for v in values]
    return (ctypes.c_double * max(1, len(vals)))(*vals)


def timing_enable(on: bool = True) -> None:
    check(lib().cxm_timing_enable(1 if on else 0))


def timing_collect() -> dict[str, tuple[float, int]]:
    """{kernel name: (total ms, launches)} since timing_enable(True)."""
    n = lib().cxm_timing_collect()
    if n < 0:
        check(-n)
    out = {}
    buf = ctypes.create_string_buffer(128)
    for i in range(n):
        ms = ctypes.c_double()
        cnt = ctypes.c_int64()
        check(lib().cxm_timing_get(i, buf, 128, ctypes.byref(ms), ctypes.byref(cnt)))
        out[buf.value.decode()] = (ms.value, cnt.value)
    return out
