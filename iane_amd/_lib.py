"""ctypes binding of libiane_hip.so (declared in include/iane_hip.h).

The library is built in-tree by ``iane_amd.build``. Loading it needs no GPU;
calling a compute entry point does. There is no fallback: if the library is
missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libiane_hip.so"

IANE_MAX_RANK = 4
U8, U16, I32, F32, F64, C64 = range(6)
OK, ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED, ERR_HIP, ERR_LENGTH, ERR_OVERFLOW = range(6)


class iane_view_desc(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("rank", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("extent", ctypes.c_int32 * IANE_MAX_RANK),
        ("stride_bytes", ctypes.c_int32 * IANE_MAX_RANK),
    ]


# ---- exceptions mirroring the reference's std:: exception types ----------
class invalid_argument(ValueError):
    """std::invalid_argument"""


class unsupported(NotImplementedError):
    """Element type / rank / size this backend does not provide."""


class length_error(RuntimeError):
    """std::length_error (scratch reservation exceeded)."""


class overflow_error(OverflowError):
    """std::overflow_error"""


class runtime_error(RuntimeError):
    """std::runtime_error carrying the HIP error text."""


_ERRORS = {
    ERR_INVALID_ARGUMENT: invalid_argument,
    ERR_UNSUPPORTED: unsupported,
    ERR_HIP: runtime_error,
    ERR_LENGTH: length_error,
    ERR_OVERFLOW: overflow_error,
}

vp = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64
f32 = ctypes.c_float
sz = ctypes.c_size_t
desc_p = ctypes.POINTER(iane_view_desc)

# name -> (argtypes); every function returns int status
SIGNATURES = {
    "iane_hip_device_count": [ctypes.POINTER(ctypes.c_int)],
    "iane_hip_tune_set": [ctypes.c_char_p, ctypes.c_int],
    "iane_hip_tune_get": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
    "iane_hip_set_device": [ctypes.c_int],
    "iane_hip_malloc": [ctypes.POINTER(vp), sz],
    "iane_hip_free": [vp],
    "iane_hip_host_malloc": [ctypes.POINTER(vp), sz],
    "iane_hip_host_free": [vp],
    "iane_hip_stream_create": [ctypes.POINTER(vp)],
    "iane_hip_stream_destroy": [vp],
    "iane_hip_stream_sync": [vp],
    "iane_hip_device_sync": [],
    "iane_hip_memcpy_async": [vp, vp, sz, ctypes.c_int, vp],
    "iane_hip_copy_view": [desc_p, desc_p, ctypes.c_int, vp],
    "iane_hip_memset_async": [vp, ctypes.c_int, sz, vp],
    "iane_hip_event_create": [ctypes.POINTER(vp)],
    "iane_hip_event_destroy": [vp],
    "iane_hip_event_record": [vp, vp],
    "iane_hip_event_elapsed_ms": [ctypes.POINTER(f32), vp, vp],
    "iane_hip_convolve_workspace_bytes": [desc_p, desc_p, desc_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(sz)],
    "iane_hip_convolve": [desc_p, desc_p, desc_p, ctypes.c_int, ctypes.c_int, vp, sz, vp],
    "iane_hip_convolve_untiled": [desc_p, desc_p, desc_p, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_convolve_separable_workspace_bytes": [desc_p, desc_p, ctypes.c_int, desc_p, ctypes.c_int, ctypes.c_int,
                                                    ctypes.POINTER(sz)],
    "iane_hip_convolve_separable": [desc_p, desc_p, ctypes.c_int, desc_p, ctypes.c_int, ctypes.c_int, vp, sz, vp],
    "iane_hip_convolve_separable_passes_workspace_bytes": [desc_p, desc_p, ctypes.c_int, desc_p, ctypes.c_int,
                                                           ctypes.c_int, ctypes.POINTER(sz)],
    "iane_hip_convolve_separable_passes": [desc_p, desc_p, ctypes.c_int, desc_p, ctypes.c_int, ctypes.c_int, vp, sz,
                                           vp],
    "iane_hip_fft_plan_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(i32), ctypes.c_int],
    "iane_hip_fft_plan_workspace_bytes": [vp, ctypes.POINTER(sz)],
    "iane_hip_fft_exec": [vp, vp, vp, vp, vp],
    "iane_hip_fft_plan_destroy": [vp],
    "iane_hip_rl_initial": [vp, vp, i64, f32, vp],
    "iane_hip_rl_ratio_workspace_bytes": [i64, ctypes.POINTER(sz)],
    "iane_hip_rl_ratio": [vp, vp, vp, i64, f32, vp, vp, sz, vp],
    "iane_hip_rl_ratio_noalloc": [vp, vp, vp, i64, f32, vp, vp],
    "iane_hip_rl_update": [vp, vp, vp, i64, f32, vp],
    "iane_hip_wiener_apply": [vp, vp, i64, f32, vp],
    "iane_hip_real_to_complex": [vp, vp, i64, vp],
    "iane_hip_complex_real_scale": [vp, vp, i64, f32, vp],
    "iane_hip_rl_plan_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(i32), desc_p, ctypes.c_int,
                                ctypes.c_int, f32, vp],
    "iane_hip_rl_plan_workspace_bytes": [vp, ctypes.POINTER(sz)],
    "iane_hip_rl_run": [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp],
    "iane_hip_rl_plan_destroy": [vp],
    "iane_hip_rl_batch_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(i32), desc_p, f32, vp],
    "iane_hip_rl_batch_workspace_bytes": [vp, ctypes.POINTER(sz)],
    "iane_hip_rl_batch_run": [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, vp, vp,
                              ctypes.c_int, vp],
    "iane_hip_rl_batch_destroy": [vp],
    "iane_hip_rl_plan_blur": [vp, vp, vp, vp, vp],
    "iane_hip_cnn_conv3d": [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_transpose_weights": [vp, vp, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_wgrad": [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_relu_backward": [vp, vp, i64, vp],
    "iane_hip_cnn_pool2": [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_pool2_backward": [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_up2_add": [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_up2_backward": [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp],
    "iane_hip_cnn_to_bf16": [vp, vp, i64, f32, vp],
    "iane_hip_cnn_adam": [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, ctypes.c_int, vp],
    "iane_hip_cnn_residual_mse": [vp, vp, vp, vp, vp, i64, vp, ctypes.POINTER(ctypes.c_int), vp],
    "iane_hip_cnn_warm_start": [vp, vp, vp, i64, f32, f32, vp],
    "iane_hip_synth_blobs": [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, f32, vp, ctypes.c_int, vp],
    "iane_hip_synth_poisson": [vp, vp, i64, ctypes.c_uint, vp],
    "iane_hip_max": [vp, i64, vp, vp, vp],
    "iane_hip_sequence_plan_create": [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp,
                                      desc_p, desc_p],
    "iane_hip_sequence_plan_workspace_bytes": [vp, ctypes.POINTER(ctypes.c_size_t)],
    "iane_hip_sequence_run": [vp, desc_p, desc_p, vp, vp],
    "iane_hip_sequence_plan_destroy": [vp],
    "iane_hip_scale": [vp, i64, f32, vp],
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (no GPU needed). Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m iane_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue  # reported by exported_symbols / tests
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.iane_hip_last_error.argtypes = []
    lib.iane_hip_last_error.restype = ctypes.c_char_p
    lib.iane_hip_version.argtypes = []
    lib.iane_hip_version.restype = ctypes.c_int
    _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status == OK:
        return
    message = load().iane_hip_last_error().decode(errors="replace")
    exc = _ERRORS.get(status, runtime_error)
    raise exc(f"{what}: {message}" if what else message)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)
