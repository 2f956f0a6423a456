"""iane_amd — MI355X (gfx950) backend of IANE's convolution / deconvolution path.

Python mirror of the reference's include/iane entry points for device images:
``convolve`` (include/iane/algorithms/convolution.hpp:97-117) with
``convolution_options{border}`` and ``border_mode`` (locator/image_locator.hpp:14),
an execution policy carrying a stream (execution/context.hpp:46-59), prepared
algorithms in an ``execution_context`` (execution/prepared.hpp:84-180) for the
FFT drop-ins ``fft_forward``/``fft_inverse``/``rfft``/``irfft``
(algorithms/fft/fft.hpp:228-231, rfft.hpp:152-161) and the Richardson–Lucy and
Wiener loops of examples/{rl_deconvolution,wiener}/main.cu.

Images are torch tensors on the GPU (torch is only the allocator and stream
provider). IANE's axis 0 is the fastest axis, which is the LAST dimension of a
C-ordered tensor: a tensor of shape (Z, Y, X) is an IANE image of extent
{X, Y, Z}. Every computation runs in libiane_hip.so; nothing here computes.
"""
from __future__ import annotations

import dataclasses
import enum

from . import _lib
from ._lib import invalid_argument, length_error, overflow_error, runtime_error, unsupported  # noqa: F401

__all__ = [
    "border_mode", "precision", "convolution_options", "hip_policy", "hip_execution",
    "convolve", "convolve_workspace_bytes", "convolve_separable", "view_desc", "library",
]


class border_mode(enum.IntEnum):
    """Sampling rule outside the source view (image_locator.hpp:14)."""
    clamp = 0
    zero = 1


class precision(enum.IntEnum):
    """Arithmetic of a convolution. ``reference`` is the reference's FP64 fma
    chain in kernel storage order (bit-exact); the others trade exactness for speed."""
    reference = 0
    fp32 = 1
    split_bf16 = 2
    split_fp16 = 3


@dataclasses.dataclass(frozen=True)
class convolution_options:
    """convolution.hpp:16-18 (border, default clamp) plus the arithmetic mode."""
    border: border_mode = border_mode.clamp
    precision: precision = precision.reference


@dataclasses.dataclass(frozen=True)
class hip_policy:
    """Queue work on a borrowed HIP stream (context.hpp:46-49). ``stream=None``
    uses torch's current stream on the current device."""
    stream: int | None = None

    def handle(self) -> int:
        if self.stream is not None:
            return int(self.stream)
        import torch
        return int(torch.cuda.current_stream().cuda_stream)


@dataclasses.dataclass(frozen=True)
class hip_execution:
    """Completion handle (context.hpp:51-59): ``wait()`` synchronises the stream."""
    stream: int

    def wait(self) -> None:
        _lib.call("iane_hip_stream_sync", self.stream)

    def policy(self) -> hip_policy:
        return hip_policy(self.stream)


def library():
    """The loaded ctypes library (raises ImportError if it was not built)."""
    return _lib.load()


_TORCH_DTYPES = None


def _dtype_code(dtype) -> int:
    global _TORCH_DTYPES
    if _TORCH_DTYPES is None:
        import torch
        _TORCH_DTYPES = {
            torch.uint8: _lib.U8, torch.uint16: _lib.U16, torch.int32: _lib.I32,
            torch.float32: _lib.F32, torch.float64: _lib.F64, torch.complex64: _lib.C64,
        }
    try:
        return _TORCH_DTYPES[dtype]
    except KeyError:
        raise unsupported(f"element type {dtype} is not one of u8, u16, i32, f32, f64, c64") from None


def view_desc(tensor, *, require_device: bool = True) -> _lib.iane_view_desc:
    """Describe a tensor as an IANE view: extent/stride order reversed (axis 0 fastest)."""
    if require_device and tensor.device.type != "cuda":
        raise invalid_argument("iane_amd computes on device images; got a host tensor")
    rank = tensor.dim()
    if not 1 <= rank <= _lib.IANE_MAX_RANK:
        raise unsupported(f"rank {rank} outside 1..{_lib.IANE_MAX_RANK}")
    d = _lib.iane_view_desc()
    d.ptr = tensor.data_ptr() if tensor.numel() else 0
    d.rank = rank
    d.dtype = _dtype_code(tensor.dtype)
    esize = tensor.element_size()
    for axis in range(rank):
        dim = rank - 1 - axis
        d.extent[axis] = tensor.shape[dim]
        stride = tensor.stride(dim) * esize
        if not -2**31 <= stride < 2**31:
            raise overflow_error("Packed image stride exceeds the supported range")
        d.stride_bytes[axis] = stride
    return d


def _stream(policy) -> int:
    if policy is None:
        policy = hip_policy()
    if isinstance(policy, hip_execution):
        policy = policy.policy()
    return policy.handle()


def _workspace(nbytes: int, workspace):
    """Caller-side scratch for an entry point that needs ``nbytes`` (torch's
    caching allocator stands in for the caller's execution_context)."""
    if nbytes == 0:
        return 0
    if workspace is None:
        import torch
        workspace = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    elif workspace.device.type != "cuda" or workspace.numel() * workspace.element_size() < nbytes:
        raise length_error("Prepared operation exceeded its scratch reservation")
    return workspace


def convolve_workspace_bytes(src, kernel, dst, options: convolution_options | None = None) -> int:
    """Scratch ``convolve`` needs for these shapes (nonzero only on the split-precision MFMA path)."""
    import ctypes
    options = options or convolution_options()
    if isinstance(src, texture_image_view):
        src = src.texels()
    nbytes = ctypes.c_size_t()
    _lib.call("iane_hip_convolve_workspace_bytes", view_desc(src), view_desc(kernel), view_desc(dst),
              int(options.border), int(options.precision), ctypes.byref(nbytes))
    return nbytes.value


def convolve(src, kernel, dst, options: convolution_options | None = None, policy=None,
             workspace=None) -> hip_execution:
    """iane::convolve for device images (convolution.hpp:97-117); ``src`` may be a
    texture view (docs/cuda_texture_images.md:5-17).

    ``dst[p] = sum_q kernel[q] * src[p + c - q]`` with ``c = extent // 2`` per axis.
    Raises ``invalid_argument`` for even/non-positive kernel extents, rank or
    extent mismatch, exactly where the reference throws. ``workspace`` (a device
    tensor of at least ``convolve_workspace_bytes`` bytes) is used by the
    split-precision path; without one it is taken from torch's allocator.
    """
    options = options or convolution_options()
    stream = _stream(policy)
    if isinstance(src, texture_image_view):
        src = src.texels()  # texture source: its texels, border rule at the view boundary
    s, k, d = view_desc(src), view_desc(kernel), view_desc(dst)
    border, prec = int(options.border), int(options.precision)
    needed = 0
    if prec in (precision.split_bf16, precision.split_fp16):  # the only path that takes scratch
        import ctypes
        query = ctypes.c_size_t()
        _lib.call("iane_hip_convolve_workspace_bytes", s, k, d, border, prec, ctypes.byref(query))
        needed = query.value
    ws = _workspace(needed, workspace)
    ptr, nbytes = (0, 0) if isinstance(ws, int) else (ws.data_ptr(), ws.numel() * ws.element_size())
    _lib.call("iane_hip_convolve", s, k, d, border, prec, ptr, nbytes, stream)
    return hip_execution(stream)


def convolve_separable(src, taps, dst, options: convolution_options | None = None, policy=None,
                       *, fused: bool = True, workspace=None) -> hip_execution:
    """Convolve with the outer product of ``taps`` (one 1-D tensor per IANE axis,
    ``taps[0]`` along axis 0 = the last tensor dimension)."""
    import ctypes
    options = options or convolution_options(precision=precision.fp32)
    stream = _stream(policy)
    arr = (_lib.iane_view_desc * len(taps))(*[view_desc(t) for t in taps])
    name = "iane_hip_convolve_separable" if fused else "iane_hip_convolve_separable_passes"
    s, d = view_desc(src), view_desc(dst)
    nbytes = ctypes.c_size_t()
    _lib.call(name + "_workspace_bytes", s, arr, len(taps), d, int(options.border), int(options.precision),
              ctypes.byref(nbytes))
    ws = _workspace(nbytes.value, workspace)
    ptr, size = (0, 0) if isinstance(ws, int) else (ws.data_ptr(), ws.numel() * ws.element_size())
    _lib.call(name, s, arr, len(taps), d, int(options.border), int(options.precision), ptr, size, stream)
    return hip_execution(stream)


def convolve_untiled(src, kernel, dst, options: convolution_options | None = None, policy=None) -> hip_execution:
    """The one-thread-per-output kernel (used for large kernels); exposed for tests."""
    options = options or convolution_options()
    stream = _stream(policy)
    _lib.call("iane_hip_convolve_untiled", view_desc(src), view_desc(kernel), view_desc(dst),
              int(options.border), int(options.precision), stream)
    return hip_execution(stream)


from .texture import copy, texture_image, texture_image_view  # noqa: E402
from .sequence import (exclusive_scan, exclusive_scan_along, inclusive_scan, inclusive_scan_along,  # noqa: E402
                       reduce, reduce_along)

__all__ += ["texture_image", "texture_image_view", "copy", "reduce", "reduce_along", "inclusive_scan",
            "exclusive_scan", "inclusive_scan_along", "exclusive_scan_along"]

from .prepared import (execution_context, fft_forward, fft_inverse, irfft, make_execution_context,  # noqa: E402
                       rfft, rfft_extent)

__all__ += ["execution_context", "make_execution_context", "fft_forward", "fft_inverse", "rfft", "irfft",
            "rfft_extent"]
