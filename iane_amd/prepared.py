"""Prepared algorithms with one shared scratch allocation and the FFT drop-ins.

Mirrors include/iane/execution/prepared.hpp:84-180 (execution_context,
prepare/run, scratch growth, operations bound to their context) and
include/iane/algorithms/fft/{fft.hpp:196-231, rfft.hpp:107-161, layout.hpp:15-21}
(fft_forward, fft_inverse, rfft, irfft, rfft_extent). Images are torch tensors
on the GPU; complex images are complex64 tensors (= fft_complex{real, imag}).
IANE axis 0 is the last tensor dimension, so ``rfft_extent`` halves the last
dimension of a tensor shape.
"""
from __future__ import annotations

import ctypes
import dataclasses

from . import _lib
from ._lib import invalid_argument, length_error


def rfft_extent(real_shape):
    """layout.hpp:15-21 for a tensor shape: last dimension n -> n // 2 + 1."""
    shape = tuple(int(s) for s in real_shape)
    if not shape or any(s <= 0 for s in shape):
        raise invalid_argument("Real FFT requires positive dimensions")
    return shape[:-1] + (shape[-1] // 2 + 1,)


def _check_packed(t, what):
    if t.device.type != "cuda":
        raise invalid_argument(f"{what}: FFT views must be device images")
    if not t.is_contiguous():
        raise invalid_argument("FFT requires packed views")
    if t.data_ptr() % 8:
        raise invalid_argument("FFT requires aligned input and output buffers")


@dataclasses.dataclass
class _fft_state:
    """A prepared FFT: the plan for one shape (fft.hpp:174-193)."""
    kind: int
    real_shape: tuple
    spec_shape: tuple
    plan: ctypes.c_void_p
    workspace: int

    def scratch_bytes(self) -> int:
        return self.workspace

    def _check(self, src, dst):
        import torch
        if self.kind in (0, 1):
            want_in, want_out = torch.complex64, torch.complex64
            in_shape, out_shape = self.spec_shape, self.spec_shape
        elif self.kind == 2:
            want_in, want_out = torch.float32, torch.complex64
            in_shape, out_shape = self.real_shape, self.spec_shape
        else:
            want_in, want_out = torch.complex64, torch.float32
            in_shape, out_shape = self.spec_shape, self.real_shape
        if src.dtype != want_in or dst.dtype != want_out:
            raise invalid_argument("FFT element types do not match the prepared transform")
        _check_packed(src, "source")
        _check_packed(dst, "destination")
        if tuple(src.shape) != in_shape or tuple(dst.shape) != out_shape:
            raise invalid_argument("FFT views must match the prepared shape")

    def run(self, src, dst, scratch, stream):
        self._check(src, dst)
        ws = scratch.bytes(self.workspace, 8) if self.workspace else 0
        _lib.call("iane_hip_fft_exec", self.plan, src.data_ptr(), dst.data_ptr(), ws, stream)

    def __del__(self):
        if self.plan:
            try:
                _lib.load().iane_hip_fft_plan_destroy(self.plan)
            except Exception:
                pass
            self.plan = None


class _fft_algorithm:
    """fft_forward / fft_inverse / rfft / irfft: prepare(src, dst, policy) -> state."""

    def __init__(self, kind: int, name: str):
        self.kind = kind
        self.name = name

    def __repr__(self):
        return f"iane_amd.{self.name}"

    def prepare(self, src, dst, stream):
        if self.kind in (0, 1):
            real_shape = spec_shape = tuple(src.shape)
            if tuple(dst.shape) != real_shape:
                raise invalid_argument("Image dimensions must match")
        elif self.kind == 2:
            real_shape, spec_shape = tuple(src.shape), tuple(dst.shape)
        else:
            real_shape, spec_shape = tuple(dst.shape), tuple(src.shape)
        if any(s <= 0 for s in real_shape):
            raise invalid_argument("FFT requires positive dimensions")
        if self.kind >= 2 and rfft_extent(real_shape) != spec_shape:
            raise invalid_argument("Real FFT requires matching real and packed spectrum extents")
        rank = len(real_shape)
        extent = (ctypes.c_int32 * 3)(*([int(s) for s in reversed(real_shape)] + [1] * (3 - rank)))
        plan = ctypes.c_void_p()
        _lib.call("iane_hip_fft_plan_create", ctypes.byref(plan), rank, extent, self.kind)
        nbytes = ctypes.c_size_t()
        _lib.call("iane_hip_fft_plan_workspace_bytes", plan, ctypes.byref(nbytes))
        state = _fft_state(self.kind, real_shape, spec_shape, plan, nbytes.value)
        state._check(src, dst)
        return state


fft_forward = _fft_algorithm(0, "fft_forward")
fft_inverse = _fft_algorithm(1, "fft_inverse")
rfft = _fft_algorithm(2, "rfft")
irfft = _fft_algorithm(3, "irfft")


class scratch_space:
    """Carves regions out of the context's allocation (prepared.hpp:18-57)."""

    def __init__(self, base: int, size: int):
        self._next = base
        self._end = base + size

    def bytes(self, size: int, alignment: int = 16) -> int:
        if size == 0:
            return 0
        if alignment <= 0 or alignment & (alignment - 1) or alignment > 16:
            raise invalid_argument("Unsupported scratch alignment")
        start = (self._next + alignment - 1) // alignment * alignment
        if start + size > self._end:
            raise length_error("Prepared operation exceeded its scratch reservation")
        self._next = start + size
        return start


class prepared_operation:
    def __init__(self, owner, state):
        self._owner = owner
        self.state = state


class execution_context:
    """One shared device scratch allocation for prepared algorithms (prepared.hpp:84-180).

    Calls must be serialised by the caller; the stream must outlive the context."""

    def __init__(self, policy=None):
        from . import hip_policy, hip_execution
        if policy is None:
            policy = hip_policy()
        if isinstance(policy, hip_execution):
            policy = policy.policy()
        self.policy = policy
        self._stream = policy.handle()
        self._scratch = None
        self._capacity = 0
        self._submitted = False

    def prepare(self, algorithm, src, dst) -> prepared_operation:
        state = algorithm.prepare(src, dst, self._stream)
        self._reserve(state.scratch_bytes())
        return prepared_operation(self, state)

    def run(self, operation: prepared_operation, src, dst):
        from . import hip_execution
        if operation._owner is not self:
            raise invalid_argument("Prepared operation belongs to a different execution context")
        self._submitted = True
        base = self._scratch.data_ptr() if self._scratch is not None else 0
        operation.state.run(src, dst, scratch_space(base, self._capacity), self._stream)
        return hip_execution(self._stream)

    def scratch_capacity(self) -> int:
        return self._capacity

    def _reserve(self, size: int) -> None:
        import torch
        if size <= self._capacity:
            return
        if self._submitted:
            _lib.call("iane_hip_stream_sync", self._stream)  # prepared.hpp:159-163
            self._submitted = False
        self._scratch = torch.empty(size, dtype=torch.uint8, device="cuda")
        self._capacity = size

    def __del__(self):
        if getattr(self, "_submitted", False):
            try:
                _lib.load().iane_hip_stream_sync(self._stream)
            except Exception:
                pass


def make_execution_context(policy=None) -> execution_context:
    return execution_context(policy)
