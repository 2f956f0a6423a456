"""Scan and reduce prepared algorithms for device images: ``reduce``,
``reduce_along``, ``inclusive_scan``, ``exclusive_scan``,
``inclusive_scan_along`` and ``exclusive_scan_along`` of the reference
(include/iane/algorithms/{reduce.hpp:7-36, scan.hpp:7-46} over
algorithms/detail/sequence.hpp:255-301), used through ``execution_context``::

    ctx = iane_amd.make_execution_context(policy)
    op = ctx.prepare(iane_amd.reduce("sum", 0.0), src, total)   # total: float64 tensor of shape (1,)
    ctx.run(op, src, total).wait()

The operator is "sum", "product", "min" or "max" (also ``operator.add``,
``operator.mul``, ``min``, ``max``). The accumulator type is the destination's
element type; ``init`` is converted to it, applied once per line and is the
result of an empty reduction. ``axis`` counts IANE axes (axis 0 = the last
tensor dimension), like the reference's template parameter. Shapes: scans keep
the source shape; a whole reduction (or one of a 1-D image) writes shape (1,);
reduce_along drops that axis. Integer sums and products wrap; floating-point
grouping may vary (as the reference allows).
"""
from __future__ import annotations

import ctypes
import dataclasses
import operator

import numpy as np

from . import _lib
from ._lib import invalid_argument

INCLUSIVE, EXCLUSIVE, REDUCE = 0, 1, 2
OPS = {"sum": 0, "product": 1, "min": 2, "max": 3}
_CALLABLES = {operator.add: 0, operator.mul: 1, min: 2, max: 3}
_NUMPY = {_lib.U8: np.uint8, _lib.U16: np.uint16, _lib.I32: np.int32, _lib.F32: np.float32, _lib.F64: np.float64}


def _op_code(op) -> int:
    if isinstance(op, str):
        if op not in OPS:
            raise invalid_argument(f"unknown sequence operator {op!r}; expected one of {sorted(OPS)}")
        return OPS[op]
    try:
        return _CALLABLES[op]
    except (KeyError, TypeError):
        raise invalid_argument("sequence operators are 'sum', 'product', 'min', 'max' "
                               "(or operator.add, operator.mul, min, max)") from None


@dataclasses.dataclass
class _sequence_state:
    plan: ctypes.c_void_p
    workspace: int

    def scratch_bytes(self) -> int:
        return self.workspace

    def run(self, src, dst, scratch, stream):
        from . import view_desc
        ws = scratch.bytes(self.workspace, 8) if self.workspace else 0
        _lib.call("iane_hip_sequence_run", self.plan, view_desc(src), view_desc(dst), ws, stream)

    def __del__(self):
        if self.plan:
            try:
                _lib.load().iane_hip_sequence_plan_destroy(self.plan)
            except Exception:
                pass
            self.plan = None


class _sequence_algorithm:
    def __init__(self, kind: int, axis: int, op, init, name: str):
        self.kind, self.axis, self.op, self.init, self.name = kind, axis, _op_code(op), init, name

    def __repr__(self):
        return f"iane_amd.{self.name}"

    def prepare(self, src, dst, stream):
        from . import view_desc
        s, d = view_desc(src), view_desc(dst)
        if d.dtype not in _NUMPY:
            raise _lib.unsupported("scan and reduce support u8, u16, i32, f32 and f64 elements")
        init = np.asarray(self.init).astype(_NUMPY[d.dtype])
        buffer = ctypes.create_string_buffer(init.tobytes(), 8)
        plan = ctypes.c_void_p()
        _lib.call("iane_hip_sequence_plan_create", ctypes.byref(plan), self.kind, self.axis, self.op, d.dtype, buffer,
                  s, d)
        nbytes = ctypes.c_size_t()
        _lib.call("iane_hip_sequence_plan_workspace_bytes", plan, ctypes.byref(nbytes))
        return _sequence_state(plan, nbytes.value)


def reduce(op, init):
    """Reduce the whole source into shape (1,) (reduce.hpp:7-22)."""
    return _sequence_algorithm(REDUCE, -1, op, init, "reduce")


def reduce_along(axis: int, op, init):
    """Independent reductions along IANE axis ``axis``; the output drops it (reduce.hpp:24-36)."""
    if axis < 0:
        raise invalid_argument("reduce_along requires a nonnegative axis")
    return _sequence_algorithm(REDUCE, axis, op, init, "reduce_along")


def inclusive_scan(op, init):
    """Inclusive prefix scan of the whole source in storage order (scan.hpp:7-24)."""
    return _sequence_algorithm(INCLUSIVE, -1, op, init, "inclusive_scan")


def exclusive_scan(op, init):
    """Exclusive prefix scan of the whole source (scan.hpp:26-31)."""
    return _sequence_algorithm(EXCLUSIVE, -1, op, init, "exclusive_scan")


def inclusive_scan_along(axis: int, op, init):
    """Independent inclusive scans along IANE axis ``axis`` (scan.hpp:33-38)."""
    if axis < 0:
        raise invalid_argument("inclusive_scan_along requires a nonnegative axis")
    return _sequence_algorithm(INCLUSIVE, axis, op, init, "inclusive_scan_along")


def exclusive_scan_along(axis: int, op, init):
    """Independent exclusive scans along IANE axis ``axis`` (scan.hpp:40-45)."""
    if axis < 0:
        raise invalid_argument("exclusive_scan_along requires a nonnegative axis")
    return _sequence_algorithm(EXCLUSIVE, axis, op, init, "exclusive_scan_along")
