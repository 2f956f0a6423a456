"""Scan and reduce on the GPU (iane_hip_sequence_*) against the oracle's serial
fold (algorithms/detail/sequence.hpp:73-150 restated in oracle.sequence).

Integer accumulators and integer-valued floating data are bit-exact (any
grouping gives the same value); random floating sums are within
1e-6 * sum|x| (the grouping may vary, as the reference allows).
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import iane_amd
from oracle import oracle

pytestmark = pytest.mark.gpu
KINDS = {"inclusive": oracle.SEQ_INCLUSIVE, "exclusive": oracle.SEQ_EXCLUSIVE, "reduce": oracle.SEQ_REDUCE}


def algorithm(kind, axis, op, init):
    if kind == "reduce":
        return iane_amd.reduce(op, init) if axis < 0 else iane_amd.reduce_along(axis, op, init)
    if kind == "inclusive":
        return iane_amd.inclusive_scan(op, init) if axis < 0 else iane_amd.inclusive_scan_along(axis, op, init)
    return iane_amd.exclusive_scan(op, init) if axis < 0 else iane_amd.exclusive_scan_along(axis, op, init)


def out_shape(shape, kind, axis):
    if kind != "reduce":
        return shape
    if axis < 0 or len(shape) == 1:
        return (1,)
    return tuple(n for i, n in enumerate(shape) if i != len(shape) - 1 - axis)


def run_gpu(src_np, kind, axis, op, init, acc):
    import torch
    src = torch.from_numpy(np.ascontiguousarray(src_np)).cuda()
    dst = torch.full(out_shape(src_np.shape, kind, axis), 77, dtype=torch.from_numpy(np.zeros(1, acc)).dtype,
                     device="cuda")
    ctx = iane_amd.make_execution_context()
    prepared = ctx.prepare(algorithm(kind, axis, op, init), src, dst)
    ctx.run(prepared, src, dst).wait()
    return dst.cpu().numpy()


def test_reference_sequence_case():
    """sequence_cases.hpp:31-117 with int32 and float64 accumulators (the device ABI has no u64)."""
    import torch
    y, x = np.mgrid[0:5, 0:263]
    src = (x % 7 + y).astype(np.uint8)
    for acc in (np.int32, np.float64):
        for kind, axis, init in [("exclusive", -1, 11), ("inclusive", -1, 11), ("reduce", -1, 11), ("reduce", 0, 11),
                                 ("exclusive", 1, 11), ("inclusive", 0, 0)]:
            got = run_gpu(src, kind, axis, "sum", init, acc)
            np.testing.assert_array_equal(got, oracle.sequence(src, KINDS[kind], axis, "sum", init, acc))
        # row scan into a padded intermediate, then column scan: 2-D prefix sums
        d_src = torch.from_numpy(src).cuda()
        tdt = torch.from_numpy(np.zeros(1, acc)).dtype
        inter = torch.zeros((5, 265), dtype=tdt, device="cuda")
        rows = inter[:, 1:264]
        out = torch.zeros((5, 263), dtype=tdt, device="cuda")
        ctx = iane_amd.make_execution_context()
        row_scan = ctx.prepare(iane_amd.inclusive_scan_along(0, "sum", 0), d_src, rows)
        column_scan = ctx.prepare(iane_amd.inclusive_scan_along(1, "sum", 0), rows, out)
        ctx.run(row_scan, d_src, rows)
        ctx.run(column_scan, rows, out).wait()
        expected = np.cumsum(np.cumsum(src.astype(np.int64), axis=1), axis=0)
        np.testing.assert_array_equal(out.cpu().numpy(), expected.astype(acc))


SHAPES = [(1000,), (5000,), (37, 129), (3, 4100), (5, 31, 70), (300, 2, 3), (3, 4, 5, 6), (2600, 7)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["inclusive", "exclusive", "reduce"])
def test_every_axis_sum_min_max_int32_exact(shape, kind):
    rng = np.random.default_rng(len(shape) * 100 + sum(shape))
    src = rng.integers(-50, 50, shape).astype(np.int32)
    for axis in [-1, *range(len(shape))]:
        for op, init in (("sum", 3), ("min", 40), ("max", -40)):
            got = run_gpu(src, kind, axis, op, init, np.int32)
            np.testing.assert_array_equal(got, oracle.sequence(src, KINDS[kind], axis, op, init, np.int32),
                                          err_msg=f"{kind} axis={axis} op={op}")


@pytest.mark.parametrize("acc", [np.uint8, np.uint16, np.int32])
@pytest.mark.parametrize("op", ["sum", "product"])
def test_integer_wraparound(acc, op):
    rng = np.random.default_rng(5)
    src = rng.integers(1, 4 if op == "product" else 250, (17, 3000)).astype(np.uint8)
    for kind, axis in itertools.product(KINDS, (-1, 0, 1)):
        got = run_gpu(src, kind, axis, op, 1, acc)
        np.testing.assert_array_equal(got, oracle.sequence(src, KINDS[kind], axis, op, 1, acc),
                                      err_msg=f"{kind} axis={axis}")


@pytest.mark.parametrize("acc", [np.float32, np.float64])
def test_float_integer_valued_data_is_exact(acc):
    rng = np.random.default_rng(6)
    src = rng.integers(0, 8, (40, 900)).astype(np.float32)
    for kind, axis in itertools.product(KINDS, (-1, 0, 1)):
        for op, init in (("sum", 2.0), ("min", 5.0), ("max", 1.0)):
            got = run_gpu(src, kind, axis, op, init, acc)
            np.testing.assert_array_equal(got, oracle.sequence(src, KINDS[kind], axis, op, init, acc))
    signs = np.where(rng.random((9, 301)) < 0.5, -1.0, 1.0).astype(np.float32)
    for kind, axis in itertools.product(KINDS, (-1, 0, 1)):
        got = run_gpu(signs, kind, axis, "product", 1.0, acc)
        np.testing.assert_array_equal(got, oracle.sequence(signs, KINDS[kind], axis, "product", 1.0, acc))


def test_random_float_sums_within_grouping_tolerance():
    rng = np.random.default_rng(7)
    src = rng.standard_normal((64, 20000)).astype(np.float32)
    for kind, axis in itertools.product(KINDS, (-1, 0, 1)):
        got = run_gpu(src, kind, axis, "sum", 0.0, np.float32).astype(np.float64)
        ref = oracle.sequence(src.astype(np.float64), KINDS[kind], axis, "sum", 0.0, np.float64)
        scale = np.abs(src).astype(np.float64).sum()
        assert np.max(np.abs(got - ref)) <= 1e-6 * scale, f"{kind} axis={axis}"


def test_strided_and_reversed_views():
    import torch
    rng = np.random.default_rng(8)
    base = rng.integers(0, 9, (12, 40, 90)).astype(np.int32)
    d_base = torch.from_numpy(base).cuda()
    src = d_base[::2, 3:35, 5:85]
    host = base[::2, 3:35, 5:85]
    for kind, axis in itertools.product(KINDS, (-1, 0, 1, 2)):
        shape = out_shape(host.shape, kind, axis)
        big = torch.full((2 * shape[0],) + tuple(shape[1:]), 99, dtype=torch.int32, device="cuda")
        dst = big[::2]
        ctx = iane_amd.make_execution_context()
        op = ctx.prepare(algorithm(kind, axis, "sum", 1), src, dst)
        ctx.run(op, src, dst).wait()
        np.testing.assert_array_equal(dst.cpu().numpy(), oracle.sequence(host, KINDS[kind], axis, "sum", 1, np.int32))
        assert (big[1::2].cpu().numpy() == 99).all()
    # a reversed axis 0 through the C ABI (negative byte stride)
    desc = iane_amd.view_desc(d_base[0])
    desc.ptr = d_base[0].data_ptr() + 89 * 4
    desc.stride_bytes[0] = -4
    out = torch.zeros((40, 90), dtype=torch.int32, device="cuda")
    from iane_amd import _lib
    import ctypes
    plan = ctypes.c_void_p()
    init = ctypes.c_int32(0)
    _lib.call("iane_hip_sequence_plan_create", ctypes.byref(plan), oracle.SEQ_INCLUSIVE, 0, 0, _lib.I32,
              ctypes.byref(init), desc, iane_amd.view_desc(out))
    _lib.call("iane_hip_sequence_run", plan, desc, iane_amd.view_desc(out), None, 0)
    iane_amd.hip_execution(0).wait()
    _lib.load().iane_hip_sequence_plan_destroy(plan)
    np.testing.assert_array_equal(out.cpu().numpy(), np.cumsum(base[0][:, ::-1], axis=1))


def test_empty_sources():
    assert run_gpu(np.zeros((0, 5), np.int32), "reduce", -1, "sum", 7, np.int32).tolist() == [7]
    np.testing.assert_array_equal(run_gpu(np.zeros((0, 5), np.int32), "reduce", 1, "max", -3, np.int32), [-3] * 5)
    assert run_gpu(np.zeros((0, 5), np.int32), "inclusive", -1, "sum", 7, np.int32).shape == (0, 5)


def test_errors():
    import torch
    src = torch.zeros((4, 6), dtype=torch.int32, device="cuda")
    ctx = iane_amd.make_execution_context()
    with pytest.raises(iane_amd.invalid_argument):
        ctx.prepare(iane_amd.reduce_along(0, "sum", 0), src, torch.zeros(6, dtype=torch.int32, device="cuda"))
    with pytest.raises(iane_amd.invalid_argument):
        ctx.prepare(iane_amd.inclusive_scan("sum", 0), src, torch.zeros((4, 5), dtype=torch.int32, device="cuda"))
    with pytest.raises(iane_amd.invalid_argument):
        iane_amd.reduce("average", 0)
    scan = ctx.prepare(iane_amd.inclusive_scan("sum", 0), src, torch.zeros((4, 6), dtype=torch.int32, device="cuda"))
    wide = torch.zeros((4, 12), dtype=torch.int32, device="cuda")
    with pytest.raises(iane_amd.invalid_argument):  # prepared on packed views, run on a strided one
        ctx.run(scan, wide[:, ::2], torch.zeros((4, 6), dtype=torch.int32, device="cuda"))
    op = ctx.prepare(iane_amd.reduce("sum", 0), src, torch.zeros(1, dtype=torch.int32, device="cuda"))
    with pytest.raises(iane_amd.invalid_argument):
        ctx.run(op, torch.zeros((4, 7), dtype=torch.int32, device="cuda"), torch.zeros(1, dtype=torch.int32,
                                                                                        device="cuda"))


def test_full_size_whole_scan_and_reduce_closed_form():
    """256^3 float32 ones: inclusive sum is index + 1 (exact in float64), the reduction is 2^24."""
    import torch
    src = torch.ones((256, 256, 256), dtype=torch.float32, device="cuda")
    total = torch.zeros(1, dtype=torch.float64, device="cuda")
    scan = torch.zeros((256, 256, 256), dtype=torch.float64, device="cuda")
    ctx = iane_amd.make_execution_context()
    r = ctx.prepare(iane_amd.reduce("sum", 0.0), src, total)
    s = ctx.prepare(iane_amd.inclusive_scan("sum", 0.0), src, scan)
    ctx.run(r, src, total)
    ctx.run(s, src, scan).wait()
    assert total.item() == 2.0 ** 24
    flat = scan.flatten()
    assert torch.equal(flat, torch.arange(1, 2 ** 24 + 1, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("acc", [np.float32, np.float64, np.int32])
def test_vector_and_scalar_tile_paths_agree(acc):
    """Lines of several 4096-element tiles whose rows are whole 16-byte vectors take
    the 16-byte tile-total and scan-output kernels; the same data through a view
    4 bytes off alignment takes the per-element kernels. Integer-valued data makes
    every grouping exact, so both must equal the oracle bit for bit (a misplaced
    element or vector shows up as a wrong prefix)."""
    import torch
    rng = np.random.default_rng(11)
    host = rng.integers(0, 8, (5, 3, 13292)).astype(np.float32 if acc != np.int32 else np.int32)
    aligned = torch.from_numpy(host).cuda()
    buf = torch.empty(host.size + 1, dtype=aligned.dtype, device="cuda")
    buf[1:] = aligned.flatten()
    shifted = buf[1:].view(host.shape)
    assert aligned.data_ptr() % 16 == 0 and shifted.data_ptr() % 16 == 4
    for kind, axis in itertools.product(KINDS, (-1, 0)):
        ref = oracle.sequence(host, KINDS[kind], axis, "sum", 3, acc)
        for src in (aligned, shifted):
            dst = torch.full(out_shape(host.shape, kind, axis), 77, dtype=torch.from_numpy(np.zeros(1, acc)).dtype,
                             device="cuda")
            ctx = iane_amd.make_execution_context()
            prepared = ctx.prepare(algorithm(kind, axis, "sum", 3), src, dst)
            ctx.run(prepared, src, dst).wait()
            np.testing.assert_array_equal(dst.cpu().numpy(), ref, err_msg=f"{kind} axis={axis}")
