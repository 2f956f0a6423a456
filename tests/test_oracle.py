"""Pins the CPU oracle (CPU only, no GPU).

Each oracle function is checked against (a) the reference's own known answers,
(b) the reference itself compiled from its headers (oracle/_ref, in the
container that has the reference), (c) the committed golden fixtures that the
reference wrote, and (d) what the mathematics fixes (brute force, closed forms,
separability), so that a dropped term, a flipped kernel, a wrong border or a
wrong rounding rule fails at least one of them.
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

from oracle import oracle
from tests import cases

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
DT = [np.uint8, np.uint16, np.int32, np.float32, np.float64]


# ---------------------------------------------------------- known answers --
@pytest.mark.parametrize("rank", [2, 3, 4])
def test_pattern_known_answer(rank):
    src, src2, kernel, first, second = cases.pattern_case(rank)
    out = oracle.convolve(src, kernel, np.int32, oracle.ZERO)
    np.testing.assert_array_equal(out, first)
    out2 = oracle.convolve(np.ascontiguousarray(src2), kernel, np.int32, oracle.ZERO)
    np.testing.assert_array_equal(out2, second)


def test_scalar_coordinates_known_answer():
    src, kernel, expected = cases.scalar_case()
    np.testing.assert_array_equal(oracle.convolve(src, kernel, np.int32, oracle.ZERO), expected)


@pytest.mark.parametrize("border", [oracle.CLAMP, oracle.ZERO])
def test_crop_flip_known_answer(border):
    _, crop, kernel, expected = cases.crop_cases()
    out = oracle.convolve(crop, kernel, np.uint8, border)
    np.testing.assert_array_equal(out, expected[border])


def test_invalid_kernel_rejected():
    src = np.zeros((4, 4), np.float32)
    with pytest.raises(ValueError):
        oracle.convolve(src, np.ones((2, 3), np.float32), np.float32)
    with pytest.raises(ValueError):
        oracle.convolve(src, np.ones((3,), np.float32), np.float32)


# ----------------------------------------------------- mathematics pins ----
def _brute(src, kernel, border):
    """Independent numpy brute force (float64 sums in kernel storage order)."""
    rank = src.ndim
    out = np.zeros(src.shape)
    c = [k // 2 for k in kernel.shape]
    for q in np.ndindex(*kernel.shape[::-1]):  # IANE storage order: numpy last axis fastest
        q = q[::-1]
        idx = []
        valid = np.ones(src.shape, bool)
        for a in range(rank):
            coord = np.arange(src.shape[a]) + c[a] - q[a]
            shape = [1] * rank
            shape[a] = -1
            inside = (coord >= 0) & (coord < src.shape[a])
            valid &= inside.reshape(shape)
            idx.append(np.clip(coord, 0, src.shape[a] - 1).reshape(shape))
        sample = src[tuple(idx)].astype(np.float64)
        if border == oracle.ZERO:
            sample = np.where(valid, sample, 0.0)
        out = out + sample * float(kernel[q])
    return out


@pytest.mark.parametrize("border", [oracle.CLAMP, oracle.ZERO])
@pytest.mark.parametrize("shape,kshape", [((7,), (3,)), ((5, 6), (3, 5)), ((4, 5, 6), (3, 1, 5)),
                                          ((3, 4, 3, 5), (3, 3, 1, 3))])
def test_integer_bruteforce(border, shape, kshape):
    rng = np.random.default_rng(11)
    src = rng.integers(-50, 50, shape).astype(np.int32)
    ker = rng.integers(-5, 6, kshape).astype(np.int32)
    out = oracle.convolve(src, ker, np.int32, border)
    np.testing.assert_array_equal(out, _brute(src, ker, border).astype(np.int64))


def test_delta_kernel_shifts():
    rng = np.random.default_rng(3)
    src = rng.random((6, 7, 8)).astype(np.float32)
    ker = np.zeros((3, 3, 5), np.float32)
    ker[1, 1, 4] = 1.0  # q = (4, 1, 1) in IANE order: sample at x + 2 - 4 = x - 2
    out = oracle.convolve(src, ker, np.float32, oracle.ZERO)
    expected = np.zeros_like(src)
    expected[:, :, 2:] = src[:, :, :-2]
    np.testing.assert_array_equal(out, expected)


def test_box_on_constant_clamp():
    src = np.full((9, 10), 3.0, np.float32)
    ker = np.full((5, 3), 0.5, np.float32)
    out = oracle.convolve(src, ker, np.float64, oracle.CLAMP)
    np.testing.assert_array_equal(out, np.full(src.shape, 3.0 * 0.5 * 15))


@pytest.mark.parametrize("border", [oracle.CLAMP, oracle.ZERO])
def test_separable_equals_dense(border):
    rng = np.random.default_rng(5)
    src = rng.random((9, 8, 10))
    taps = [rng.random(k) for k in (3, 5, 7)]  # IANE axes 0,1,2
    dense = np.multiply.outer(np.multiply.outer(taps[2], taps[1]), taps[0])
    out = oracle.convolve(src, dense, np.float64, border)
    step = src
    for axis, t in enumerate(taps):
        kshape = [1, 1, 1]
        kshape[2 - axis] = t.size
        step = oracle.convolve(np.ascontiguousarray(step), t.reshape(kshape), np.float64, border)
    np.testing.assert_allclose(out, step, rtol=1e-13, atol=1e-13)


def test_output_conversions():
    """u8 rounds half up and saturates; other integers truncate and saturate; NaN -> 0."""
    src = np.array([-1.0, 0.49, 0.5, 1.5, 254.5, 300.0, np.nan], np.float64)
    ker = np.ones(1, np.float64)
    np.testing.assert_array_equal(oracle.convolve(src, ker, np.uint8),
                                  np.array([0, 0, 1, 2, 255, 255, 0], np.uint8))
    np.testing.assert_array_equal(oracle.convolve(src, ker, np.uint16),
                                  np.array([0, 0, 0, 1, 254, 300, 0], np.uint16))
    big = np.array([-3e9, -2.7, 2.7, 3e9], np.float64)
    np.testing.assert_array_equal(oracle.convolve(big, ker, np.int32),
                                  np.array([-2**31, -2, 2, 2**31 - 1], np.int32))


def test_policy_partition_matches_serial():
    rng = np.random.default_rng(9)
    src = rng.random((13, 17)).astype(np.float32)
    ker = rng.random((3, 5)).astype(np.float32)
    serial = oracle.convolve(src, ker, np.float32, oracle.CLAMP)
    threaded = oracle.convolve(src, ker, np.float32, oracle.CLAMP, threads=3, grain=4)
    np.testing.assert_array_equal(serial, threaded)
    with pytest.raises(ValueError):
        oracle.convolve(src, ker, np.float32, oracle.CLAMP, threads=0)


# ----------------------------------------------- reference and fixtures ----
def _random_case(rng, rank, sdt, kdt, ddt):
    shape = tuple(int(x) for x in rng.integers(1, 7, rank))
    kshape = tuple(int(x) for x in rng.choice([1, 3, 5], rank))
    if np.issubdtype(sdt, np.integer):
        hi = {np.uint8: 256, np.uint16: 65536, np.int32: 1000}[sdt]
        lo = -1000 if sdt == np.int32 else 0
        src = rng.integers(lo, hi, shape).astype(sdt)
    else:
        src = (rng.random(shape) * 4 - 1).astype(sdt)
    if np.issubdtype(kdt, np.integer):
        ker = rng.integers(-3, 4, kshape).astype(kdt)
    else:
        ker = (rng.random(kshape) - 0.3).astype(kdt)
    return src, ker


@pytest.mark.skipif(not oracle.have_reference(), reason="oracle/_ref not built (reference headers absent)")
@pytest.mark.parametrize("sdt", DT)
@pytest.mark.parametrize("kdt", [np.int32, np.float32, np.float64])
def test_oracle_matches_reference_build(sdt, kdt):
    rng = np.random.default_rng(hash((str(sdt), str(kdt))) % 2**32)
    for rank in (1, 2, 3, 4):
        for ddt in DT:
            for border in (oracle.CLAMP, oracle.ZERO):
                src, ker = _random_case(rng, rank, sdt, kdt, ddt)
                expected = oracle.convolve_ref(src, ker, ddt, border)
                got = oracle.convolve(src, ker, ddt, border)
                np.testing.assert_array_equal(got.view(np.uint8), expected.view(np.uint8))


@pytest.mark.skipif(not oracle.have_reference(), reason="oracle/_ref not built (reference headers absent)")
def test_oracle_matches_reference_strided_views():
    rng = np.random.default_rng(21)
    base = rng.random((9, 11, 13)).astype(np.float32)
    views = [base[1:8, 2:9, 3:12], base[::2, ::-1, 1::3], base[:, 5, :]]
    for src in views:
        ker = rng.random((3,) * src.ndim).astype(np.float32)
        for border in (oracle.CLAMP, oracle.ZERO):
            np.testing.assert_array_equal(oracle.convolve(src, ker, np.float32, border),
                                          oracle.convolve_ref(src, ker, np.float32, border))


def _golden_files():
    return sorted(GOLDEN.glob("conv_*.npz"))


@pytest.mark.parametrize("path", _golden_files(), ids=lambda p: p.stem)
def test_oracle_matches_golden(path):
    data = np.load(path, allow_pickle=False)
    src, ker, expected = data["src"], data["kernel"], data["expected"]
    border = int(data["border"])
    got = oracle.convolve(src, ker, expected.dtype, border)
    np.testing.assert_array_equal(got.view(np.uint8), expected.view(np.uint8))


def test_golden_fixtures_present():
    assert len(_golden_files()) >= 20, "run tests/golden/make_golden.py where the reference is available"
