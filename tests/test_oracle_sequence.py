"""The scan/reduce oracle against the known answers of the reference's
test/unit/prepared/sequence_cases.hpp:31-117 (a 263 x 5 u8 image x % 7 + y:
row and column scans composed into 2-D prefix sums, flat exclusive/inclusive
scans seeded with 11, whole and per-row reductions) and closed forms."""
from __future__ import annotations

import numpy as np

from oracle import oracle

W, H, SEED = 263, 5, 11


def image():
    y, x = np.mgrid[0:H, 0:W]
    return (x % 7 + y).astype(np.uint8)


def test_reference_sequence_case_closed_forms():
    src = image()
    pixel = src.astype(np.int64)
    flat = pixel.ravel()
    exclusive = oracle.sequence(src, oracle.SEQ_EXCLUSIVE, -1, "sum", SEED, np.float64)
    inclusive = oracle.sequence(src, oracle.SEQ_INCLUSIVE, -1, "sum", SEED, np.float64)
    prefix = SEED + np.concatenate([[0], np.cumsum(flat)[:-1]])
    np.testing.assert_array_equal(exclusive.ravel(), prefix)
    np.testing.assert_array_equal(inclusive.ravel(), prefix + flat)
    total = oracle.sequence(src, oracle.SEQ_REDUCE, -1, "sum", SEED, np.float64)
    assert total.shape == (1,) and total[0] == SEED + flat.sum()
    rows = oracle.sequence(src, oracle.SEQ_REDUCE, 0, "sum", SEED, np.float64)
    np.testing.assert_array_equal(rows, SEED + pixel.sum(axis=1))
    columns = oracle.sequence(src, oracle.SEQ_EXCLUSIVE, 1, "sum", SEED, np.float64)
    for x in range(W):
        running = SEED
        for yy in range(H):
            assert columns[yy, x] == running
            running += x % 7 + yy
    # row scan then column scan = 2-D prefix (rectangle) sums
    row_scan = oracle.sequence(src, oracle.SEQ_INCLUSIVE, 0, "sum", 0, np.float64)
    rect = oracle.sequence(row_scan, oracle.SEQ_INCLUSIVE, 1, "sum", 0, np.float64)
    for yy in range(H):
        for x in (0, 1, 100, W - 1):
            assert rect[yy, x] == sum(xx % 7 + r for xx in range(x + 1) for r in range(yy + 1))


def test_integer_wrap_min_max_and_empty():
    a = np.array([200, 100, 7], np.uint8)
    np.testing.assert_array_equal(oracle.sequence(a, oracle.SEQ_INCLUSIVE, -1, "sum", 0, np.uint8), [200, 44, 51])
    np.testing.assert_array_equal(oracle.sequence(a, oracle.SEQ_INCLUSIVE, -1, "product", 1, np.uint8),
                                  [200, (200 * 100) % 256, (200 * 100 * 7) % 256])
    np.testing.assert_array_equal(oracle.sequence(a, oracle.SEQ_EXCLUSIVE, -1, "min", 150, np.int32), [150, 150, 100])
    np.testing.assert_array_equal(oracle.sequence(a, oracle.SEQ_REDUCE, -1, "max", 0, np.int32), [200])
    empty = np.zeros((0, 3), np.float32)
    assert oracle.sequence(empty, oracle.SEQ_REDUCE, -1, "sum", 5.0, np.float32)[0] == 5.0
    np.testing.assert_array_equal(oracle.sequence(empty, oracle.SEQ_REDUCE, 1, "sum", 5.0, np.float32), [5, 5, 5])
