"""Known-answer convolution cases restated from the reference's own tests, as
numpy inputs and expected outputs (arrays in C order: IANE axis 0 = last axis).

  * pattern_cases   — test/unit/algorithms/convolution_cases.hpp:12-34 with
                      algorithm_test_views.hpp:72-103 (3^N input, kernel
                      [2,-1,0] along the last IANE axis, zero border), N = 2, 3, 4
  * scalar_case     — test/unit/algorithms/test_convolution.cpp:10-23
  * crop_cases      — test/unit/algorithms/test_convolution.cpp:25-50
  * texture_case    — test/unit/algorithms/test_convolution_cuda.cu:53-86 inputs
                      (the expected values there are the CPU path's results)
"""
from __future__ import annotations

import numpy as np

CLAMP, ZERO = 0, 1


def pattern_case(rank: int):
    """Returns (src_first, src_second, kernel, expected_first, expected_second).

    The reference element is a pair {v, -v}; the pair's sum is componentwise,
    so each component is an independent integer convolution."""
    count = 3 ** rank
    values = np.arange(1, count + 1, dtype=np.int32)
    src = values.reshape((3,) * rank)
    kshape = [1] * rank
    kshape[0] = 3  # IANE axis rank-1 is numpy axis 0
    kernel = np.array([2, -1, 0], dtype=np.int32).reshape(kshape)
    plane = 3 ** (rank - 1)
    index = np.arange(count)
    layer = index // plane
    value = index + 1
    nxt = np.where(layer == 2, 0, value + plane)
    first = (2 * nxt - value).astype(np.int32).reshape(src.shape)
    return src, -src, kernel, first, -first


def scalar_case():
    src = np.array([1, 2, 3], dtype=np.int32)
    kernel = np.array([1, 0, 0], dtype=np.int32)
    return src, kernel, np.array([2, 3, 0], dtype=np.int32)


def crop_cases():
    """u8 4x2 source 10x + 100y, crop corner (1,0) size (2,2), float kernel [1,0,0]."""
    x = np.arange(4)
    y = np.arange(2)
    source = (10 * x[None, :] + 100 * y[:, None]).astype(np.uint8)  # shape (2, 4)
    crop = source[:, 1:3]
    kernel = np.array([[1.0, 0.0, 0.0]], dtype=np.float32)  # extent {3, 1}
    return source, crop, kernel, {CLAMP: np.array([[20, 20], [120, 120]], np.uint8),
                                  ZERO: np.array([[20, 0], [120, 0]], np.uint8)}


def texture_case():
    """5x4 float source i-6, 3x3 kernel (i-3)/8, crop origin (1,1) size 3x3."""
    source = (np.arange(20, dtype=np.float32) - 6).reshape(4, 5)
    kernel = ((np.arange(9, dtype=np.float32) - 3) / 8.0).reshape(3, 3)
    crop = source[1:4, 1:4]
    return source, crop, kernel
