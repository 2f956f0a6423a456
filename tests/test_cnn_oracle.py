"""Pins the FP64 oracle of the warm-start network (CPU only)."""
from __future__ import annotations

import numpy as np

from oracle import cnn_oracle as cnn
from oracle import oracle


def test_conv3d_equals_sum_of_reference_convolves_with_flipped_kernels():
    """One layer = sum over input channels of the reference convolve with the flipped weight, zero border."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 6, 7, 3))
    w = rng.standard_normal((2, 27, 3))
    out = cnn.conv3d(x, w)
    for co in range(2):
        ref = np.zeros((5, 6, 7))
        for ci in range(3):
            k = w[co, :, ci].reshape(3, 3, 3)[::-1, ::-1, ::-1].copy()
            ref += oracle.convolve(np.ascontiguousarray(x[..., ci]), k, np.float64, oracle.ZERO)
        np.testing.assert_allclose(out[..., co], ref, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 6, 8, 3))
    y = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose(np.sum(cnn.pool2(x) * y), np.sum(x * cnn.pool2_backward(y)))
    assert np.isclose(np.sum(cnn.up2(y) * x), np.sum(y * cnn.up2_backward(x)))


def test_input_gradient_is_the_adjoint_of_conv3d():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 6, 3))
    w = rng.standard_normal((4, 27, 3))
    dy = rng.standard_normal((4, 5, 6, 4))
    assert np.isclose(np.sum(cnn.conv3d(x, w) * dy), np.sum(x * cnn.conv3d_input_grad(dy, w)))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = cnn.random_params(4)
    u = rng.random((8, 8, 8)) * 4
    target = u + rng.standard_normal(u.shape) * 0.3
    _, grads = cnn.backward(params, u, target)
    h = 1e-6
    for name in ("c1a", "c2a", "c3b", "c1c"):
        w, b = params[name]
        for idx in [(0, 13, 0), (w.shape[0] - 1, 0, w.shape[2] - 1), (w.shape[0] // 2, 26, w.shape[2] // 2)]:
            wp, wm = w.copy(), w.copy()
            wp[idx] += h
            wm[idx] -= h
            lp = cnn.loss({**params, name: (wp, b)}, u, target)
            lm = cnn.loss({**params, name: (wm, b)}, u, target)
            fd = (lp - lm) / (2 * h)
            an = grads[name][0][idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)) + 1e-9, (name, idx, fd, an)
        bp, bm = b.copy(), b.copy()
        bp[0] += h
        bm[0] -= h
        fd = (cnn.loss({**params, name: (w, bp)}, u, target) - cnn.loss({**params, name: (w, bm)}, u, target)) / (2 * h)
        assert abs(fd - grads[name][1][0]) <= 1e-5 * max(1.0, abs(fd)) + 1e-9


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 3.14159], np.float32)
    r = cnn.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -6        # tie -> even (odd lower neighbour)
    assert abs(r[3] - 3.140625) < 1e-9
