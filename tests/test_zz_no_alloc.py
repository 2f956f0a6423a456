"""The allocation-free RL ratio entry point (iane_hip_rl_ratio_noalloc) against
iane_hip_rl_ratio with a caller workspace: the same kernels, so the same bits
for the ratio image and the log-likelihood, on two calls in a row (the second
on fewer elements, so stale partial sums of the first call would show)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from iane_amd import _lib, deconvolution

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    y = rng.poisson(7.0, n).astype(np.float32)
    p = (rng.random(n) * 12).astype(np.float32)
    p[: n // 50] = 0.0  # below the epsilon floor
    return y, p


def _noalloc(dy, dp, eps):
    import torch
    ratio = torch.empty_like(dp)
    ll = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.call("iane_hip_rl_ratio_noalloc", dy.data_ptr(), dp.data_ptr(), ratio.data_ptr(), dy.numel(),
              ctypes.c_float(eps), ll.data_ptr(), stream)
    torch.cuda.synchronize()
    return ratio.cpu().numpy(), float(ll.cpu().numpy()[0])


def _with_workspace(dy, dp, eps):
    import torch
    ratio = torch.empty_like(dp)
    ll = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    deconvolution.rl_ratio(dy, dp, ratio, eps, ll).wait()
    return ratio.cpu().numpy(), float(ll.cpu().numpy()[0])


def test_noalloc_ratio_matches_workspace_entry_point():
    import torch
    for n, seed in ((3_000_017, 1), (40_961, 2)):  # 1024 blocks, then 161
        y, p = _inputs(n, seed)
        dy, dp = torch.from_numpy(y).cuda(), torch.from_numpy(p).cuda()
        r_free, ll_free = _noalloc(dy, dp, 1e-6)
        r_ws, ll_ws = _with_workspace(dy, dp, 1e-6)
        np.testing.assert_array_equal(r_free.view(np.uint32), r_ws.view(np.uint32))
        assert ll_free == ll_ws
        floor = np.where(p > np.float32(1e-6), p, np.float32(1e-6))
        np.testing.assert_array_equal(r_free, y / floor)


def test_noalloc_ratio_without_likelihood_and_empty():
    import torch
    y, p = _inputs(1000, 3)
    dy, dp = torch.from_numpy(y).cuda(), torch.from_numpy(p).cuda()
    ratio = torch.empty_like(dp)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.call("iane_hip_rl_ratio_noalloc", dy.data_ptr(), dp.data_ptr(), ratio.data_ptr(), 1000,
              ctypes.c_float(1e-6), 0, stream)
    _lib.call("iane_hip_rl_ratio_noalloc", 0, 0, 0, 0, ctypes.c_float(1e-6), 0, stream)
    torch.cuda.synchronize()
    floor = np.where(p > np.float32(1e-6), p, np.float32(1e-6))
    np.testing.assert_array_equal(ratio.cpu().numpy(), y / floor)
    with pytest.raises(_lib.invalid_argument):
        _lib.call("iane_hip_rl_ratio_noalloc", 0, 0, 0, 5, ctypes.c_float(1e-6), 0, stream)
