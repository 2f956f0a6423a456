"""What the trained warm start gains over plain RL (committed weights), and its plumbing.

Gain: on held-out volumes of the training distribution (blobs, PSF sigma
(3, 1.5, 1.5) on 63x31x31, peak 1000, Poisson), RL started from the network's
prediction must reach the Poisson log-likelihood of plain 200-iteration RL in
at most 50 iterations (>= 4x fewer), and the prediction alone must already be
far closer to that likelihood than the plain start max(y, eps).
"""
from __future__ import annotations

import numpy as np
import pytest

from iane_amd import learned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    if not learned.WEIGHTS.exists():
        pytest.skip("no trained weights")
    return learned.warm_start_net().load()


def test_warm_start_reaches_plain_200_likelihood_4x_sooner(net):
    import torch
    psf = torch.from_numpy(learned.psf_kernel()).cuda()
    ys = list(learned.observed_volumes(2, (128, 128, 128), seed=77))
    results = learned.iterations_to_likelihood(net, ys, psf, iterations=200, max_iterations=60)
    for r in results:
        print(r)
        assert 0 <= r["k"] <= 50
        gap_warm = r["target"] - r["ll_warm0"]
        gap_plain = r["target"] - r["ll_plain0"]
        assert gap_warm < 0.05 * gap_plain


def test_uneven_extents_are_zero_padded_then_cropped(net):
    """warm_start on extents not divisible by 4 equals the network on the
    zero-padded volume, cropped (bit-exact)."""
    import torch
    rng = np.random.default_rng(3)
    y = (rng.poisson(50.0, (30, 45, 62))).astype(np.float32)
    yd = torch.from_numpy(y).cuda()
    x0 = learned.warm_start(net, yd).cpu().numpy()
    padded = np.zeros((32, 48, 64), np.float32)
    padded[:30, :45, :62] = y
    xp = learned.warm_start(net, torch.from_numpy(padded).cuda()).cpu().numpy()
    np.testing.assert_array_equal(x0.view(np.uint32), xp[:30, :45, :62].view(np.uint32))
    assert np.all(x0 >= 1e-6)


def test_batched_time_to_likelihood_runs(net):
    import torch
    psf = torch.from_numpy(learned.psf_kernel()).cuda()
    ys = list(learned.observed_volumes(2, (64, 96, 96), seed=78))
    out = learned.batch_time_to_likelihood(net, ys, [psf, psf], iterations=200, max_iterations=80, repeats=1)
    print(out)
    assert out["k_max"] >= 0 and out["plain_ms"] > 0 and out["learned_ms"] > 0
    assert out["speedup"] > 1.0
    # the timed run stops at the first check where every volume has reached its target
    assert out["k_max"] % out["check_every"] == 0
    assert max(out["k_per_volume"]) <= out["k_max"] < max(out["k_per_volume"]) + out["check_every"]


def test_timed_gain_at_256x256x128(net):
    """Four 256x256x128 volumes: the learned path, timed with its stopping rule
    (likelihood read back every 2 iterations and compared with the plain-RL
    target), takes at most a quarter of the GPU time of plain 200-iteration RL."""
    import torch
    psf = torch.from_numpy(learned.psf_kernel()).cuda()
    ys = list(learned.observed_volumes(4, (128, 256, 256), seed=3))
    out = learned.batch_time_to_likelihood(net, ys, [psf] * 4, iterations=200, max_iterations=120, repeats=2)
    print(out)
    assert 0 < out["k_max"] <= 50
    assert out["speedup"] >= 4.0
