"""Batched Richardson–Lucy: many uneven volumes with their own PSFs in one
launch per pass, replayed from a hipGraph.

Parity: every volume of the batch equals the single-volume FFT plan bit for bit
(same per-volume passes), with and without the graph, on repeated graph replays
with eager work in between, and its log-likelihood history equals the single
plan's. The single plan itself is pinned against the FP64 oracle in test_rl_gpu.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from iane_amd import deconvolution, synth

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def _single(observed, psf, iterations, stream_policy=None):
    import torch
    plan = deconvolution.rl_plan(observed.shape, dev(psf), deconvolution.FFT)
    est = torch.empty_like(dev(observed))
    ll = torch.zeros(max(iterations, 1), dtype=torch.float64, device="cuda")
    plan.run(dev(observed), est, iterations, ll_history=ll).wait()
    return host(est), host(ll)[:iterations]


def _stream_policy():
    import ctypes
    from iane_amd import _lib, hip_policy
    s = ctypes.c_void_p()
    _lib.call("iane_hip_stream_create", ctypes.byref(s))
    return hip_policy(s.value)


def test_batch_equals_single_plans_with_and_without_graph():
    import torch
    data = synth.rl_batch_data(count=6, seed=11, edge_range=(20, 60))
    iterations = 7
    policy = _stream_policy()
    batch = deconvolution.rl_batch([o.shape for o, _ in data], [dev(p) for _, p in data], policy=policy)
    obs = [dev(o) for o, _ in data]
    results = {}
    for graph in (False, True):
        est = [torch.empty_like(o) for o in obs]
        ll = torch.zeros((len(data), iterations), dtype=torch.float64, device="cuda")
        batch.run(obs, est, iterations, ll_history=ll, use_graph=graph).wait()
        results[graph] = ([host(e) for e in est], host(ll))
    for v, (o, p) in enumerate(data):
        ref, ref_ll = _single(o, p, iterations)
        for graph in (False, True):
            np.testing.assert_array_equal(results[graph][0][v].view(np.uint32), ref.view(np.uint32))
            np.testing.assert_array_equal(results[graph][1][v], ref_ll)


def test_graph_replays_are_identical():
    import torch
    data = synth.rl_batch_data(count=4, seed=12, edge_range=(24, 48))
    policy = _stream_policy()
    batch = deconvolution.rl_batch([o.shape for o, _ in data], [dev(p) for _, p in data], policy=policy)
    obs = [dev(o) for o, _ in data]
    outs = []
    for replay in range(3):
        est = [torch.empty_like(o) for o in obs]
        batch.run(obs, est, 5, use_graph=True).wait()
        outs.append([host(e) for e in est])
        # eager work between replays
        scratch = torch.randn(1000, device="cuda")
        torch.cuda.synchronize()
        del scratch
    for replay in outs[1:]:
        for a, b in zip(outs[0], replay):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_c4_sized_batch_spot_checks():
    """64 volumes, edges 32..192, own PSFs: spot-check three volumes against single plans."""
    import torch
    data = synth.rl_batch_data(count=64, seed=4)
    policy = _stream_policy()
    batch = deconvolution.rl_batch([o.shape for o, _ in data], [dev(p) for _, p in data], policy=policy)
    obs = [dev(o) for o, _ in data]
    est = [torch.empty_like(o) for o in obs]
    batch.run(obs, est, 3, use_graph=True).wait()
    for v in (0, 31, 63):
        ref, _ = _single(data[v][0], data[v][1], 3)
        np.testing.assert_array_equal(host(est[v]).view(np.uint32), ref.view(np.uint32))
