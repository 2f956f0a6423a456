"""Output invariance across every autotuner candidate (bit equality) and a real tuning run."""
from __future__ import annotations

import numpy as np
import pytest

import iane_amd
from iane_amd import autotune, border_mode, convolution_options, deconvolution, precision, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def reset_knobs(tmp_path, monkeypatch):
    monkeypatch.setenv("IANE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    yield
    for key in (list(autotune.VECTOR_SEPARABLE_SPACE) + list(autotune.SEPARABLE_SPACE)
                + list(autotune.STAGED_SEPARABLE_SPACE) + list(autotune.RL_BATCH_SPACE) + ["sep3.engine"]):
        autotune.set(key, -1)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("border", [border_mode.zero, border_mode.clamp])
@pytest.mark.parametrize("width", [132, 130], ids=["x4", "ragged"])
def test_separable_candidates_bit_identical(border, width):
    """Every launch shape of each engine gives the same bits (a width that is a
    multiple of 4 with the zero border runs the barrier-free engine)."""
    import torch
    rng = np.random.default_rng(0)
    vol = dev(rng.random((70, 45, width)).astype(np.float32))
    taps = [dev(synth.gaussian_1d(s, 15).astype(np.float32)) for s in (2.0, 2.5, 3.0)]
    opts = convolution_options(border, precision.fp32)
    for engine, space in ((2, autotune.SEPARABLE_SPACE), (0, autotune.STAGED_SEPARABLE_SPACE),
                          (1, autotune.VECTOR_SEPARABLE_SPACE)):
        with autotune.applied({"sep3.engine": engine}):
            ref = torch.empty_like(vol)
            iane_amd.convolve_separable(vol, taps, ref, opts).wait()
            ref = ref.cpu().numpy()
            for config in autotune.candidates(space):
                out = torch.empty_like(vol)
                with autotune.applied(config):
                    iane_amd.convolve_separable(vol, taps, out, opts).wait()
                np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32),
                                              err_msg=f"engine {engine} {config}")


def _rl_case():
    data = synth.rl_batch_data(count=3, seed=11, edge_range=(24, 70))
    return [dev(o) for o, _ in data], [dev(p) for _, p in data]


def test_fft_candidates_bit_identical_single_and_batched():
    import torch
    obs, psfs = _rl_case()
    results = {}
    for config in autotune.candidates(autotune.RL_BATCH_SPACE):
        with autotune.applied(config):
            single = []
            for o, p in zip(obs, psfs):
                plan = deconvolution.rl_plan(tuple(o.shape), p, deconvolution.FFT)
                x = torch.empty_like(o)
                plan.run(o, x, 5).wait()
                single.append(x.cpu().numpy())
            batch = deconvolution.rl_batch([tuple(o.shape) for o in obs], psfs)
            xs = [torch.empty_like(o) for o in obs]
            batch.run(obs, xs, 5).wait()
            batched = [x.cpu().numpy() for x in xs]
        for a, b in zip(single, batched):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        results[config["fft.lds_elems"]] = single
    first = results[4096]
    for budget, got in results.items():
        for a, b in zip(got, first):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"lds_elems={budget}")


def test_tuning_runs_and_applies_a_candidate():
    import torch
    obs, psfs = _rl_case()
    best = autotune.tune_rl_batch(obs, psfs, iterations=2)
    assert best in autotune.candidates(autotune.RL_BATCH_SPACE)
    assert autotune.get("fft.lds_elems") == best["fft.lds_elems"]
    vol = dev(np.random.default_rng(1).random((40, 40, 64)).astype(np.float32))
    taps = [dev(synth.gaussian_1d(2.5, 15).astype(np.float32))] * 3
    out = torch.empty_like(vol)
    best = autotune.tune_separable(vol, taps, out, convolution_options(border_mode.zero, precision.fp32), repeats=2)
    assert best in autotune.candidates(autotune.SEPARABLE_SPACE)
