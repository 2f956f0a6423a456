"""Autotuner logic on the CPU: candidate order, knob round trips, search and cache."""
from __future__ import annotations

import json

import pytest

from iane_amd import autotune


@pytest.fixture
def cache(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("IANE_AMD_TUNE_CACHE", str(path))
    yield path
    for key in list(autotune.VECTOR_SEPARABLE_SPACE) + list(autotune.FFT_SPACE):
        autotune.set(key, -1)


def test_candidates_cover_the_space_default_first():
    cands = autotune.candidates(autotune.VECTOR_SEPARABLE_SPACE)
    assert len(cands) == 2 * 2 * 4
    assert cands[0] == {"sep3.threads": 512, "sep3.tile_y": 16, "sep3.chunk": 0}
    assert autotune.candidates(autotune.SEPARABLE_SPACE)[0] == {"sep3.waves": 1, "sep3.chunk": 0}
    assert autotune.candidates(autotune.STAGED_SEPARABLE_SPACE)[0] == {"sep3.waves": 2, "sep3.chunk": 0}
    assert len({tuple(sorted(c.items())) for c in cands}) == len(cands)


def test_knob_round_trip_and_reset(cache):
    autotune.set("fft.lds_elems", 2048)
    assert autotune.get("fft.lds_elems") == 2048
    with autotune.applied({"fft.lds_elems": 1024}):
        assert autotune.get("fft.lds_elems") == 1024
    assert autotune.get("fft.lds_elems") == 2048
    autotune.set("fft.lds_elems", -1)
    assert autotune.get("fft.lds_elems") == -1


def test_search_picks_fastest_and_caches(cache):
    seen = []

    def measure(config):
        seen.append(dict(config))
        assert autotune.get("fft.lds_elems") == config["fft.lds_elems"]  # applied while measured
        return {4096: 3.0, 3072: 2.5, 2048: 1.0, 1536: 1.5, 1024: 2.0}[config["fft.lds_elems"]]

    best = autotune.search("k", autotune.FFT_SPACE, measure)
    assert best == {"fft.lds_elems": 2048}
    assert autotune.get("fft.lds_elems") == 2048
    assert len(seen) == len(autotune.FFT_SPACE["fft.lds_elems"])
    stored = json.loads(cache.read_text())["k"]
    assert stored["config"] == {"fft.lds_elems": 2048}
    # cached: applied without measuring
    autotune.set("fft.lds_elems", -1)
    again = autotune.search("k", autotune.FFT_SPACE, lambda c: pytest.fail("measured a cached key"))
    assert again == best and autotune.get("fft.lds_elems") == 2048


def test_ties_keep_the_earlier_candidate(cache):
    best = autotune.search("tie", autotune.FFT_SPACE, lambda c: 1.0, use_cache=False)
    assert best == autotune.candidates(autotune.FFT_SPACE)[0]


def test_stale_cache_entry_outside_the_space_is_ignored(cache):
    cache.write_text(json.dumps({"k": {"config": {"fft.lds_elems": 777}, "timings_ms": []}}))
    best = autotune.search("k", autotune.FFT_SPACE, lambda c: float(c["fft.lds_elems"]))
    assert best == {"fft.lds_elems": 1024}
