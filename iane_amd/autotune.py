"""Launch-shape autotuner (no reference counterpart).

The knobs are process-wide values that kernels read at launch (or plans read
at creation) through ``iane_hip_tune_set`` / ``iane_hip_tune_get``. No knob
changes a result: every candidate computes each output sample with the same
operations in the same order (tests check bit equality across candidates),
so tuning only trades occupancy, LDS use and tile overlap.

    sep3.waves      1 | 2 (engine 2)       waves per SIMD the matrix engines' registers are fitted to
                    2 | 3 (engine 0)
    sep3.chunk      0 (auto) | 16 .. 80    z planes streamed per workgroup (every engine)
    sep3.threads    256 | 512              threads per workgroup (vector engine)
    sep3.tile_y     16 | 32                output rows per workgroup tile (vector engine)

    fft.lds_elems   1024 .. 4096           LDS budget per buffer (complex elements) of the
                                           FFT line passes and RL row steps: lines per workgroup

``sep3.engine`` (2: x/y filters on FP32 MFMA with per-wave operand loads, the
default for zero borders and 4-aligned rows; 0: the same on LDS-staged planes;
1: vector ALUs) is not tuned: the engines sum the same FP32 products in
different orders, so it changes bits.
Direct-convolution switches, read at launch and not searched (every setting
gives the same bits; tests compare them):

    conv.stream         1 | 0              z-marching kernels for small 3-D kernels (0: tiled kernel)
    conv.column         1 | 0              register column kernel for 3^3 kernels on 4-aligned rows
                                           (0: the LDS streaming kernel)
    conv.column_rows    4 | 2              output rows per lane of the column kernel
    conv.column_chunk   0 (auto) | >= 1    output planes per column workgroup
    conv.column_dot2    1 | 0              u16 column sums on packed v_dot2_u32_u16 pairs
    conv.column_alternate 1 | 0            exact integer column sums: odd z chunks march down
    conv.plane          1 | 0              register plane kernel for square single-plane kernels
                                           (3 .. 9 taps) on 4-aligned rows (0: the tiled kernel)
    conv.plane_rows     0 (auto) | 1, 2, 4 output rows per lane of the plane kernel
    conv.stream_chunk   0 (auto) | 1 .. 64 output planes per streaming workgroup
    conv.stream_unroll  1 | 0              fully unrolled FP32 3^3 taps (0: row loop)

Choices are cached per (device, workload) in a JSON file: ``$IANE_AMD_TUNE_CACHE``
or ``~/.cache/iane_amd/tune.json``. A cached choice is applied without timing.
"""
from __future__ import annotations

import contextlib
import ctypes
import itertools
import json
import os
import pathlib

from . import _lib

SEPARABLE_SPACE = {"sep3.waves": [1, 2], "sep3.chunk": [0, 32, 48, 64, 80]}  # barrier-free matrix engine (default)
STAGED_SEPARABLE_SPACE = {"sep3.waves": [2, 3], "sep3.chunk": [0, 16, 21, 24, 32, 43, 64]}  # LDS-staged matrix engine
VECTOR_SEPARABLE_SPACE = {"sep3.threads": [512, 256], "sep3.tile_y": [16, 32], "sep3.chunk": [0, 16, 32, 64]}
FFT_SPACE = {"fft.lds_elems": [4096, 3072, 2048, 1536, 1024]}
RL_BATCH_SPACE = FFT_SPACE


def get(key: str) -> int:
    """Current value of a knob (-1 = unset, the kernel's default)."""
    value = ctypes.c_int()
    _lib.call("iane_hip_tune_get", key.encode(), ctypes.byref(value))
    return value.value


def set(key: str, value: int) -> None:  # noqa: A001 - mirrors iane_hip_tune_set
    """Set a knob; a negative value restores the default."""
    _lib.call("iane_hip_tune_set", key.encode(), int(value))


@contextlib.contextmanager
def applied(config: dict):
    """Apply ``config`` for the duration of a with-block, then restore."""
    old = {k: get(k) for k in config}
    try:
        for k, v in config.items():
            set(k, v)
        yield
    finally:
        for k, v in old.items():
            set(k, v)


def candidates(space: dict) -> list[dict]:
    """Every combination of a knob space, in a fixed order (the first is the default)."""
    keys = list(space)
    return [dict(zip(keys, values)) for values in itertools.product(*(space[k] for k in keys))]


# ------------------------------------------------------------------- cache --
def cache_path() -> pathlib.Path:
    env = os.environ.get("IANE_AMD_TUNE_CACHE")
    return pathlib.Path(env) if env else pathlib.Path.home() / ".cache" / "iane_amd" / "tune.json"


def _load_cache() -> dict:
    path = cache_path()
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, config: dict, timings: list) -> None:
    path = cache_path()
    data = _load_cache()
    data[key] = {"config": config, "timings_ms": timings}
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass  # a read-only home only loses the cache


def device_name() -> str:
    import torch
    return torch.cuda.get_device_name() if torch.cuda.is_available() else "cpu"


# ------------------------------------------------------------------ search --
def search(key: str, space: dict, measure, *, use_cache: bool = True) -> dict:
    """Time ``measure(config) -> ms`` for every candidate, keep the fastest (ties:
    the earlier candidate), apply and cache it. A cached key is applied as is."""
    if use_cache:
        hit = _load_cache().get(key)
        if hit and set_of(hit["config"]) <= set_of_space(space):
            for k, v in hit["config"].items():
                set(k, v)
            return dict(hit["config"])
    best, best_ms, timings = None, float("inf"), []
    for config in candidates(space):
        with applied(config):
            ms = float(measure(config))
        timings.append([config, ms])
        if ms < best_ms:
            best, best_ms = config, ms
    for k, v in best.items():
        set(k, v)
    _store_cache(key, best, timings)
    return dict(best)


def set_of(config: dict) -> set:
    return {(k, v) for k, v in config.items()}


def set_of_space(space: dict) -> set:
    return {(k, v) for k, values in space.items() for v in values}


def _event_timer(stream: int):
    start, stop = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.call("iane_hip_event_create", ctypes.byref(start))
    _lib.call("iane_hip_event_create", ctypes.byref(stop))

    def time(fn, repeats: int) -> float:
        fn()  # warm-up: plan / graph creation and first-touch
        _lib.call("iane_hip_stream_sync", stream)
        _lib.call("iane_hip_event_record", start, stream)
        for _ in range(repeats):
            fn()
        _lib.call("iane_hip_event_record", stop, stream)
        ms = ctypes.c_float()
        _lib.call("iane_hip_event_elapsed_ms", ctypes.byref(ms), start, stop)
        return ms.value / repeats

    def close():
        _lib.call("iane_hip_event_destroy", start)
        _lib.call("iane_hip_event_destroy", stop)

    return time, close


def tune_separable(src, taps, dst, options=None, policy=None, repeats: int = 10, use_cache: bool = True) -> dict:
    """Pick the fused separable kernel's launch shape for this image shape and tap count."""
    import iane_amd
    policy = policy or iane_amd.hip_policy()
    stream = policy.handle()
    engine = get("sep3.engine")
    engine = 2 if engine < 0 else engine
    space = {0: STAGED_SEPARABLE_SPACE, 1: VECTOR_SEPARABLE_SPACE}.get(engine, SEPARABLE_SPACE)
    key = f"{device_name()}|sep3.engine{engine}|{tuple(src.shape)}|{tuple(int(t.numel()) for t in taps)}"
    time, close = _event_timer(stream)
    try:
        return search(key, space,
                      lambda _: time(lambda: iane_amd.convolve_separable(src, taps, dst, options, policy), repeats),
                      use_cache=use_cache)
    finally:
        close()


def tune_rl_batch(observed, psfs, iterations: int = 4, epsilon: float = 1e-6, policy=None,
                  use_cache: bool = True) -> dict:
    """Pick the FFT/RL row LDS budget for a batch of RL problems (a batch reads
    the knob when it is created: create it after tuning)."""
    import torch
    import iane_amd
    from . import deconvolution
    policy = policy or iane_amd.hip_policy()
    stream = policy.handle()
    shapes = [tuple(o.shape) for o in observed]
    key = f"{device_name()}|rl_batch|{shapes}|{[tuple(p.shape) for p in psfs]}"
    estimates = [torch.empty_like(o) for o in observed]
    time, close = _event_timer(stream)

    def measure(_config):
        batch = deconvolution.rl_batch(shapes, psfs, epsilon, policy=policy)
        ms = time(lambda: batch.run(observed, estimates, iterations), 2)
        del batch
        return ms

    try:
        return search(key, RL_BATCH_SPACE, measure, use_cache=use_cache)
    finally:
        close()


def tune_rl_plan(shape, psf, iterations: int = 4, epsilon: float = 1e-6, policy=None, use_cache: bool = True) -> dict:
    """Pick the FFT/RL row LDS budget for single-volume FFT RL plans of this shape."""
    import torch
    import iane_amd
    from . import deconvolution
    policy = policy or iane_amd.hip_policy()
    stream = policy.handle()
    key = f"{device_name()}|rl_plan|{tuple(shape)}|{tuple(psf.shape)}"
    observed = torch.ones(tuple(shape), dtype=torch.float32, device="cuda")
    torch.cuda.current_stream().synchronize()
    estimate = torch.empty_like(observed)
    time, close = _event_timer(stream)

    def measure(_config):
        plan = deconvolution.rl_plan(tuple(shape), psf, deconvolution.FFT, epsilon=epsilon, policy=policy)
        ms = time(lambda: plan.run(observed, estimate, iterations), 2)
        del plan
        return ms

    try:
        return search(key, FFT_SPACE, measure, use_cache=use_cache)
    finally:
        close()
