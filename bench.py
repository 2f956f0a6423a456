#!/usr/bin/env python3
"""Benchmark of the IANE convolution hot path on one MI355X.

Headline workload (the one the metric is quoted on): 3-D FP32 convolution of a
256^3 volume (U[0,1) noise + 200 Gaussian blobs, seed 2) with the separable
15x15x15 Gaussian (sigma 2.5), zero border. One step = one
iane_hip_convolve_separable call over the whole volume, inputs resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]

Prints ONE JSON line. ``roofline`` is for the dominant kernel
(sep3_direct_kernel<15, W>: x/y filters on FP32 matrix cores with operands
loaded per wave, z filter on the vector ALUs; ``sep3.engine=0`` selects the
LDS-staged sep3_mfma_kernel, ``sep3.engine=1`` the all-vector sep3_fused_kernel):
achieved = algorithmic bytes per launch
(256^3 voxels x (4 B read + 4 B written)) / average launch time measured
with HIP events on the stream the kernel runs on. ``cpu_baseline`` times
the reference's own CPU convolve (compiled from its headers into
oracle/_ref) on a 48-plane z-slab of the same workload with the dense
15^3 outer-product kernel, which is what the reference runs for it.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
MFMA16_PEAK_TFLOPS = 2500.0
LEGS = ("separable_split", "dense15_reference", "dense15_fp32", "dense15_split_fp16", "dense15_split_bf16", "u16_binomial3",
        "dense3_fp32", "image_2d", "sequence", "rl_fft", "rl_split", "rl_error", "rl_batch", "learned")
METRIC = "Mvoxel/s 3-D conv (%HBM roofline); RL time-to-equal-likelihood, learned vs plain"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class Timer:
    """HIP events recorded on an explicit stream through the C-ABI."""

    def __init__(self, stream):
        from iane_amd import _lib
        import ctypes
        self._lib = _lib
        self.stream = stream
        self.start = ctypes.c_void_p()
        self.stop = ctypes.c_void_p()
        _lib.call("iane_hip_event_create", ctypes.byref(self.start))
        _lib.call("iane_hip_event_create", ctypes.byref(self.stop))

    def time(self, fn, steps: int) -> float:
        import ctypes
        _lib = self._lib
        _lib.call("iane_hip_stream_sync", self.stream)
        _lib.call("iane_hip_event_record", self.start, self.stream)
        for _ in range(steps):
            fn()
        _lib.call("iane_hip_event_record", self.stop, self.stream)
        ms = ctypes.c_float()
        _lib.call("iane_hip_event_elapsed_ms", ctypes.byref(ms), self.start, self.stop)
        return ms.value


def settle_clocks(timer, fn, ms: float = 20.0, max_calls: int = 2000):
    """Run `fn` untimed until it has kept the GPU busy for `ms` milliseconds (at
    most `max_calls` calls): clocks take tens of milliseconds of sustained work
    to settle, and a short leg timed straight after a cold start reads them low."""
    busy, calls = 0.0, 0
    while busy < ms and calls < max_calls:
        busy += timer.time(fn, 50)
        calls += 50


def traffic_from_profiles(kernel_tag: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if present."""
    path = ROOT / "profiles" / f"{kernel_tag}_traffic.json"
    if not path.exists():
        return None
    try:
        return float(json.loads(path.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(volume: np.ndarray, taps: np.ndarray, planes: int = 48):
    """Reference CPU convolve on a z-slab. Threads: min(16, os.cpu_count()). The
    reference's default is std::thread::hardware_concurrency(), which on the GPU
    host reports every core of the machine, while a one-GPU job there is given a
    share of 16 CPUs, so 16 threads is what a user of that one GPU gets."""
    from oracle import oracle
    cores = min(16, os.cpu_count() or 1)
    t = taps.astype(np.float64)
    dense = np.multiply.outer(np.multiply.outer(t, t), t).astype(np.float32)
    z0 = volume.shape[0] // 2 - planes // 2
    slab = np.ascontiguousarray(volume[z0:z0 + planes])
    kind = "reference" if oracle.have_reference() else "port"
    run = (lambda: oracle.convolve_ref(slab, dense, np.float32, oracle.ZERO, threads=cores)) if kind == "reference" \
        else (lambda: oracle.convolve(slab, dense, np.float32, oracle.ZERO, threads=cores))
    run()  # warm-up (run_on.hpp:43-59: warmup then samples)
    samples = []
    for _ in range(3):
        t0 = time.perf_counter()
        run()
        samples.append(time.perf_counter() - t0)
    median = sorted(samples)[1]
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": slab.size / median / 1e6, "unit": "Mvoxel/s", "cores": cores, "kind": kind,
            "cpu_model": model, "min_s": min(samples), "max_s": max(samples),
            "sample": f"{planes}-plane z-slab ({slab.shape[2]}x{slab.shape[1]}x{planes}) of the 256^3 workload, "
                      f"dense 15^3 outer-product kernel (the reference has no separable path), zero border, "
                      f"cpu_policy{{threads={cores}, grain=4096}}, median of 3 after 1 warm-up",
            "median_s": median,
            "cores_note": "min(16, os.cpu_count()): one GPU's share of the host's CPUs; the reference's "
                          "default std::thread::hardware_concurrency() would count the whole host"}


def _lib_sync(stream) -> None:
    """Wait for our stream before torch (on its own stream) reads results back."""
    from iane_amd import _lib
    _lib.call("iane_hip_stream_sync", stream)


def bench_rl(timer, stream, iterations: int = 200):
    """Plain FFT Richardson–Lucy on the 256x256x128 volume with the 31x31x63 PSF:
    GPU time of `iterations` iterations (inputs resident) and the final log-likelihood."""
    import torch
    from iane_amd import deconvolution, hip_policy, synth
    log("[bench] RL: generating 256x256x128 volume (seed 3)")
    _, psf, y = synth.rl_volume_data()
    dy, dpsf = torch.from_numpy(y).cuda(), torch.from_numpy(psf).cuda()
    from iane_amd import autotune
    tuned = autotune.tune_rl_plan(y.shape, dpsf, policy=hip_policy(stream))
    t0 = time.perf_counter()
    plan = deconvolution.rl_plan(y.shape, dpsf, deconvolution.FFT, policy=hip_policy(stream))
    est = torch.empty_like(dy)
    ll = torch.empty(iterations, dtype=torch.float64, device="cuda")  # every entry is written by the run
    plan.run(dy, est, 1).wait()
    prepare_s = time.perf_counter() - t0
    ms = timer.time(lambda: plan.run(dy, est, iterations, ll_history=ll), 1)
    _lib_sync(stream)
    hist = ll.cpu().numpy()
    voxels = y.size
    log(f"[bench] RL: {iterations} iterations in {ms:.2f} ms ({ms / iterations * 1e3:.1f} us/iteration)")
    return {"iterations": iterations, "ms_total": round(ms, 3), "us_per_iteration": round(ms / iterations * 1e3, 2),
            "Mvoxel_iter_s": round(voxels * iterations / ms / 1e3, 1), "ll_final": float(hist[-1]),
            "ll_first": float(hist[0]), "prepare_s": round(prepare_s, 3),
            "padded_domain": "280x280x160 (x, y, z) R2C", "workspace_bytes": plan.workspace_bytes,
            "launch_shape": tuned}


def bench_image_2d(timer, stream, repeats: int = 50):
    """The 2-D workload: 1024x1024 U[0,1) image with the 7x7 Gaussian (sigma 1.5, clamp border)
    in reference (FP64, bit-exact) and FP32 arithmetic, and its u8 quantisation with the
    integer 5x5 box into i32 (bit-exact). Two figures per case: `us`, calls issued one by one
    from Python (a 1024^2 call is shorter than the host's per-call cost, so this is the
    host-paced rate), and `graph_us`, the GPU time per call of 50 calls captured once in a
    HIP graph and replayed (every replay checked equal to the eager output)."""
    import torch
    import iane_amd
    from iane_amd import border_mode, convolution_options, hip_policy, precision, synth
    img, kernel, img_u8, box = synth.image_2d_uniform()
    policy = hip_policy(stream)
    d_img, d_k = torch.from_numpy(img).cuda(), torch.from_numpy(kernel).cuda()
    d_u8, d_box = torch.from_numpy(img_u8).cuda(), torch.from_numpy(box).cuda()
    out = torch.empty_like(d_img)
    out_i32 = torch.empty(img.shape, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    pixels = img.size
    capture = torch.cuda.Stream()
    capture_policy = hip_policy(capture.cuda_stream)
    result = {}
    for name, make, dst in (
            ("f32_gauss7_reference", lambda p: lambda: iane_amd.convolve(
                d_img, d_k, out, convolution_options(border_mode.clamp, precision.reference), p), out),
            ("f32_gauss7_fp32", lambda p: lambda: iane_amd.convolve(
                d_img, d_k, out, convolution_options(border_mode.clamp, precision.fp32), p), out),
            ("u8_box5_i32", lambda p: lambda: iane_amd.convolve(
                d_u8, d_box, out_i32, convolution_options(border_mode.clamp, precision.reference), p), out_i32)):
        fn = make(policy)
        fn()
        _lib_sync(stream)
        ms = timer.time(fn, repeats) / repeats
        eager = dst.clone()
        graph = torch.cuda.CUDAGraph()
        call = make(capture_policy)
        with torch.cuda.graph(graph, stream=capture):
            for _ in range(repeats):
                call()
        for _ in range(2):
            dst.zero_()
            graph.replay()
            torch.cuda.synchronize()
            if not torch.equal(dst, eager):
                raise RuntimeError(f"2-D {name}: graph replay differs from the eager call")
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(capture):
            start.record()
            graph.replay()
            graph.replay()
            stop.record()
        stop.synchronize()
        graph_ms = start.elapsed_time(stop) / (2 * repeats)
        result[name] = {"us": round(ms * 1e3, 2), "Mpixel_s": round(pixels / ms / 1e3, 1),
                        "graph_us": round(graph_ms * 1e3, 2), "graph_Mpixel_s": round(pixels / graph_ms / 1e3, 1)}
        log(f"[bench] 2-D {name}: {ms * 1e3:.1f} us per call, {graph_ms * 1e3:.2f} us in a graph")
    return result


def bench_rl_split(timer, stream, iterations: int = 2):
    """RL with the DIRECT method on split-precision MFMA convolution (hi + lo FP16 / BF16
    operands, FP32 accumulation) on the same 256x256x128 volume and 31x31x63 PSF: GPU time
    per iteration, and the largest difference from the FFT method's estimate after the same
    iterations (the FFT method is checked against the FP64 oracle in the tests)."""
    import torch
    from iane_amd import deconvolution, hip_policy, precision, synth
    _, psf, y = synth.rl_volume_data()
    dy, dpsf = torch.from_numpy(y).cuda(), torch.from_numpy(psf).cuda()
    policy = hip_policy(stream)
    fft_plan = deconvolution.rl_plan(y.shape, dpsf, deconvolution.FFT, policy=policy)
    ref = torch.empty_like(dy)
    fft_plan.run(dy, ref, iterations).wait()
    scale = float(ref.abs().max())
    out = {"iterations": iterations, "psf": "31x31x63", "compared_with": "FFT method, same iterations"}
    for name, prec in (("split_fp16", precision.split_fp16), ("split_bf16", precision.split_bf16)):
        plan = deconvolution.rl_plan(y.shape, dpsf, deconvolution.DIRECT, prec, policy=policy)
        est = torch.empty_like(dy)
        plan.run(dy, est, 1).wait()  # warm-up
        ms = timer.time(lambda: plan.run(dy, est, iterations), 1)
        _lib_sync(stream)
        diff = float((est - ref).abs().max()) / scale
        out[name] = {"ms_per_iteration": round(ms / iterations, 2), "max_rel_diff_vs_fft": float(f"{diff:.3g}")}
        log(f"[bench] RL direct {name}: {ms / iterations:.1f} ms/iteration, max diff vs FFT {diff:.2e}")
    return out


def bench_separable_split(timer, vol, taps, out, policy, stream, steps: int, headline_out: np.ndarray):
    """The headline workload on the split-FP16 separable engine (precision.split_fp16:
    every filter on v_mfma_f32_16x16x32_f16 with hi + lo operands, FP32 sums), and the
    error of both engines over the whole volume against float64 per-axis passes
    (scipy.ndimage.convolve1d, zero border: a separable kernel's zero border factors
    over the axes)."""
    import iane_amd
    from iane_amd import border_mode, convolution_options, precision
    from scipy import ndimage
    opts = convolution_options(border_mode.zero, precision.split_fp16)

    def step():
        iane_amd.convolve_separable(vol, [taps, taps, taps], out, opts, policy)

    settle = 0.0
    while settle < 30.0:
        settle += timer.time(step, 50)
    ms = timer.time(step, steps) / steps
    step()
    _lib_sync(stream)
    split_out = out.cpu().numpy()
    t = taps.cpu().numpy().astype(np.float64)
    ref = vol.cpu().numpy().astype(np.float64)
    for axis in (2, 1, 0):
        ref = ndimage.convolve1d(ref, t, axis=axis, mode="constant", cval=0.0)
    voxels = ref.size
    result = {"us": round(ms * 1e3, 2), "Mvoxel_s": round(voxels / ms / 1e3, 1),
              "GB_s": round(voxels * 8 / ms / 1e6, 1), "frac_hbm": round(voxels * 8 / ms / 1e6 / HBM_PEAK_GBS, 4),
              "kernel": "sep3_split_kernel<15>",
              "error_vs_fp64": _errors(split_out, ref), "fp32_engine_error_vs_fp64": _errors(headline_out, ref)}
    log(f"[bench] separable split_fp16: {ms * 1e3:.1f} us, error {result['error_vs_fp64']}, "
        f"fp32 engine error {result['fp32_engine_error_vs_fp64']}")
    return result


def _errors(est: np.ndarray, ref: np.ndarray) -> dict:
    diff = est.astype(np.float64) - ref
    return {"l2_rel": float(f"{np.linalg.norm(diff) / np.linalg.norm(ref):.3g}"),
            "max_abs_rel": float(f"{np.abs(diff).max() / np.abs(ref).max():.3g}")}


def _conv_zero_fp64(x: np.ndarray, psf: np.ndarray) -> np.ndarray:
    """Zero-border convolution with an odd centred kernel in float64 (a padded FFT
    product; scipy's pocketfft)."""
    from scipy import fft as sfft
    shape = [n + k - 1 for n, k in zip(x.shape, psf.shape)]
    axes = list(range(x.ndim))
    workers = min(16, os.cpu_count() or 1)
    full = sfft.irfftn(sfft.rfftn(x, shape, axes, workers=workers) * sfft.rfftn(psf, shape, axes, workers=workers),
                       shape, axes, workers=workers)
    return full[tuple(slice(k // 2, k // 2 + n) for n, k in zip(x.shape, psf.shape))]


def _rl_fp64(observed: np.ndarray, psf: np.ndarray, iterations: int, epsilon: float = 1e-6) -> np.ndarray:
    """Float64 RL (examples/rl_deconvolution/main.cu:51-81), the reference the error
    figures below are taken against; the tests pin the same formulas (the oracle's rl)
    to the reference's own golden runs."""
    y = observed.astype(np.float64)
    h = psf.astype(np.float64)
    flipped = h[::-1, ::-1, ::-1].copy()
    sensitivity = _conv_zero_fp64(np.ones_like(y), flipped)
    x = np.where(y > epsilon, y, epsilon)
    for _ in range(iterations):
        p = _conv_zero_fp64(x, h)
        c = _conv_zero_fp64(y / np.where(p > epsilon, p, epsilon), flipped)
        x = x * np.where(sensitivity > 0, c / np.where(sensitivity > epsilon, sensitivity, epsilon), 0.0)
    return x


def bench_rl_fp64_error(stream):
    """Error of the GPU RL estimates against float64 RL (scipy FFT convolutions
    restating examples/rl_deconvolution/main.cu:51-81): normwise ||x - x64|| / ||x64||
    and max|x - x64| / max|x64|, for the FFT method and the split-precision direct
    method. Full 256x256x128 volume with the 31x31x63 PSF after 2 iterations, and a
    64^3 volume of the same construction (40 blobs) after 200 iterations, where the
    FP32 rounding has accumulated."""
    import torch
    from iane_amd import deconvolution, hip_policy, precision, synth
    policy = hip_policy(stream)
    out = {"reference": "float64 RL on scipy FFT convolutions, zero border"}
    for label, shape, blobs, iterations in (("256x256x128_2_iterations", (128, 256, 256), 300, 2),
                                            ("64x64x64_200_iterations", (64, 64, 64), 40, 200)):
        _, psf, y = synth.rl_volume_data(shape=shape, blobs=blobs)
        t0 = time.perf_counter()
        ref = _rl_fp64(y, psf, iterations)
        log(f"[bench] RL error {label}: float64 reference in {time.perf_counter() - t0:.1f} s")
        dy, dpsf = torch.from_numpy(y).cuda(), torch.from_numpy(psf).cuda()
        leg = {}
        for name, method, prec in (("fft", deconvolution.FFT, precision.fp32),
                                   ("direct_split_fp16", deconvolution.DIRECT, precision.split_fp16),
                                   ("direct_split_bf16", deconvolution.DIRECT, precision.split_bf16)):
            plan = deconvolution.rl_plan(y.shape, dpsf, method, prec, policy=policy)
            est = torch.empty_like(dy)
            plan.run(dy, est, iterations).wait()
            leg[name] = _errors(est.cpu().numpy(), ref)
            log(f"[bench] RL error {label} {name}: {leg[name]}")
        out[label] = leg
    return out


def bench_rl_batch(timer, stream, iterations: int = 100):
    """64 RL problems (edges U{32..192}, own PSFs, seed 4), 100 iterations:
    one hipGraph launch of the batched engine vs the same volumes run one plan at a time."""
    import torch
    from iane_amd import deconvolution, hip_policy, synth
    log("[bench] RL batch: generating 64 volumes (seed 4)")
    data = synth.rl_batch_data(count=64, seed=4)
    policy = hip_policy(stream)
    obs = [torch.from_numpy(o).cuda() for o, _ in data]
    psfs = [torch.from_numpy(p).cuda() for _, p in data]
    est = [torch.empty_like(o) for o in obs]
    from iane_amd import autotune
    tuned = autotune.tune_rl_batch(obs, psfs, policy=policy)
    log(f"[bench] RL batch: autotuned {tuned}")
    t0 = time.perf_counter()
    batch = deconvolution.rl_batch([o.shape for o in obs], psfs, policy=policy)
    prepare_s = time.perf_counter() - t0
    batch.run(obs, est, iterations, use_graph=True).wait()  # capture + first replay
    graph_ms = timer.time(lambda: batch.run(obs, est, iterations, use_graph=True), 1)
    eager_ms = timer.time(lambda: batch.run(obs, est, iterations, use_graph=False), 1)
    plans = [deconvolution.rl_plan(o.shape, p, deconvolution.FFT, policy=policy) for o, p in zip(obs, psfs)]
    plans[0].run(obs[0], est[0], 1).wait()
    single_ms = timer.time(lambda: [pl.run(o, e, iterations) for pl, o, e in zip(plans, obs, est)], 1)
    voxels = sum(o.numel() for o in obs)
    log(f"[bench] RL batch: graph {graph_ms:.1f} ms, batched eager {eager_ms:.1f} ms, one plan at a time {single_ms:.1f} ms")
    return {"volumes": len(obs), "voxels": voxels, "iterations": iterations, "graph_ms": round(graph_ms, 2),
            "batched_eager_ms": round(eager_ms, 2), "per_volume_plans_ms": round(single_ms, 2),
            "speedup_vs_per_volume": round(single_ms / graph_ms, 2),
            "Mvoxel_iter_s": round(voxels * iterations / graph_ms / 1e3, 1), "prepare_s": round(prepare_s, 2),
            "launch_shape": tuned}


def bench_sequence(timer, stream, vol):
    """Scan / reduce on the 256^3 f32 volume: a whole f64-accumulated sum (the
    log-likelihood reduction pattern; 4 B read per voxel), an f32 inclusive scan
    in storage order (4 B read + 4 B written) and per-row sums along x. `us` is
    the host-paced time per call (one Python call each); `graph_us` the GPU time
    per call of 20 calls captured in a HIP graph and replayed (each replay checked
    equal to the eager result), which `frac_hbm` uses."""
    import torch
    import iane_amd
    policy = iane_amd.hip_policy(stream)
    ctx = iane_amd.make_execution_context(policy)
    capture = torch.cuda.Stream()
    capture_ctx = iane_amd.make_execution_context(iane_amd.hip_policy(capture.cuda_stream))
    voxels = vol.numel()
    total = torch.empty(1, dtype=torch.float64, device="cuda")
    scan = torch.empty_like(vol)
    rows = torch.empty(vol.shape[:-1], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    legs = (("reduce_sum_f32_to_f64", lambda: iane_amd.reduce("sum", 0.0), total, 4.0),
            ("inclusive_scan_sum_f32", lambda: iane_amd.inclusive_scan("sum", 0.0), scan, 8.0),
            ("reduce_along_x_sum_f32", lambda: iane_amd.reduce_along(0, "sum", 0.0), rows, 4.0 + 4.0 / vol.shape[-1]))
    out = {}
    k = 20
    for name, make, dst, bytes_per_voxel in legs:
        op = ctx.prepare(make(), vol, dst)
        fn = lambda: ctx.run(op, vol, dst)  # noqa: E731
        fn()
        _lib_sync(stream)
        t = timer.time(fn, k) / k
        eager = dst.clone()
        gop = capture_ctx.prepare(make(), vol, dst)
        with torch.cuda.stream(capture):
            capture_ctx.run(gop, vol, dst)
        capture.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=capture):
            for _ in range(k):
                capture_ctx.run(gop, vol, dst)
        for _ in range(2):
            dst.zero_()
            graph.replay()
            torch.cuda.synchronize()
            if not torch.equal(dst, eager):
                raise RuntimeError(f"{name}: graph replay differs from the eager call")
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(capture):
            start.record()
            graph.replay()
            graph.replay()
            stop.record()
        stop.synchronize()
        tg = start.elapsed_time(stop) / (2 * k)
        gbs = voxels * bytes_per_voxel / tg / 1e6
        out[name] = {"us": round(t * 1e3, 2), "graph_us": round(tg * 1e3, 2), "GB_s": round(gbs, 1),
                     "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)}
        log(f"[bench] {name}: {t * 1e3:.1f} us per call, {tg * 1e3:.1f} us in a graph, {gbs:.0f} GB/s")
    return out


def bench_learned(stream, volumes: int = 4):
    """Learned warm start + batched RL vs plain batched RL on 256x256x128 volumes
    (blobs, PSF sigma (3, 1.5, 1.5) on 63x31x31, peak 1000, Poisson; made on the
    GPU from seed 3): GPU time until every volume's Poisson log-likelihood reaches
    that of plain 200-iteration RL."""
    import torch
    from iane_amd import hip_policy, learned
    if not learned.WEIGHTS.exists():
        return {"skipped": f"no trained weights at {learned.WEIGHTS}"}
    policy = hip_policy(stream)
    net = learned.warm_start_net(policy=policy).load()
    psf = torch.from_numpy(learned.psf_kernel()).cuda()
    log(f"[bench] learned: {volumes} observed volumes 256x256x128")
    ys = list(learned.observed_volumes(volumes, (128, 256, 256), seed=3, policy=policy))
    from iane_amd import autotune
    tuned = autotune.tune_rl_batch(ys, [psf] * volumes, policy=policy)
    log(f"[bench] learned: autotuned batch {tuned}")
    out = learned.batch_time_to_likelihood(net, ys, [psf] * volumes, iterations=200, max_iterations=120,
                                           policy=policy)
    log(f"[bench] learned: plain {out['plain_ms']:.1f} ms, learned {out['learned_ms']:.1f} ms "
        f"(inference {out['inference_ms']:.1f} ms, k={out['k_max']}), speedup {out['speedup']:.2f}x")
    out["plain_ms"] = round(out["plain_ms"], 2)
    out["learned_ms"] = round(out["learned_ms"], 2)
    out["inference_ms"] = round(out["inference_ms"], 2)
    out["speedup"] = round(out["speedup"], 2)
    out["network_parameters"] = net.parameter_count()
    out["launch_shape"] = tuned
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--direct-only", action="store_true",
                    help="headline plus the two streaming direct-convolution legs (u16 3^3, f32 dense 3^3)")
    ap.add_argument("--legs", default=None, metavar="NAME[,NAME...]",
                    help="run only these secondary legs (after the headline): " + ", ".join(LEGS))
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes (no autotuner launches)")
    ap.add_argument("--launch-shape", default=None, metavar="WAVES,CHUNK",
                    help="fix the separable kernel's launch shape (sep3.waves, sep3.chunk) instead of searching: "
                         "profiles of one autotuned shape without the candidates' launches")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy the 256^3 volume once (device to device, 64 MiB read + 64 MiB written): "
                         "the reference dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.quick:
        legs = set()
    elif args.legs is not None:
        legs = set(args.legs.split(",")) - {""}
        unknown = legs - set(LEGS)
        if unknown:
            ap.error(f"unknown legs {sorted(unknown)}; choose from {', '.join(LEGS)}")
    elif args.direct_only:
        legs = {"u16_binomial3", "dense3_fp32"}
    else:
        legs = set(LEGS)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(slab decomposition with halo exchange is designed, not built)", file=sys.stderr)
        return 2

    import ctypes
    import torch
    import iane_amd
    from iane_amd import _lib, synth, border_mode, convolution_options, precision

    torch.cuda.init()
    n = args.size
    log(f"[bench] generating 256^3 workload (seed 2)")
    data = synth.volume_blobs(n, seed=2)
    vol_h, taps_h = data["volume"], data["taps"]
    stream_p = ctypes.c_void_p()
    _lib.call("iane_hip_stream_create", ctypes.byref(stream_p))
    stream = stream_p.value
    policy = iane_amd.hip_policy(stream)
    timer = Timer(stream)

    vol = torch.from_numpy(vol_h).cuda()
    taps = torch.from_numpy(taps_h).cuda()
    out = torch.empty_like(vol)
    torch.cuda.synchronize()
    opts = convolution_options(border_mode.zero, precision.fp32)

    def step():
        iane_amd.convolve_separable(vol, [taps, taps, taps], out, opts, policy)

    from iane_amd import autotune
    if args.launch_shape:
        waves, chunk = (int(v) for v in args.launch_shape.split(","))
        tuned = {"sep3.waves": waves, "sep3.chunk": chunk}
        for key, value in tuned.items():
            autotune.set(key, value)
    else:
        tuned = {} if args.no_tune else autotune.tune_separable(vol, [taps, taps, taps], out, opts, policy)
    log(f"[bench] autotuned launch shape: {tuned}")

    # Steady clocks: the GPU needs tens of milliseconds of sustained work before
    # its clocks settle (measured on this kernel: 74 us per launch after 5 launches,
    # 68 us after 300). The autotuner's candidate launches usually provide that; a
    # cached or fixed launch shape does not, so run the step untimed for at least
    # 30 ms of GPU time first. The timed region below is still exactly `steps` steps.
    settle_ms, settle_steps = 0.0, 0
    while settle_ms < 30.0 and settle_steps < 2000:
        settle_ms += timer.time(step, 50)
        settle_steps += 50
    log(f"[bench] separable 15^3 fp32: {settle_steps} settling steps ({settle_ms:.1f} ms), warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.call("iane_hip_stream_sync", stream)
    log(f"[bench] separable 15^3 fp32: timing {args.steps} steps")
    t0 = time.perf_counter()
    ms = timer.time(step, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ms / args.steps
    voxels = vol_h.size
    value = voxels / (ms_per_step / 1e3) / 1e6
    bytes_per_launch = voxels * 8.0
    achieved = bytes_per_launch / (ms_per_step / 1e3) / 1e9
    engine = autotune.get("sep3.engine")
    if engine == 1:
        kernel, tag = "sep3_fused_kernel<15>", "sep3_fused_256"
    elif engine == 0:
        kernel, tag = f"sep3_mfma_kernel<15, {3 if autotune.get('sep3.waves') == 3 else 2}>", "sep3_mfma_256"
    else:
        kernel, tag = f"sep3_direct_kernel<15, {2 if autotune.get('sep3.waves') == 2 else 1}>", "sep3_direct_256"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(tag),
                "kernel": kernel, "launch_shape": tuned, "bytes_per_launch": bytes_per_launch,
                "avg_launch_us": round(ms_per_step * 1e3, 2)}
    log(f"[bench] separable: {ms_per_step * 1e3:.1f} us/step, {value:.0f} Mvoxel/s, {achieved:.0f} GB/s")

    extra = {}
    if "separable_split" in legs:
        _lib.call("iane_hip_stream_sync", stream)
        headline_out = out.cpu().numpy()
        extra["separable_split_fp16"] = bench_separable_split(timer, vol, taps, out, policy, stream, args.steps, headline_out)
    if args.copy_calibration:
        _lib.call("iane_hip_memcpy_async", out.data_ptr(), vol.data_ptr(), voxels * 4, 3, stream)
        _lib.call("iane_hip_stream_sync", stream)
        extra["copy_calibration_bytes"] = {"read": voxels * 4, "written": voxels * 4}
    dense_legs = [(name, prec) for name, prec in (
            ("dense15_reference", precision.reference), ("dense15_fp32", precision.fp32),
            ("dense15_split_fp16", precision.split_fp16), ("dense15_split_bf16", precision.split_bf16))
            if name in legs]
    if dense_legs:
        # dense random 15^3 kernel in each arithmetic: compute bound
        dense = torch.from_numpy(data["dense"]).cuda()
        for name, prec in dense_legs:
            o = convolution_options(border_mode.zero, prec)
            fn = lambda: iane_amd.convolve(vol, dense, out, o, policy)  # noqa: E731
            fn()
            _lib.call("iane_hip_stream_sync", stream)
            k = 3
            t = timer.time(fn, k) / k
            flops = 2.0 * voxels * 15 ** 3
            tflops = flops / t / 1e9
            entry = {"ms": round(t, 3), "Mvoxel_s": round(voxels / t / 1e3, 1), "TFLOP_s": round(tflops, 1)}
            if prec in (precision.split_fp16, precision.split_bf16):
                # algorithmic 2*N*K flops against the dense FP16/BF16 MFMA peak
                entry["frac_of_mfma_peak"] = round(tflops / MFMA16_PEAK_TFLOPS, 4)
            elif prec == precision.fp32:
                entry["frac_of_fp32_vector_peak"] = round(tflops / FP32_PEAK_TFLOPS, 3)
            else:
                entry["frac_of_fp64_vector_peak"] = round(tflops / FP64_PEAK_TFLOPS, 3)
            extra[name] = entry
            log(f"[bench] {name}: {t:.2f} ms")
    if "u16_binomial3" in legs:
        # u16 volume, 3^3 binomial int kernel, int32 result (bit-exact)
        v16 = torch.from_numpy(data["volume_u16"].view(np.int16)).cuda().view(torch.uint16)
        binom = torch.from_numpy(data["binomial"]).cuda()
        o32 = torch.empty(vol.shape, dtype=torch.int32, device="cuda")
        o = convolution_options(border_mode.zero, precision.reference)
        fn = lambda: iane_amd.convolve(v16, binom, o32, o, policy)  # noqa: E731
        fn()
        _lib.call("iane_hip_stream_sync", stream)
        settle_clocks(timer, fn)
        k = 20
        t = timer.time(fn, k) / k
        extra["u16_binomial3"] = {"ms": round(t, 4), "Mvoxel_s": round(voxels / t / 1e3, 1),
                                  "GB_s": round(voxels * 6 / t / 1e6, 1),
                                  "frac_hbm": round(voxels * 6 / t / 1e6 / HBM_PEAK_GBS, 3),
                                  "kernel": "conv_column_kernel<unsigned short, int, true, 4, 2>",
                                  "traffic": traffic_from_profiles("conv_column_u16_256")}
        log(f"[bench] u16 3^3: {t * 1e3:.1f} us")
    if "dense3_fp32" in legs:
        # f32 volume, dense random 3^3 kernel, FP32 arithmetic (register column kernel)
        k = 20
        k3 = torch.from_numpy(data["dense3"]).cuda()
        o = convolution_options(border_mode.zero, precision.fp32)
        fn = lambda: iane_amd.convolve(vol, k3, out, o, policy)  # noqa: E731
        fn()
        _lib.call("iane_hip_stream_sync", stream)
        settle_clocks(timer, fn)
        t = timer.time(fn, k) / k
        extra["dense3_fp32"] = {"ms": round(t, 4), "Mvoxel_s": round(voxels / t / 1e3, 1),
                                "GB_s": round(voxels * 8 / t / 1e6, 1),
                                "frac_hbm": round(voxels * 8 / t / 1e6 / HBM_PEAK_GBS, 3),
                                "kernel": "conv_column_kernel<float, float, false, 4, 2>",
                                "traffic": traffic_from_profiles("conv_column_dense3_256")}
        log(f"[bench] f32 dense 3^3: {t * 1e3:.1f} us")

    if "image_2d" in legs:
        extra["image_2d_1024"] = bench_image_2d(timer, stream)
    if "sequence" in legs:
        extra["sequence_256"] = bench_sequence(timer, stream, vol)
    if "rl_fft" in legs:
        extra["rl_fft_256x256x128_psf31x31x63"] = bench_rl(timer, stream)
    if "rl_split" in legs:
        extra["rl_direct_split_256x256x128_psf31x31x63"] = bench_rl_split(timer, stream)
    if "rl_error" in legs:
        extra["rl_error_vs_fp64"] = bench_rl_fp64_error(stream)
    if "rl_batch" in legs:
        extra["rl_batch_64_volumes"] = bench_rl_batch(timer, stream)
    if "learned" in legs:
        extra["rl_learned_vs_plain_256x256x128"] = bench_learned(stream)

    base = None
    if not args.no_cpu:
        log("[bench] cpu baseline (reference CPU convolve, z-slab)")
        base = cpu_baseline(vol_h, taps_h)
        log(f"[bench] cpu baseline: {base['value']:.2f} Mvoxel/s on {base['cores']} cores")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mvoxel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (Philox4x32-10 seed 2: U[0,1) + 200 Gaussian blobs)",
        "config": {"workload": f"3-D FP32 convolution {n}^3, separable 15x15x15 Gaussian sigma 2.5, zero border",
                   "voxels": voxels, "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": base, "wall_s_timed": round(wall, 3), "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
