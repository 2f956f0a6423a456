"""GPU parity of the split-FP16 separable engine (precision.split_fp16).

Every filter runs on the FP16 matrix instructions with operands split into
hi + lo parts (22 significant bits) after exact power-of-two scales: one per
plane from the plane's largest sample, one per axis from the taps, and a
per-window compensation of the z taps. The bound is the FP32 one against the
FP64 oracle (a separable kernel with the zero border factors over axes, so the
oracle is the dense outer-product kernel or three 1-D passes):
    max|gpu - oracle64| <= TOL * max|oracle64|,   TOL = 1e-5.
The scale handling is exercised with planes of very different magnitude,
very small and very large images, tiny and huge taps, and all-zero inputs;
every case runs with both z filters (register ring, default; split MFMA
windows over an LDS ring).
"""
from __future__ import annotations

import numpy as np
import pytest

import iane_amd
from iane_amd import autotune, border_mode, convolution_options, precision
from oracle import oracle
from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(autouse=True, params=[1, 0], ids=["ring_z", "mfma_z"])
def z_mode(request):
    """Every case runs with the z filter on the register ring (default) and on
    split MFMA windows over the LDS ring."""
    with autotune.applied({"sep3.split_z": request.param}):
        yield request.param


def dev(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def rel_err(got, ref64):
    return float(np.max(np.abs(got.astype(np.float64) - ref64)) / max(np.max(np.abs(ref64)), 1e-300))


def separable_ref(src, taps, border=cases.ZERO):
    """Three FP64 1-D passes of the C oracle (x, then y, then z)."""
    ref = src.astype(np.float64)
    for axis, t in enumerate(taps):
        shape = [1, 1, 1]
        shape[2 - axis] = t.size
        ref = oracle.convolve(np.ascontiguousarray(ref), t.astype(np.float64).reshape(shape), np.float64, border,
                              threads=8)
    return ref


def run_split(src, taps, border=cases.ZERO, out=None):
    if out is None:
        out = dev(np.full(src.shape, 77, np.float32))
    iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                convolution_options(border_mode(border), precision.split_fp16)).wait()
    return out


@pytest.mark.parametrize("k", [1, 3, 5, 7, 9, 11, 13, 15])
def test_split_every_tap_length(k):
    rng = np.random.default_rng(100 + k)
    src = rng.random((37, 45, 72)).astype(np.float32)
    taps = [rng.random(k).astype(np.float32) for _ in range(3)]
    got = host(run_split(src, taps))
    err = rel_err(got, separable_ref(src, taps))
    print(f"k={k}: {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("shape", [(5, 16, 64), (40, 17, 68), (3, 100, 132), (70, 33, 260), (97, 33, 256),
                                   (130, 7, 4), (16, 16, 16), (33, 1, 8)])
def test_split_ragged_tiles_signed_taps(shape):
    """Partial 16 x 16 wave tiles, z extents that are not whole 16-plane windows,
    signed taps of unequal lengths' worth of cancellation."""
    rng = np.random.default_rng(sum(shape))
    src = (rng.random(shape) - 0.3).astype(np.float32)
    taps = [rng.standard_normal(7).astype(np.float32) for _ in range(3)]
    got = host(run_split(src, taps))
    assert np.isfinite(got).all()
    assert rel_err(got, separable_ref(src, taps)) <= TOL


def test_split_planes_of_very_different_magnitude():
    """Per-plane scales from 2^-40 to 2^40 meet in every z window: each plane's
    operands are scaled on their own and the z taps compensate. Checked over
    the whole volume and, per output plane, against that plane's own magnitude
    where the z neighbourhood is within 2^10 of it."""
    rng = np.random.default_rng(7)
    nz = 64
    expo = np.concatenate([np.zeros(16), rng.integers(-40, 41, nz - 32), np.full(16, 3.0)])
    src = ((rng.random((nz, 40, 64)) - 0.5) * np.exp2(expo)[:, None, None]).astype(np.float32)
    taps = [rng.random(15).astype(np.float32) + 0.1 for _ in range(3)]
    got = host(run_split(src, taps)).astype(np.float64)
    ref = separable_ref(src, taps)
    assert rel_err(got, ref) <= TOL
    # planes 0..15 hold unit-scale data, their outputs draw on planes up to 22
    near = slice(0, 8)
    assert rel_err(got[near], ref[near]) <= TOL


@pytest.mark.parametrize("scale", [2.0 ** -100, 1e-30, 1e30, 2.0 ** 100])
def test_split_tiny_and_huge_images(scale):
    rng = np.random.default_rng(3)
    src = (rng.random((30, 24, 48)) * scale).astype(np.float32)
    taps = [rng.random(9).astype(np.float32) for _ in range(3)]
    got = host(run_split(src, taps))
    assert np.isfinite(got).all()
    assert rel_err(got, separable_ref(src, taps)) <= TOL


@pytest.mark.parametrize("tap_scale", [1e-12, 1e12])
def test_split_tiny_and_huge_taps(tap_scale):
    rng = np.random.default_rng(4)
    src = rng.random((20, 20, 40)).astype(np.float32)
    taps = [(rng.random(11) * tap_scale).astype(np.float32) for _ in range(3)]
    got = host(run_split(src, taps))
    assert np.isfinite(got).all()
    assert rel_err(got, separable_ref(src, taps)) <= TOL


def test_split_zero_image_and_zero_taps():
    rng = np.random.default_rng(5)
    zero = np.zeros((20, 18, 36), np.float32)
    taps = [rng.random(5).astype(np.float32) for _ in range(3)]
    np.testing.assert_array_equal(host(run_split(zero, taps)), zero)
    src = rng.random((20, 18, 36)).astype(np.float32)
    ztaps = [taps[0], np.zeros(5, np.float32), taps[2]]
    np.testing.assert_array_equal(host(run_split(src, ztaps)), zero)


@pytest.mark.parametrize("chunk", [16, 32, 48, 64, 96])
def test_split_chunk_lengths(chunk):
    """Tuned z chunks (whole 16-plane windows) change where windows start, not the bound."""
    rng = np.random.default_rng(6)
    src = rng.random((100, 20, 32)).astype(np.float32)
    taps = [rng.random(15).astype(np.float32) for _ in range(3)]
    with autotune.applied({"sep3.chunk": chunk}):
        got = host(run_split(src, taps))
    assert rel_err(got, separable_ref(src, taps)) <= TOL


def test_split_falls_back_where_the_engine_does_not_apply():
    """Clamp border and a strided destination run the FP32 engines (more accurate)."""
    rng = np.random.default_rng(8)
    src = rng.random((12, 20, 30)).astype(np.float32)  # width not a multiple of 4
    taps = [rng.random(5).astype(np.float32) for _ in range(3)]
    for border in (cases.CLAMP, cases.ZERO):
        got = host(run_split(src, taps, border))
        assert rel_err(got, separable_ref(src, taps, border)) <= TOL
    src4 = rng.random((12, 20, 32)).astype(np.float32)
    import torch
    big = torch.zeros((12, 20, 64), dtype=torch.float32, device="cuda")
    view = big[:, :, ::2]
    run_split(src4, taps, cases.ZERO, out=view)
    assert rel_err(view.cpu().numpy(), separable_ref(src4, taps)) <= TOL


def test_split_full_size_headline_vs_fp64_oracle():
    """The headline workload (separable 15^3 Gaussian on the 256^3 blob volume,
    zero border) over the whole volume against three FP64 per-axis passes."""
    from iane_amd import synth
    vol = synth.volume_blobs(256, seed=2)
    v = vol["volume"]
    taps = vol["taps"]
    got = host(run_split(v, [taps] * 3)).astype(np.float64)
    ref = v.astype(np.float64)
    t = taps.astype(np.float64)
    for shape in ((1, 1, 15), (1, 15, 1), (15, 1, 1)):
        ref = oracle.convolve(ref, t.reshape(shape), np.float64, oracle.ZERO, threads=16)
    err = rel_err(got, ref)
    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"split separable 15^3 at 256^3 vs FP64 oracle: max {err:.2e}, normwise {l2:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("waves,chunk", [(1, 0), (2, 0), (2, 32), (1, 48)])
@pytest.mark.parametrize("k", [3, 9, 15])
def test_split_register_ring_z(k, waves, chunk):
    """The same engine with the z filter on a register ring of y-filtered planes
    (FP32 vector FMAs, no LDS; one or two waves per SIMD), ragged tiles and planes
    of different magnitude."""
    rng = np.random.default_rng(200 + k)
    expo = rng.integers(-20, 21, 41)
    src = ((rng.random((41, 37, 68)) - 0.4) * np.exp2(expo)[:, None, None]).astype(np.float32)
    taps = [rng.standard_normal(k).astype(np.float32) for _ in range(3)]
    with autotune.applied({"sep3.split_z": 1, "sep3.waves": waves, "sep3.chunk": chunk}):
        got = host(run_split(src, taps))
    assert np.isfinite(got).all()
    assert rel_err(got, separable_ref(src, taps)) <= TOL


@pytest.mark.parametrize("lengths", [(3, 7, 15), (15, 5, 1), (9, 1, 13)])
def test_split_unequal_tap_lengths(lengths):
    rng = np.random.default_rng(sum(lengths) + 1)
    src = (rng.random((40, 33, 64)) - 0.3).astype(np.float32)
    taps = [rng.standard_normal(k).astype(np.float32) for k in lengths]
    got = host(run_split(src, taps))
    assert rel_err(got, separable_ref(src, taps)) <= TOL


def test_split_rank2_runs_the_fp32_engines():
    rng = np.random.default_rng(9)
    src = rng.random((50, 64)).astype(np.float32)
    taps = [rng.random(7).astype(np.float32), rng.random(5).astype(np.float32)]
    out = dev(np.full(src.shape, 77, np.float32))
    iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                convolution_options(border_mode.zero, precision.split_fp16)).wait()
    dense = np.multiply.outer(taps[1].astype(np.float64), taps[0].astype(np.float64))
    assert rel_err(host(out), oracle.convolve(src, dense, np.float64, oracle.ZERO)) <= TOL
