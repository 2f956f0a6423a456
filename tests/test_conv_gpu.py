"""GPU parity of iane_hip_convolve / iane_hip_convolve_separable against the oracle.

REFERENCE precision must be bit-identical to the reference's CPU convolve
(oracle, golden fixtures written by the reference). FP32 precision and the
separable FP32 path must agree with the FP64 oracle to
    max|gpu - oracle64| <= TOL_FP32 * max|oracle64|      (TOL_FP32 = 1e-5).
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

import iane_amd
from iane_amd import border_mode, convolution_options, precision
from oracle import oracle
from tests import cases

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
TOL_FP32 = 1e-5
DT = [np.uint8, np.uint16, np.int32, np.float32, np.float64]


def dev(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def empty(shape, dtype):
    """Device image pre-filled with a sentinel so unwritten outputs show."""
    return dev(np.full(tuple(shape), 77, dtype=dtype))


def gpu_convolve(src, ker, ddt, border, prec=precision.reference, untiled=False):
    d = empty(src.shape, ddt)
    fn = iane_amd.convolve_untiled if untiled else iane_amd.convolve
    fn(dev(src), dev(ker), d, convolution_options(border_mode(border), prec)).wait()
    return host(d)


def same_bits(a, b):
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


# ------------------------------------------------------ known answers -----
@pytest.mark.parametrize("rank", [2, 3, 4])
def test_pattern_known_answer(rank):
    src, src2, kernel, first, second = cases.pattern_case(rank)
    same_bits(gpu_convolve(src, kernel, np.int32, cases.ZERO), first)
    same_bits(gpu_convolve(src2, kernel, np.int32, cases.ZERO), second)


def test_scalar_coordinates_known_answer():
    src, kernel, expected = cases.scalar_case()
    same_bits(gpu_convolve(src, kernel, np.int32, cases.ZERO), expected)


@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_crop_flip_known_answer(border):
    source, _, kernel, expected = cases.crop_cases()
    d_source = dev(source)
    crop = d_source[:, 1:3]  # device subview: pointer offset, parent strides
    out = empty((2, 2), np.uint8)
    iane_amd.convolve(crop, dev(kernel), out, convolution_options(border_mode(border))).wait()
    same_bits(host(out), expected[border])


@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_texture_case_exact(border):
    source, crop, kernel = cases.texture_case()
    expected = oracle.convolve(np.ascontiguousarray(crop), kernel, np.float32, border)
    d = dev(source)[1:4, 1:4]
    out = empty((3, 3), np.float32)
    iane_amd.convolve(d, dev(kernel), out, convolution_options(border_mode(border))).wait()
    same_bits(host(out), expected)


# ---------------------------------------------------- golden fixtures -----
@pytest.mark.parametrize("path", sorted(GOLDEN.glob("conv_*.npz")), ids=lambda p: p.stem)
def test_golden_bit_exact(path):
    data = np.load(path, allow_pickle=False)
    got = gpu_convolve(data["src"], data["kernel"], data["expected"].dtype, int(data["border"]))
    same_bits(got, data["expected"])


# ---------------------------------------------------- random matrix -------
def _case(rng, shape, kshape, sdt, kdt):
    if np.issubdtype(sdt, np.integer):
        hi = {np.uint8: 256, np.uint16: 65536, np.int32: 3000}[sdt]
        lo = -3000 if sdt == np.int32 else 0
        src = rng.integers(lo, hi, shape).astype(sdt)
    else:
        src = (rng.random(shape) * 100 - 10).astype(sdt)
    if np.issubdtype(kdt, np.integer):
        ker = rng.integers(-4, 5, kshape).astype(kdt)
    else:
        ker = (rng.random(kshape) - 0.3).astype(kdt)
    return src, ker


@pytest.mark.parametrize("sdt", DT)
@pytest.mark.parametrize("ddt", DT)
def test_dtype_matrix_bit_exact(sdt, ddt):
    rng = np.random.default_rng(abs(hash((str(sdt), str(ddt)))) % 2**32)
    for kdt in (np.int32, np.float32, np.float64):
        for shape, kshape in [((70,), (5,)), ((33, 70), (3, 7)), ((5, 40, 67), (3, 5, 9)),
                              ((3, 4, 20, 66), (1, 3, 3, 5))]:
            for border in (cases.CLAMP, cases.ZERO):
                src, ker = _case(rng, shape, kshape, sdt, kdt)
                expected = oracle.convolve(src, ker, ddt, border)
                same_bits(gpu_convolve(src, ker, ddt, border), expected)


@pytest.mark.parametrize("kx", [1, 3, 5, 7, 9, 11, 13, 15, 17, 21, 31])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_kernel_widths_tiled_and_untiled(kx, border):
    rng = np.random.default_rng(kx)
    src, ker = _case(rng, (3, 37, 131), (3, 5, kx), np.float32, np.float32)
    expected = oracle.convolve(src, ker, np.float32, border)
    same_bits(gpu_convolve(src, ker, np.float32, border), expected)
    same_bits(gpu_convolve(src, ker, np.float32, border, untiled=True), expected)


def test_large_kernel_untiled_fallback():
    rng = np.random.default_rng(4)
    src, ker = _case(rng, (40, 130), (41, 101), np.float32, np.float64)  # tile does not fit in LDS
    expected = oracle.convolve(src, ker, np.float64, cases.ZERO)
    same_bits(gpu_convolve(src, ker, np.float64, cases.ZERO), expected)


def test_strided_and_reversed_views():
    from iane_amd import _lib
    rng = np.random.default_rng(8)
    base = (rng.random((9, 40, 70)) * 10).astype(np.float32)
    ker = (rng.random((3, 3, 5)) - 0.5).astype(np.float32)
    dbase = dev(base)
    dker = dev(ker)
    views = [(base[1:8, 3:35, 2:69], dbase[1:8, 3:35, 2:69]), (base[::2, :, 1::3], dbase[::2, :, 1::3])]
    for border in (cases.CLAMP, cases.ZERO):
        for h, d in views:
            expected = oracle.convolve(h, ker, np.float32, border)
            out_base = empty((h.shape[0], h.shape[1], h.shape[2] * 2), np.float32)
            out = out_base[:, :, ::2]  # strided destination
            iane_amd.convolve(d, dker, out, convolution_options(border_mode(border))).wait()
            same_bits(host(out), expected)
        # negative stride along IANE axis 1 (numpy axis 1): a flipped view of the same storage
        desc = iane_amd.view_desc(dbase)
        desc.ptr = dbase.data_ptr() + (base.shape[1] - 1) * desc.stride_bytes[1]
        desc.stride_bytes[1] = -desc.stride_bytes[1]
        out = empty(base.shape, np.float32)
        _lib.call("iane_hip_convolve", desc, iane_amd.view_desc(dker), iane_amd.view_desc(out), border, 0, 0, 0, 0)
        iane_amd.hip_execution(0).wait()
        same_bits(host(out), oracle.convolve(base[:, ::-1, :], ker, np.float32, border))


def test_streaming_kernel_strided_and_reversed_views():
    """The streaming kernels address planes through buffer resources: sliced and strided
    sources, a strided destination and a source with a negative z stride give the same
    bits as the tiled kernel (FP32 precision, and u16 with an integer kernel)."""
    from iane_amd import _lib, autotune
    rng = np.random.default_rng(9)
    cases_ = [((rng.random((20, 40, 140)) * 10).astype(np.float32), (rng.random((3, 3, 3)) - 0.3).astype(np.float32),
               np.float32, precision.fp32),
              (rng.integers(0, 65536, (20, 40, 140)).astype(np.uint16), rng.integers(-4, 5, (3, 5, 3)).astype(np.int32),
               np.int32, precision.reference)]
    for base, ker, ddt, prec in cases_:
        dbase, dker = dev(base), dev(ker)
        for border in (cases.CLAMP, cases.ZERO):
            for h, d in ((base[1:19, 3:35, 2:131], dbase[1:19, 3:35, 2:131]), (base[::2, :, 1::3], dbase[::2, :, 1::3])):
                outs = []
                for stream in (1, 0):
                    out_base = empty((h.shape[0], h.shape[1], h.shape[2] * 2), ddt)
                    out = out_base[:, :, ::2]  # strided destination
                    with autotune.applied({"conv.stream": stream}):
                        iane_amd.convolve(d, dker, out, convolution_options(border_mode(border), prec)).wait()
                    outs.append(host(out))
                same_bits(outs[0], outs[1])
                if prec == precision.reference:
                    same_bits(outs[0], oracle.convolve(h, ker, ddt, border))
            # negative stride along IANE axis 2 (numpy axis 0): planes in reverse order
            desc = iane_amd.view_desc(dbase)
            desc.ptr = dbase.data_ptr() + (base.shape[0] - 1) * desc.stride_bytes[2]
            desc.stride_bytes[2] = -desc.stride_bytes[2]
            outs = []
            for stream in (1, 0):
                out = empty(base.shape, ddt)
                with autotune.applied({"conv.stream": stream}):
                    _lib.call("iane_hip_convolve", desc, iane_amd.view_desc(dker), iane_amd.view_desc(out), border,
                              int(prec), 0, 0, 0)
                    iane_amd.hip_execution(0).wait()
                outs.append(host(out))
            same_bits(outs[0], outs[1])


def test_empty_image_is_a_no_op():
    import torch
    src = torch.zeros((0, 5), device="cuda")
    dst = torch.zeros((0, 5), device="cuda")
    iane_amd.convolve(src, dev(np.ones((3, 3), np.float32)), dst).wait()


def test_errors_match_reference():
    import torch
    a = torch.zeros((8, 8), device="cuda")
    with pytest.raises(iane_amd.invalid_argument, match="odd"):
        iane_amd.convolve(a, torch.ones((3, 4), device="cuda"), a.clone())
    with pytest.raises(iane_amd.invalid_argument, match="dimensions"):
        iane_amd.convolve(a, torch.ones((3, 3), device="cuda"), torch.zeros((8, 9), device="cuda"))
    with pytest.raises(iane_amd.invalid_argument):
        iane_amd.convolve(a, torch.ones((3,), device="cuda"), a.clone())
    with pytest.raises(iane_amd.invalid_argument, match="host"):
        iane_amd.convolve(a.cpu(), torch.ones((3, 3)), a.cpu())


# ---------------------------------------------------- FP32 precision ------
def rel_err(got, ref64):
    return float(np.max(np.abs(got.astype(np.float64) - ref64)) / max(np.max(np.abs(ref64)), 1e-300))


@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_fp32_precision_dense15(border):
    vol = __import__("iane_amd.synth", fromlist=["x"]).volume_blobs(40, seed=2, blobs=8)
    ref64 = oracle.convolve(vol["volume"], vol["dense"], np.float64, border, threads=8)
    got = gpu_convolve(vol["volume"], vol["dense"], np.float32, border, prec=precision.fp32)
    assert rel_err(got, ref64) <= TOL_FP32


# ---------------------------------------------------- separable -----------
@pytest.mark.parametrize("k", [1, 3, 5, 7, 9, 11, 13, 15])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
@pytest.mark.parametrize("engine", [2, 0, 1], ids=["direct", "staged", "vector"])
def test_separable_fused_fp32(k, border, engine):
    """Every fused engine (x/y filters on FP32 MFMA with per-wave loads or LDS
    staging, or on the vector ALUs); the width is a multiple of 4, so the zero
    border runs the barrier-free engine."""
    from iane_amd import autotune
    rng = np.random.default_rng(k)
    src = rng.random((37, 45, 72)).astype(np.float32)
    taps = [rng.random(k).astype(np.float32) for _ in range(3)]
    t64 = [t.astype(np.float64) for t in taps]
    dense = np.multiply.outer(np.multiply.outer(t64[2], t64[1]), t64[0])
    ref64 = oracle.convolve(src, dense, np.float64, border, threads=8)
    out = empty(src.shape, np.float32)
    with autotune.applied({"sep3.engine": engine}):
        iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                    convolution_options(border_mode(border), precision.fp32)).wait()
    assert rel_err(host(out), ref64) <= TOL_FP32


@pytest.mark.parametrize("shape", [(5, 16, 64), (40, 17, 65), (3, 100, 130), (70, 33, 257), (97, 33, 260),
                                   (130, 7, 4)])
def test_separable_matrix_engine_ragged_tiles(shape):
    """Tile edges in x, y and z (partial 64 x 16 tiles, short and long z chunks);
    widths that are multiples of 4 take the barrier-free engine for the zero border."""
    rng = np.random.default_rng(sum(shape))
    src = (rng.random(shape) - 0.3).astype(np.float32)
    taps = [rng.standard_normal(7).astype(np.float32) for _ in range(3)]
    t64 = [t.astype(np.float64) for t in taps]
    dense = np.multiply.outer(np.multiply.outer(t64[2], t64[1]), t64[0])
    for border in (cases.CLAMP, cases.ZERO):
        ref64 = oracle.convolve(src, dense, np.float64, border, threads=8)
        out = empty(src.shape, np.float32)
        iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                    convolution_options(border_mode(border), precision.fp32)).wait()
        assert rel_err(host(out), ref64) <= TOL_FP32


@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_separable_passes_reference_bit_exact(border):
    """REFERENCE-precision passes equal the oracle's per-axis FP64 convolutions."""
    rng = np.random.default_rng(12)
    src = (rng.random((6, 9, 21)) * 50).astype(np.float32)
    taps = [rng.random(k).astype(np.float64) for k in (5, 3, 7)]
    step = src
    for axis, t in enumerate(taps):
        kshape = [1, 1, 1]
        kshape[2 - axis] = t.size
        last = axis == 2
        step = oracle.convolve(np.ascontiguousarray(step), t.reshape(kshape), np.float32 if last else np.float64,
                               border)
    out = empty(src.shape, np.float32)
    iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                convolution_options(border_mode(border), precision.reference)).wait()
    same_bits(host(out), step)


# ------------------------------------- full size (256^3) properties -------
def test_full_size_delta_shift_and_integer_linearity():
    """At the benchmark size: a delta kernel is an exact shift, and the u16 3^3
    binomial convolution is exactly linear."""
    n = 256
    rng = np.random.default_rng(1)
    vol = rng.random((n, n, n), dtype=np.float32)
    ker = np.zeros((15, 15, 15), np.float32)
    ker[7 + 2, 7 - 3, 7 + 1] = 1.0  # q - c = (+1, -3, +2) in IANE (x, y, z): sample at p - (q - c)
    out = empty(vol.shape, np.float32)
    iane_amd.convolve(dev(vol), dev(ker), out, convolution_options(border_mode.zero)).wait()
    expected = np.zeros_like(vol)
    expected[2:, :n - 3, 1:] = vol[:-2, 3:, :-1]
    same_bits(host(out), expected)
    a = rng.integers(0, 30000, (n, n, n)).astype(np.uint16)
    b = rng.integers(0, 30000, (n, n, n)).astype(np.uint16)
    b3 = np.array([1, 2, 1], np.int32)
    binom = np.multiply.outer(np.multiply.outer(b3, b3), b3).astype(np.int32)
    outs = []
    for x in (a, b, (a.astype(np.int64) + b).astype(np.int32)):
        o = empty(x.shape, np.int32)
        iane_amd.convolve(dev(x), dev(binom), o, convolution_options(border_mode.zero)).wait()
        outs.append(host(o).astype(np.int64))
    np.testing.assert_array_equal(outs[0] + outs[1], outs[2])


def test_full_size_workloads_bit_exact_vs_oracle():
    """The benchmark workloads that run on the streaming and tiled direct kernels, at full
    size, against the C oracle: the 1024x1024 image with the 7x7 Gaussian (clamp,
    reference precision) and its u8 quantisation with the 5x5 box into i32, and the
    u16 256^3 volume with the 3^3 binomial kernel into i32 (zero border)."""
    from iane_amd import synth
    img, gauss7, img_u8, box5 = synth.image_2d_uniform()
    got = gpu_convolve(img, gauss7, np.float32, cases.CLAMP)
    same_bits(got, oracle.convolve(img, gauss7, np.float32, oracle.CLAMP, threads=8))
    got = gpu_convolve(img_u8, box5, np.int32, cases.CLAMP)
    same_bits(got, oracle.convolve(img_u8, box5, np.int32, oracle.CLAMP, threads=8))
    data = synth.volume_blobs(256, seed=2)
    got = gpu_convolve(data["volume_u16"], data["binomial"], np.int32, cases.ZERO)
    same_bits(got, oracle.convolve(data["volume_u16"], data["binomial"], np.int32, oracle.ZERO, threads=8))


def test_full_size_separable_vs_fp64_oracle():
    """The headline workload (separable 15^3 Gaussian on the 256^3 blob volume,
    zero border, FP32) over the whole volume against three FP64 per-axis passes
    of the C oracle (a separable kernel with the zero border factors over axes):
    max|gpu - ref| <= TOL_FP32 * max|ref| (measured: ~2.2e-7; normwise ~7e-8)."""
    from iane_amd import synth
    vol = synth.volume_blobs(256, seed=2)
    v = vol["volume"]
    taps = vol["taps"]
    out = empty(v.shape, np.float32)
    iane_amd.convolve_separable(dev(v), [dev(taps)] * 3, out,
                                convolution_options(border_mode.zero, precision.fp32)).wait()
    got = host(out).astype(np.float64)
    ref = v.astype(np.float64)
    t = taps.astype(np.float64)
    for shape in ((1, 1, 15), (1, 15, 1), (15, 1, 1)):
        ref = oracle.convolve(ref, t.reshape(shape), np.float64, oracle.ZERO, threads=16)
    err = rel_err(got, ref)
    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"separable 15^3 at 256^3 vs FP64 oracle: max {err:.2e}, normwise {l2:.2e}")
    assert err <= TOL_FP32


@pytest.mark.parametrize("shape", [(20, 37, 68), (5, 33, 128), (70, 18, 64), (1, 5, 4), (13, 100, 196)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_column_kernel_equals_tiled(shape, border):
    """The column kernel (3x3x3 kernels, rows aligned for 4-sample vectors, x extent
    a multiple of 4) against the streaming and the tiled kernel: identical bits for
    FP32 sums and for integer images with integer kernels, whether the int32 sum is
    exact in FP32 (binomial-like kernels), in int32 only, or coefficients up to 2^21
    force the FP64 chain; both rows-per-lane settings and z chunks of 1, 5 and the
    default (ragged last chunks included); integer sums with odd chunks marching
    down (the default) and with every chunk marching up."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(shape) + border)
    for sdt, kdt, ddt, prec in [(np.float32, np.float32, np.float32, precision.fp32),
                                (np.uint16, np.int32, np.int32, precision.reference),
                                (np.uint16, np.uint8, np.uint16, precision.reference),
                                (np.uint8, np.uint8, np.int32, precision.reference),
                                (np.uint8, np.int32, np.uint8, precision.reference),
                                (np.uint8, np.uint16, np.float32, precision.reference),
                                (np.uint16, "mid", np.int32, precision.reference),
                                (np.uint16, "big", np.int32, precision.reference)]:
        if kdt in ("mid", "big"):
            src, _ = _case(rng, shape, (3, 3, 3), sdt, np.int32)
            bound = 1000 if kdt == "mid" else 1 << 21
            ker = rng.integers(-bound, bound, (3, 3, 3)).astype(np.int32)
        else:
            src, ker = _case(rng, shape, (3, 3, 3), sdt, kdt)
        outs = []
        for config in ({"conv.stream": 0}, {"conv.column": 0}, {}, {"conv.column_rows": 2},
                       {"conv.column_chunk": 1}, {"conv.column_chunk": 5},
                       {"conv.column_chunk": 5, "conv.column_alternate": 0}):
            with autotune.applied(config):
                outs.append(gpu_convolve(src, ker, ddt, border, prec))
        for out in outs[1:]:
            same_bits(out, outs[0])
        if prec == precision.reference:
            same_bits(outs[0], oracle.convolve(src, ker, ddt, border))


@pytest.mark.parametrize("shape", [(37, 68), (33, 128), (18, 64), (5, 4), (100, 196), (3, 21, 44), (300, 1024)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
@pytest.mark.parametrize("k", [3, 5, 7, 9])
def test_plane_kernel_equals_tiled(shape, border, k):
    """The plane kernel (square single-plane kernels on rows aligned for 4-sample
    vectors, x extent a multiple of 4; 2-D images and stacks of planes) against the
    tiled kernel: identical bits for FP32 sums, for the reference's FP64 chain, and
    for integer images with integer kernels whether the sum is exact in FP32, in
    int32 only, or coefficients up to 2^21 force the FP64 chain; 1, 2 and 4 rows per
    lane. REFERENCE precision and integer images also equal the oracle bit for bit,
    FP32 is within TOL_FP32 of the FP64 oracle."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(shape) + 10 * border + k)
    kshape = (1,) * (len(shape) - 2) + (k, k)
    combos = [(np.float32, np.float32, np.float32, precision.fp32),
              (np.float32, np.float32, np.float32, precision.reference),
              (np.float32, np.float64, np.float32, precision.reference),
              (np.uint8, np.int32, np.int32, precision.reference),
              (np.uint8, np.uint8, np.uint8, precision.reference),
              (np.uint16, np.uint16, np.uint16, precision.reference),
              (np.uint16, "mid", np.int32, precision.reference),
              (np.uint8, "big", np.int32, precision.reference)]
    for sdt, kdt, ddt, prec in combos:
        if kdt in ("mid", "big"):
            src, _ = _case(rng, shape, kshape, sdt, np.int32)
            bound = 400 if kdt == "mid" else 1 << 21
            ker = rng.integers(-bound, bound, kshape).astype(np.int32)
        else:
            src, ker = _case(rng, shape, kshape, sdt, kdt)
        outs = []
        for config in ({"conv.plane": 0}, {}, {"conv.plane_rows": 1}, {"conv.plane_rows": 2},
                       {"conv.plane_rows": 4}):
            with autotune.applied(config):
                outs.append(gpu_convolve(src, ker, ddt, border, prec))
        for out in outs[1:]:
            same_bits(out, outs[0])
        if prec == precision.reference:
            same_bits(outs[0], oracle.convolve(src, ker, ddt, border))
        else:
            ref = oracle.convolve(src, ker, np.float64, border)
            assert np.max(np.abs(outs[0] - ref)) <= TOL_FP32 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape,kshape", [((70, 90), (31, 31)), ((9, 40, 52), (3, 21, 27)), ((45, 77), (17, 23))])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_tiled_kernel_tile_heights(shape, kshape, border):
    """Kernels large enough that the FP64 64 x 64 tile's LDS would not fit run the
    tiled kernel with 64 x 32 tiles (and FP32 still with 64 x 64): REFERENCE
    precision bit for bit against the oracle, FP32 within TOL_FP32."""
    rng = np.random.default_rng(sum(shape) + sum(kshape) + border)
    src = (rng.random(shape) * 10 - 2).astype(np.float32)
    ker = (rng.random(kshape) - 0.4).astype(np.float32)
    same_bits(gpu_convolve(src, ker, np.float32, border, precision.reference),
              oracle.convolve(src, ker, np.float32, border))
    ref = oracle.convolve(src, ker, np.float64, border)
    got = gpu_convolve(src, ker, np.float32, border, precision.fp32)
    assert np.max(np.abs(got - ref)) <= TOL_FP32 * np.max(np.abs(ref))


def test_plane_kernel_views():
    """Views the plane kernel must decline or handle: rows that are not 16-byte
    aligned, a strided x axis, an x extent not a multiple of 4, a strided
    destination, row offsets into a larger image: the tiled kernel's bits."""
    from iane_amd import autotune
    rng = np.random.default_rng(6)
    base = (rng.random((90, 136)) * 10).astype(np.float32)
    ker = (rng.random((7, 7)) - 0.3).astype(np.float32)
    dbase, dker = dev(base), dev(ker)
    for prec in (precision.fp32, precision.reference):
        opts = convolution_options(border_mode.clamp, prec)
        for h, d in ((base[2:88, 1:129], dbase[2:88, 1:129]), (base[:, ::2], dbase[:, ::2]),
                     (base[:, 4:130], dbase[:, 4:130]), (base[:, 0:126], dbase[:, 0:126]),
                     (base[3:, 8:], dbase[3:, 8:])):
            outs = []
            for config in ({"conv.plane": 0}, {}):
                out_base = empty((h.shape[0], h.shape[1] * 2), np.float32)
                for out in (out_base[:, ::2], out_base[:, :h.shape[1]]):
                    with autotune.applied(config):
                        iane_amd.convolve(d, dker, out, opts).wait()
                    outs.append(host(out))
            for out in outs[1:]:
                same_bits(out, outs[0])


def test_column_kernel_views():
    """Views the column kernel must decline or handle: a source slice whose rows
    are not 16-byte aligned, a strided x axis, an x extent not a multiple of 4, a
    strided destination, and a negative z stride give the tiled kernel's bits."""
    from iane_amd import _lib, autotune
    rng = np.random.default_rng(5)
    base = (rng.random((12, 40, 136)) * 10).astype(np.float32)
    ker = (rng.random((3, 3, 3)) - 0.3).astype(np.float32)
    dbase, dker = dev(base), dev(ker)
    opts = convolution_options(border_mode.zero, precision.fp32)
    for h, d in ((base[1:11, 2:38, 1:129], dbase[1:11, 2:38, 1:129]), (base[:, :, ::2], dbase[:, :, ::2]),
                 (base[:, :, 4:130], dbase[:, :, 4:130]), (base[:, :, 0:126], dbase[:, :, 0:126]),
                 (base[:, 3:, 8:], dbase[:, 3:, 8:])):
        outs = []
        for config in ({"conv.stream": 0}, {}):
            out_base = empty((h.shape[0], h.shape[1], h.shape[2] * 2), np.float32)
            for out in (out_base[:, :, ::2], out_base[:, :, :h.shape[2]]):
                with autotune.applied(config):
                    iane_amd.convolve(d, dker, out, opts).wait()
                outs.append(host(out))
        for out in outs[1:]:
            same_bits(out, outs[0])
    desc = iane_amd.view_desc(dbase)
    desc.ptr = dbase.data_ptr() + (base.shape[0] - 1) * desc.stride_bytes[2]
    desc.stride_bytes[2] = -desc.stride_bytes[2]
    outs = []
    for config in ({"conv.stream": 0}, {}):
        out = empty(base.shape, np.float32)
        with autotune.applied(config):
            _lib.call("iane_hip_convolve", desc, iane_amd.view_desc(dker), iane_amd.view_desc(out), cases.ZERO,
                      int(precision.fp32), 0, 0, 0)
            iane_amd.hip_execution(0).wait()
        outs.append(host(out))
    same_bits(outs[0], outs[1])
    same_bits(outs[1], gpu_convolve(base[::-1], ker, np.float32, cases.ZERO, precision.fp32))


@pytest.mark.parametrize("shape,kshape", [((40, 37, 70), (3, 3, 3)), ((9, 20, 130), (5, 3, 7)), ((5, 33, 64), (7, 5, 1)),
                                          ((3, 17, 65), (5, 5, 5)), ((70, 16, 64), (3, 7, 3))])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_streaming_kernel_equals_tiled(shape, kshape, border):
    """The z-streaming direct kernels (small 3-D kernels; unrolled or looped taps, chunk
    lengths with and without a partial pipeline tail) and the per-plane tiled kernel
    give identical bits in every precision and dtype pair. Integer images with integer
    kernels accumulate in int32 when the sum provably fits, else in FP64: both cases run
    (the last pair's coefficients up to 2^21 break the int32 bound)."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(shape) + sum(kshape))
    for sdt, kdt, ddt, prec in [(np.float32, np.float32, np.float32, precision.reference),
                                (np.float32, np.float32, np.float32, precision.fp32),
                                (np.uint16, np.int32, np.int32, precision.reference),
                                (np.uint8, np.uint8, np.int32, precision.reference),
                                (np.uint8, np.float32, np.uint8, precision.reference),
                                (np.uint16, "big", np.float64, precision.reference)]:
        if kdt == "big":
            src, _ = _case(rng, shape, kshape, sdt, np.int32)
            ker = rng.integers(-(1 << 21), 1 << 21, kshape).astype(np.int32)
        else:
            src, ker = _case(rng, shape, kshape, sdt, kdt)
        outs = []
        for config in ({"conv.stream": 0}, {}, {"conv.stream_unroll": 0}, {"conv.stream_chunk": 5},
                       {"conv.stream_chunk": 3}):
            with autotune.applied(config):
                outs.append(gpu_convolve(src, ker, ddt, border, prec))
        for out in outs[1:]:
            same_bits(out, outs[0])
        if prec == precision.reference:
            same_bits(outs[0], oracle.convolve(src, ker, ddt, border))



@pytest.mark.parametrize("lengths", [(3, 7, 15), (15, 5, 1), (1, 13, 9), (11, 11, 3)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
@pytest.mark.parametrize("engine", [2, 0, 1], ids=["direct", "staged", "vector"])
def test_separable_unequal_tap_lengths(lengths, border, engine):
    """Tap vectors of different odd lengths (x, y, z) run on the single-kernel
    engines, each centred in the window of the longest (anisotropic Gaussians)."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(lengths))
    src = (rng.random((23, 37, 68)) - 0.2).astype(np.float32)
    taps = [rng.standard_normal(k).astype(np.float32) for k in lengths]
    t64 = [t.astype(np.float64) for t in taps]
    dense = np.multiply.outer(np.multiply.outer(t64[2], t64[1]), t64[0])
    ref64 = oracle.convolve(src, dense, np.float64, border, threads=8)
    out = empty(src.shape, np.float32)
    with autotune.applied({"sep3.engine": engine}):
        iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                    convolution_options(border_mode(border), precision.fp32)).wait()
    assert rel_err(host(out), ref64) <= TOL_FP32


@pytest.mark.parametrize("shape", [(1024, 1024), (45, 72), (300, 20), (7, 4), (64, 130)])
@pytest.mark.parametrize("lengths", [(7, 7), (15, 3), (1, 9)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_separable_rank2(shape, lengths, border):
    """2-D images on the single-kernel engines (one plane, a one-tap z filter)."""
    rng = np.random.default_rng(shape[0] + 7 * lengths[0])
    src = rng.random(shape).astype(np.float32)
    taps = [rng.random(k).astype(np.float32) for k in lengths]
    dense = np.multiply.outer(taps[1].astype(np.float64), taps[0].astype(np.float64))
    ref64 = oracle.convolve(src, dense, np.float64, border, threads=8)
    for engine in (2, 0, 1):
        from iane_amd import autotune
        out = empty(src.shape, np.float32)
        with autotune.applied({"sep3.engine": engine}):
            iane_amd.convolve_separable(dev(src), [dev(t) for t in taps], out,
                                        convolution_options(border_mode(border), precision.fp32)).wait()
        assert rel_err(host(out), ref64) <= TOL_FP32, engine


@pytest.mark.parametrize("shape", [(20, 37, 68), (5, 33, 128), (70, 18, 64), (1, 5, 4), (13, 100, 196), (9, 64, 256)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
@pytest.mark.parametrize("taps", ["binomial", "random", "large", "sparse"])
def test_column_u16_dot2_path(shape, border, taps):
    """u16 images with 3^3 kernels of nonnegative coefficients below 2^16 (sums below
    2^31) run on packed v_dot2_u32_u16 pairs: the same integers as the exact FP32 /
    int32 column paths and as the oracle (bit-exact, i32 and u16 destinations)."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(shape) + len(taps))
    src = rng.integers(0, 65536, shape).astype(np.uint16)
    if taps == "binomial":
        b = np.array([1, 2, 1])
        ker = np.multiply.outer(np.multiply.outer(b, b), b).astype(np.int32)
    elif taps == "random":
        ker = rng.integers(0, 300, (3, 3, 3)).astype(np.int32)
    elif taps == "large":
        ker = rng.integers(1000, 1200, (3, 3, 3)).astype(np.int32)  # 27 * 1199 * 65535 < 2^31
    else:
        ker = np.zeros((3, 3, 3), np.int32)
        ker[0, 1, 2] = 65535 // 64
        ker[2, 0, 0] = 7
    ref = oracle.convolve(src, ker, np.int32, border, threads=8)
    with autotune.applied({"conv.column": 1, "conv.column_dot2": 1}):
        got = gpu_convolve(src, ker, np.int32, border)
    same_bits(got, ref)
    with autotune.applied({"conv.column": 1, "conv.column_dot2": 0}):
        same_bits(gpu_convolve(src, ker, np.int32, border), ref)
    with autotune.applied({"conv.column": 1, "conv.column_chunk": 3, "conv.column_alternate": 0}):
        same_bits(gpu_convolve(src, ker, np.int32, border), ref)
    with autotune.applied({"conv.column": 1, "conv.column_chunk": 3}):
        same_bits(gpu_convolve(src, ker, np.int32, border), ref)
    if taps == "binomial":
        with autotune.applied({"conv.column": 1}):
            same_bits(gpu_convolve(src, ker, np.uint16, border), oracle.convolve(src, ker, np.uint16, border, threads=8))
