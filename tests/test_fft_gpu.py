"""GPU parity of the Stockham FFT drop-ins (fft_forward / fft_inverse / rfft / irfft).

Known answers restated from test/unit/prepared/prepared_cases.hpp:73-123,
test_prepared.cpp:39-59,160-183 and real_fft_cases.hpp:12-77; larger sizes
against the FP64 oracle (numpy.fft, itself pinned by a brute-force DFT in
tests/test_oracle_fft.py). Tolerance for FP32 transforms:
    max|gpu - oracle64| <= 1e-5 * max|oracle64|   (TOL_FFT)
plus the reference's own absolute tolerances on its small cases.
"""
from __future__ import annotations

import numpy as np
import pytest

import iane_amd
from iane_amd import execution_context, fft_forward, fft_inverse, irfft, rfft, rfft_extent
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL_FFT = 1e-5


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def cempty(shape):
    import torch
    return torch.full(tuple(shape), complex(7, 7), dtype=torch.complex64, device="cuda")


def rempty(shape):
    import torch
    return torch.full(tuple(shape), 7.0, dtype=torch.float32, device="cuda")


def rel(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


# ------------------------------------------------------ known answers -----
def test_c2c_delta_known_answer_and_round_trip():
    """prepared_cases.hpp:73-123: delta at index 1 of a 4x2 image."""
    x = np.zeros((2, 4), np.complex64)
    x[0, 1] = 1.0
    src, out, restored = dev(x), cempty((2, 4)), cempty((2, 4))
    ctx = execution_context()
    fwd = ctx.prepare(fft_forward, src, out)
    inv = ctx.prepare(fft_inverse, out, restored)
    ctx.run(fwd, src, out)
    ctx.run(inv, out, restored)
    ctx.run(fwd, src, out).wait()
    got = host(out)
    real = np.array([1, 0, -1, 0], np.float32)
    imag = np.array([0, -1, 0, 1], np.float32)
    for row in range(2):
        np.testing.assert_allclose(got[row].real, real, atol=1e-5)
        np.testing.assert_allclose(got[row].imag, imag, atol=1e-5)
    rt = host(restored)
    np.testing.assert_allclose(rt.real, 8 * x.real, atol=1e-5)
    np.testing.assert_allclose(rt.imag, 0, atol=1e-5)
    assert ctx.scratch_capacity() == 0


def test_c2c_constant_dc():
    """test_prepared.cpp:39-59: constant 2 on 4x2 -> DC 16, zeros elsewhere."""
    x = np.full((2, 4), 2.0, np.complex64)
    src, out = dev(x), cempty((2, 4))
    ctx = execution_context()
    op = ctx.prepare(fft_forward, src, out)
    ctx.run(op, src, out).wait()
    got = host(out)
    expected = np.zeros((2, 4), np.complex64)
    expected[0, 0] = 16
    np.testing.assert_allclose(got, expected, atol=1e-5)


def test_c2c_aliasing_and_shape_errors():
    x = dev(np.zeros((2, 4), np.complex64))
    ctx = execution_context()
    with pytest.raises(iane_amd.invalid_argument):
        op = ctx.prepare(fft_forward, x, x)
        ctx.run(op, x, x)
    out = cempty((2, 4))
    op = ctx.prepare(fft_forward, x, out)
    with pytest.raises(iane_amd.invalid_argument, match="prepared shape"):
        ctx.run(op, dev(np.zeros((4, 2), np.complex64)), cempty((4, 2)))
    other = execution_context()
    with pytest.raises(iane_amd.invalid_argument, match="different execution context"):
        other.run(op, x, out)
    with pytest.raises(iane_amd.invalid_argument, match="packed"):
        ctx.prepare(fft_forward, dev(np.zeros((2, 8), np.complex64))[:, ::2], out)


@pytest.mark.parametrize("shape", [(3, 4), (3, 5), (4, 6)])
def test_real_fft_known_answer(shape):
    """real_fft_cases.hpp:12-77: sizes 4x3, 5x3, 6x4 (IANE {w, h})."""
    count = shape[0] * shape[1]
    x = ((np.arange(count) * 7) % 11 - 4).astype(np.float32).reshape(shape)
    packed_shape = rfft_extent(shape)
    src, packed, out = dev(x), cempty(packed_shape), rempty(shape)
    ctx = execution_context()
    fwd = ctx.prepare(rfft, src, packed)
    inv = ctx.prepare(irfft, packed, out)
    ctx.run(fwd, src, packed)
    before = host(packed)
    ctx.run(inv, packed, out).wait()
    after = host(packed)
    np.testing.assert_allclose(host(out), count * x, atol=2e-4)
    np.testing.assert_array_equal(host(src), x)
    brute = oracle.dft_bruteforce(x.astype(np.float64))
    np.testing.assert_allclose(before.real, brute[:, :packed_shape[1]].real, atol=3e-5)
    np.testing.assert_allclose(before.imag, brute[:, :packed_shape[1]].imag, atol=3e-5)
    np.testing.assert_array_equal(after.view(np.uint32), before.view(np.uint32))  # C2R preserved input


def test_real_fft_layout_errors():
    ctx = execution_context()
    x = dev(np.zeros((3, 4), np.float32))
    with pytest.raises(iane_amd.invalid_argument, match="matching real and packed"):
        ctx.prepare(rfft, x, cempty((3, 4)))
    op = ctx.prepare(rfft, x, cempty((3, 3)))
    with pytest.raises(iane_amd.invalid_argument):
        ctx.run(op, dev(np.zeros((3, 6), np.float32)), cempty((3, 4)))


# ------------------------------------------------------ oracle sweeps -----
LENGTHS = [1, 2, 3, 5, 7, 8, 12, 16, 17, 30, 31, 64, 90, 128, 143, 160, 224, 280, 288, 512, 1000, 2048]


@pytest.mark.parametrize("n", LENGTHS)
def test_c2c_1d_lengths(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))).astype(np.complex64)
    ctx = execution_context()
    for alg, ref in ((fft_forward, np.fft.fft(x.astype(np.complex128), axis=-1)),
                     (fft_inverse, np.fft.ifft(x.astype(np.complex128), axis=-1) * n)):
        # a batch of 5 independent 1-D lines = rank-1 plans on each row
        for row in range(5):
            src, out = dev(x[row]), cempty((n,))
            op = ctx.prepare(alg, src, out)
            ctx.run(op, src, out).wait()
            assert rel(host(out), ref[row]) <= TOL_FFT * max(1.0, np.log2(max(n, 2)) / 4)


@pytest.mark.parametrize("shape", [(6, 10), (9, 16, 20), (16, 30, 28), (7, 5, 3), (64, 64, 64)])
def test_c2c_nd(shape):
    rng = np.random.default_rng(len(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ctx = execution_context()
    src, out = dev(x), cempty(shape)
    op = ctx.prepare(fft_forward, src, out)
    ctx.run(op, src, out).wait()
    assert rel(host(out), oracle.fft_forward(x)) <= TOL_FFT
    back = cempty(shape)
    op2 = ctx.prepare(fft_inverse, out, back)
    ctx.run(op2, out, back).wait()
    assert rel(host(back), x * np.prod(shape)) <= TOL_FFT


@pytest.mark.parametrize("shape", [(7,), (8,), (31,), (3, 9), (10, 12), (5, 6, 7), (9, 20, 30), (40, 56, 70),
                                   (160, 280, 280)])
def test_r2c_c2r_nd(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape).astype(np.float32)
    spec_shape = rfft_extent(shape)
    ctx = execution_context()
    src, spec, back = dev(x), cempty(spec_shape), rempty(shape)
    f = ctx.prepare(rfft, src, spec)
    b = ctx.prepare(irfft, spec, back)
    ctx.run(f, src, spec)
    saved = host(spec)
    ctx.run(b, spec, back).wait()
    assert rel(saved, oracle.rfft(x)) <= TOL_FFT
    assert rel(host(back), x.astype(np.float64) * np.prod(shape)) <= TOL_FFT
    np.testing.assert_array_equal(host(spec).view(np.uint32), saved.view(np.uint32))
    # Parseval on the full spectrum (size independent)
    if len(shape) == 3 and shape[0] >= 100:
        full = oracle.rfft(x)
        energy = np.sum(np.abs(saved[..., 0]) ** 2) + 2 * np.sum(np.abs(saved[..., 1:]) ** 2)
        if shape[-1] % 2 == 0:
            energy -= np.sum(np.abs(saved[..., -1]) ** 2)
        assert abs(energy / np.prod(shape) - np.sum(x.astype(np.float64) ** 2)) <= 1e-4 * np.sum(x ** 2)
        del full


# ------------------------------------- Bluestein and long lines -----
def _dft_direct(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """O(n^2) float64 DFT of one line, in blocks of 256 bins, with the angle of
    each term reduced exactly (j k mod n) before scaling."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    j = np.arange(n, dtype=np.int64)
    out = np.empty(n, dtype=np.complex128)
    for k0 in range(0, n, 256):
        k = np.arange(k0, min(n, k0 + 256), dtype=np.int64)
        phase = (np.outer(k, j) % n).astype(np.float64) * (sign * 2.0 * np.pi / n)
        out[k0:k0 + k.size] = np.exp(1j * phase) @ x
    return out


@pytest.mark.parametrize("n", [67, 127, 1009, 2053, 4099, 4096, 8192, 10000])
def test_c2c_prime_and_long_lengths_vs_direct_dft(n):
    """Lengths with a prime factor above 64 run Bluestein's chirp-z form (67, 127,
    1009, 2053, 4099 are prime); lines past 4 095 points use the whole LDS with the
    twiddles read from memory (4096, 8192, 10000). Forward and inverse against the
    direct float64 DFT: max|gpu - dft| <= 1e-5 max|dft|."""
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ctx = execution_context()
    for alg, sign in ((fft_forward, -1), (fft_inverse, 1)):
        src, out = dev(x), cempty((n,))
        op = ctx.prepare(alg, src, out)
        ctx.run(op, src, out).wait()
        err = rel(host(out), _dft_direct(x, sign))
        print(f"n={n} sign={sign}: {err:.2e}")
        assert err <= TOL_FFT


@pytest.mark.parametrize("shape", [(1009,), (2018,), (8192,), (8190,), (5, 4099), (131, 20), (3, 67, 10), (12, 2, 1031)])
def test_prime_and_long_axes_nd(shape):
    """Bluestein and long lines on every axis and through the real transforms: an
    odd prime real length (1009), an even one whose half length is prime (2018),
    long even lines (8192, and 8190 = 2 x 4095 with half length 4095), and prime
    extents on strided axes; C2C and R2C/C2R against the float64 FFT (itself
    pinned by the brute-force DFT), round trips included."""
    rng = np.random.default_rng(sum(shape))
    ctx = execution_context()
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    src, out, back = dev(x), cempty(shape), cempty(shape)
    ctx.run(ctx.prepare(fft_forward, src, out), src, out).wait()
    assert rel(host(out), oracle.fft_forward(x)) <= TOL_FFT
    ctx.run(ctx.prepare(fft_inverse, out, back), out, back).wait()
    assert rel(host(back), x * np.prod(shape)) <= TOL_FFT
    r = rng.standard_normal(shape).astype(np.float32)
    spec_shape = rfft_extent(shape)
    rsrc, spec, rback = dev(r), cempty(spec_shape), rempty(shape)
    ctx.run(ctx.prepare(rfft, rsrc, spec), rsrc, spec).wait()
    assert rel(host(spec), oracle.rfft(r)) <= TOL_FFT
    ctx.run(ctx.prepare(irfft, spec, rback), spec, rback).wait()
    assert rel(host(rback), r.astype(np.float64) * np.prod(shape)) <= TOL_FFT


@pytest.mark.parametrize("shape", [(3, 128), (4, 256), (2, 512), (1024,), (12, 280), (5, 160, 140), (20, 512)])
def test_compile_time_plans_give_the_run_time_bits(shape):
    """Line lengths with a compile-time plan (140, 160, 280, 128 .. 1024) run the
    same operations in the same order as the run-time stages: identical bits for
    C2C both ways and for R2C/C2R ("fft.fixed" switches the table off)."""
    from iane_amd import autotune
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    r = rng.standard_normal(shape).astype(np.float32)
    outs = []
    for fixed in (0, 1):
        with autotune.applied({"fft.fixed": fixed}):
            ctx = execution_context()
            src, out, back = dev(x), cempty(shape), cempty(shape)
            ctx.run(ctx.prepare(fft_forward, src, out), src, out)
            ctx.run(ctx.prepare(fft_inverse, out, back), out, back)
            rsrc, spec, rback = dev(r), cempty(rfft_extent(shape)), rempty(shape)
            ctx.run(ctx.prepare(rfft, rsrc, spec), rsrc, spec)
            ctx.run(ctx.prepare(irfft, spec, rback), spec, rback).wait()
            outs.append([host(t) for t in (out, back, spec, rback)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert rel(outs[1][0], oracle.fft_forward(x)) <= TOL_FFT


@pytest.mark.parametrize("n", [16384, 20000, 10006, 65536])
def test_c2c_four_step_lines(n):
    """Lines longer than two copies fit in LDS run as two passes (four-step:
    n = n1 n2, a twiddle between them, the second pass storing the transposed
    order through the plan's workspace): 16384 = 128 x 128, 20000, 10006 =
    2 x 5003 (the prime factor in Bluestein's form), 65536 = 256 x 256. Forward
    and inverse against the direct float64 DFT (numpy's FFT for 65536)."""
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ctx = execution_context()
    for alg, sign in ((fft_forward, -1), (fft_inverse, 1)):
        src, out = dev(x), cempty((n,))
        op = ctx.prepare(alg, src, out)
        assert op.state.scratch_bytes() == n * 8
        ctx.run(op, src, out).wait()
        if n <= 20000:
            ref = _dft_direct(x, sign)
        else:
            ref = np.fft.fft(x.astype(np.complex128)) if sign < 0 else np.fft.ifft(x.astype(np.complex128)) * n
        err = rel(host(out), ref)
        print(f"n={n} sign={sign}: {err:.2e}")
        assert err <= TOL_FFT


@pytest.mark.parametrize("shape", [(12000, 3), (2, 10006, 5), (13000, 4), (3, 16384)])
def test_four_step_axes_nd(shape):
    """Long strided axes (12000 and 10006 along axis 1, 13000 along the last
    numpy axis of a real transform's spectrum) and a long axis 0 (16384): C2C
    both ways and, where axis 0 is short, R2C/C2R (the inverse's second
    workspace region), against the float64 FFT, round trips included."""
    rng = np.random.default_rng(sum(shape))
    ctx = execution_context()
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    src, out, back = dev(x), cempty(shape), cempty(shape)
    ctx.run(ctx.prepare(fft_forward, src, out), src, out).wait()
    assert rel(host(out), oracle.fft_forward(x)) <= TOL_FFT
    ctx.run(ctx.prepare(fft_inverse, out, back), out, back).wait()
    assert rel(host(back), x * np.prod(shape)) <= TOL_FFT
    if shape[-1] <= 20000:
        r = rng.standard_normal(shape).astype(np.float32)
        spec_shape = rfft_extent(shape)
        rsrc, spec, rback = dev(r), cempty(spec_shape), rempty(shape)
        ctx.run(ctx.prepare(rfft, rsrc, spec), rsrc, spec).wait()
        assert rel(host(spec), oracle.rfft(r)) <= TOL_FFT
        ctx.run(ctx.prepare(irfft, spec, rback), spec, rback).wait()
        assert rel(host(rback), r.astype(np.float64) * np.prod(shape)) <= TOL_FFT


def test_line_too_long_for_lds_is_refused():
    """A line with no factorisation into LDS-sized transforms (a prime above the
    Bluestein limit) is refused with the unsupported status, not run; so is a
    real line too long for one pass (real lines are not split)."""
    from iane_amd import _lib
    ctx = execution_context()
    x = dev(np.zeros(40009, np.complex64))  # prime: Bluestein length > 80 000
    out = cempty((40009,))
    with pytest.raises(_lib.unsupported):
        ctx.prepare(fft_forward, x, out)
    r = dev(np.zeros(50000, np.float32))
    spec = cempty(rfft_extent((50000,)))
    with pytest.raises(_lib.unsupported):
        ctx.prepare(rfft, r, spec)


def test_fft_workspace_reported_for_c2r():
    ctx = execution_context()
    spec, back = cempty((6, 9)), rempty((6, 16))
    op = ctx.prepare(irfft, spec, back)
    assert ctx.scratch_capacity() == 6 * 9 * 8
    assert op.state.scratch_bytes() == 6 * 9 * 8
