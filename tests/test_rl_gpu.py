"""GPU parity of the Richardson–Lucy and Wiener loops.

* DIRECT method, REFERENCE precision: bit-identical to the reference RL loop
  (golden fixtures written by the reference's own convolve/zip/copy).
* FFT method: against the FP64 oracle (FFT restatement of the loop, itself
  pinned in tests/test_oracle_fft.py). The FP32 FFT's rounding is absolute (a
  fraction of the signal's norm), so the per-voxel relative error grows with
  the iteration count, linearly in the measurements (about 1.1-2.2e-7 per
  iteration: 1.0-1.6e-6 after 10 iterations, 1.4-4.3e-5 after 200; the
  reference's own FP32 loop with FP64-accumulated direct convolutions measures
  2.2-4.1e-6 after 200; 6.4e-7 after 2 iterations at 256x256x128). The bound is
      max|gpu - oracle64| <= tol_rl(i) * max|oracle64|,  tol_rl(i) = 1e-6 + 2.5e-7 * i,
  and |ll_gpu - ll64| <= 1e-6 |ll64| for the log-likelihood.
* DIRECT method with split-precision MFMA convolution (hi + lo FP16 / BF16 operands,
  FP32 accumulation): against the FP64 oracle, max|gpu - oracle64| <= 2e-5 (FP16) and
  <= 1e-4 (BF16) of max|oracle64| after 5 iterations with a 9x7x21 PSF (measured:
  about 3e-6 for both); after 200 iterations with the 31x31x63 PSF on a 64^3
  volume, <= 3e-4 (measured 1.6e-4 FP16, 2.2e-4 BF16).
* Pointwise steps: bit-identical to the same float32 operations on the host.
* 200 iterations with the 31x31x63 PSF (sigma 1.5, 1.5, 3.0) on a 64^3 volume of
  the same construction as the 256x256x128 workload: FFT method and split FP16 / BF16
  direct method against the FP64 oracle.
* Full size (256x256x128, PSF 31x31x63): FFT RL vs the FP64 oracle for 2 iterations,
  and the EM property (non-decreasing log-likelihood) over 20 iterations.
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

from iane_amd import deconvolution, precision, synth
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def tol_rl(iterations: int) -> float:
    return 1e-6 + 2.5e-7 * iterations


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def rel(got, ref):
    return float(np.max(np.abs(got.astype(np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("rl_*.npz")), ids=lambda p: p.stem)
def test_direct_reference_precision_bit_exact(path):
    data = np.load(path, allow_pickle=False)
    est = deconvolution.richardson_lucy(dev(data["observed"]), dev(data["psf"]), int(data["iterations"]),
                                        float(data["epsilon"]), method=deconvolution.DIRECT,
                                        precision=precision.reference)
    np.testing.assert_array_equal(host(est).view(np.uint32), data["expected"].view(np.uint32))


def _case(shape, psf_sigma, psf_shape, peak=500.0, seed=3):
    truth = np.full(shape, 0.05)
    synth.add_blobs(truth, 6, (1.0, 3.0), (0.3, 1.0), seed, 1)
    psf = synth.gaussian_kernel(psf_sigma, psf_shape)
    blurred = oracle.conv_zero_fft(truth * peak, psf)
    y = synth.poisson(np.maximum(blurred, 0), seed, 2).reshape(shape).astype(np.float32)
    return y, psf.astype(np.float32)


@pytest.mark.parametrize("shape,psf_shape", [((40, 50), (7, 9)), ((33, 21), (5, 15)), ((20, 24, 28), (5, 7, 9)),
                                             ((17, 30, 26), (11, 5, 3))])
def test_fft_method_vs_fp64_oracle(shape, psf_shape):
    y, psf = _case(shape, [1.2] * len(shape), psf_shape)
    iters = 10
    est, ll = deconvolution.richardson_lucy(dev(y), dev(psf), iters, method=deconvolution.FFT, return_ll=True)
    ref, ll_ref = oracle.rl(y, psf.astype(np.float64), iters)
    assert rel(host(est), ref) <= tol_rl(iters)
    np.testing.assert_allclose(host(ll), ll_ref, rtol=1e-6)


@pytest.mark.parametrize("shape,psf_shape", [((30, 34), (7, 5)), ((12, 16, 20), (3, 5, 7))])
def test_direct_fp32_and_fft_agree(shape, psf_shape):
    y, psf = _case(shape, [1.0] * len(shape), psf_shape)
    a = deconvolution.richardson_lucy(dev(y), dev(psf), 8, method=deconvolution.DIRECT, precision=precision.fp32)
    b = deconvolution.richardson_lucy(dev(y), dev(psf), 8, method=deconvolution.FFT)
    ref, _ = oracle.rl(y, psf.astype(np.float64), 8)
    assert rel(host(a), ref) <= tol_rl(8)
    assert rel(host(b), ref) <= tol_rl(8)


@pytest.mark.parametrize("prec,tol", [(precision.split_fp16, 2e-5), (precision.split_bf16, 2e-5)],
                         ids=["split_fp16", "split_bf16"])
def test_direct_split_precision_vs_fp64_oracle(prec, tol):
    y, psf = _case((18, 26, 70), (1.2, 1.0, 2.5), (9, 7, 21))
    iters = 5
    est = deconvolution.richardson_lucy(dev(y), dev(psf), iters, method=deconvolution.DIRECT, precision=prec)
    ref, _ = oracle.rl(y, psf.astype(np.float64), iters)
    err = rel(host(est), ref)
    print(f"RL direct, precision code {prec}: max error {err:.2e} of max|oracle64|")
    assert err <= tol


def _c3_like(shape, seed=3):
    """0.01 + blobs, blurred by the 31x31x63 anisotropic Gaussian PSF, scaled to
    peak 1000 photons, Poisson noise (the construction of the 256x256x128 workload)."""
    psf = synth.gaussian_kernel((3.0, 1.5, 1.5), (63, 31, 31)).astype(np.float32)
    truth = np.full(shape, 0.01)
    synth.add_blobs(truth, 40, (1.0, 4.0), (0.1, 1.0), seed, 1)
    blurred = oracle.conv_zero_fft(truth, psf.astype(np.float64))
    blurred *= 1000.0 / blurred.max()
    y = synth.poisson(np.maximum(blurred, 0), seed, 2).reshape(shape).astype(np.float32)
    return y, psf


def test_rl_200_iterations_real_psf():
    """200 iterations, FFT method and split-precision direct method, against the FP64 oracle."""
    y, psf = _c3_like((64, 64, 64))
    ref, ll_ref = oracle.rl(y, psf.astype(np.float64), 200)
    est, ll = deconvolution.richardson_lucy(dev(y), dev(psf), 200, method=deconvolution.FFT, return_ll=True)
    err = rel(host(est), ref)
    print(f"FFT RL, 200 iterations, 64^3, PSF 31x31x63: max {err:.2e} of max|oracle64|")
    assert err <= tol_rl(200)
    np.testing.assert_allclose(host(ll), ll_ref, rtol=1e-6)
    # The direct method sums 60 543 products per output in FP32 accumulators (the
    # split operands themselves hold 22 / 24 bits: one convolution is within 1e-6);
    # its per-iteration rounding is several times the FFT method's and grows the same
    # way with iterations (measured 1.6e-4 FP16, 2.1e-4 BF16): 8 x the FFT bound.
    for prec, bound in ((precision.split_fp16, 8 * tol_rl(200)), (precision.split_bf16, 8 * tol_rl(200))):
        est = deconvolution.richardson_lucy(dev(y), dev(psf), 200, method=deconvolution.DIRECT, precision=prec)
        err = rel(host(est), ref)
        print(f"direct RL, precision code {int(prec)}, 200 iterations: max {err:.2e} of max|oracle64|")
        assert err <= bound


def test_compile_time_fft_plans_keep_the_rl_bits():
    """The 256 x 256 x 128 RL domain (rows 140, y 280, z 160) runs compile-time FFT
    plans; with them switched off ("fft.fixed" = 0) the estimate and the
    log-likelihoods are bit-identical (a 256 x 256 x 122 volume: the same padded
    domain 280 x 280 x 160 with the 31 x 31 x 63 PSF)."""
    import torch
    from iane_amd import autotune
    y, psf = _c3_like((122, 256, 256))
    outs = []
    for fixed in (0, 1):
        with autotune.applied({"fft.fixed": fixed}):
            plan = deconvolution.rl_plan(y.shape, dev(psf), deconvolution.FFT)
            est = torch.empty_like(dev(y))
            ll = torch.empty(3, dtype=torch.float64, device="cuda")
            plan.run(dev(y), est, 3, ll_history=ll).wait()
            outs.append((host(est), host(ll)))
    np.testing.assert_array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_warm_start_continues_the_iteration():
    import torch
    y, psf = _case((24, 28), (1.3, 1.1), (7, 7))
    plan = deconvolution.rl_plan(y.shape, dev(psf), deconvolution.FFT)
    full = torch.empty_like(dev(y))
    plan.run(dev(y), full, 6).wait()
    part = torch.empty_like(full)
    plan.run(dev(y), part, 4).wait()
    plan.run(dev(y), part, 2, init=False).wait()
    np.testing.assert_array_equal(host(full), host(part))


def test_pointwise_steps_match_float32_semantics():
    import torch
    rng = np.random.default_rng(0)
    n = 10007
    y = rng.poisson(5.0, n).astype(np.float32)
    p = (rng.random(n) * 10).astype(np.float32)
    p[:50] = 0.0
    s = rng.random(n).astype(np.float32)
    s[:30] = 0.0
    s[30:60] = 1e-9
    c = rng.random(n).astype(np.float32)
    x = (rng.random(n) + 0.5).astype(np.float32)
    eps = np.float32(1e-6)
    dy, dp, ds, dc, dx = map(dev, (y, p, s, c, x))
    ratio = torch.empty_like(dp)
    ll = torch.zeros(1, dtype=torch.float64, device="cuda")
    deconvolution.rl_ratio(dy, dp, ratio, 1e-6, ll).wait()
    floor = np.where(p > eps, p, eps)
    np.testing.assert_array_equal(host(ratio), y / floor)
    # the log-likelihood is formed from FP32 terms summed as a compensated FP32 pair:
    # against the FP64 oracle within 1e-6 relative (the terms' own log rounding), and
    # against the exact sum of FP32 terms within a few ulp of log per term (the sum
    # itself adds no error a plain FP32 sum of 10^4 terms would)
    ll_ref = oracle.poisson_ll(y, p, 1e-6)
    got = host(ll)[0]
    assert abs(got - ll_ref) <= 1e-6 * abs(ll_ref)
    terms32 = (y * np.log(floor) - p).astype(np.float32)
    exact32 = float(np.sum(terms32.astype(np.float64)))
    ulp_budget = float(np.sum(np.abs(y.astype(np.float64) * np.log(floor.astype(np.float64))))) * 2.0 ** -22
    assert abs(got - exact32) <= ulp_budget
    deconvolution.rl_update(dx, dc, ds, 1e-6).wait()
    factor = np.where(s > 0, c / np.where(s > eps, s, eps), np.float32(0)).astype(np.float32)
    np.testing.assert_array_equal(host(dx), x * factor)
    init = torch.empty_like(dy)
    y2 = y.copy()
    y2[:10] = 0
    deconvolution.rl_initial(dev(y2), init, 1e-6).wait()
    np.testing.assert_array_equal(host(init), np.where(y2 > eps, y2, eps))


def test_errors():
    import iane_amd
    y, psf = _case((16, 16), (1.0, 1.0), (5, 5))
    with pytest.raises(iane_amd.invalid_argument, match="odd"):
        deconvolution.rl_plan(y.shape, dev(np.ones((4, 5), np.float32)))
    with pytest.raises(iane_amd.invalid_argument, match="epsilon"):
        deconvolution.rl_plan(y.shape, dev(psf), epsilon=0.0)
    plan = deconvolution.rl_plan(y.shape, dev(psf))
    with pytest.raises(iane_amd.invalid_argument, match="shape"):
        plan.run(dev(np.zeros((16, 17), np.float32)), dev(np.zeros((16, 17), np.float32)), 1)


# ------------------------------------------------------------ full size ----
def test_full_size_fft_rl_vs_oracle_and_em_property():
    """256x256x128 volume, PSF 31x31x63 (sigma 1.5, 1.5, 3.0), peak 1000 photons."""
    import torch
    truth, psf, y = synth.rl_volume_data()
    dy, dpsf = dev(y), dev(psf)
    plan = deconvolution.rl_plan(y.shape, dpsf, deconvolution.FFT)
    est = torch.empty_like(dy)
    ll = torch.zeros(20, dtype=torch.float64, device="cuda")
    plan.run(dy, est, 2, ll_history=ll).wait()
    ref, ll_ref = oracle.rl(y, psf.astype(np.float64), 2)
    assert rel(host(est), ref) <= tol_rl(2)
    np.testing.assert_allclose(host(ll)[:2], ll_ref, rtol=1e-6)
    plan.run(dy, est, 20, ll_history=ll).wait()
    history = host(ll)
    assert np.all(np.diff(history) >= -1e-7 * np.abs(history[:-1]))


# -------------------------------------------------------------- Wiener -----
def test_wiener_matches_oracle_and_delta_identity():
    rng = np.random.default_rng(2)
    img = (rng.random((48, 40)) * 200).astype(np.float32)
    psf = synth.gaussian_kernel((1.5, 1.2), (9, 7))
    padded = oracle.pad_psf_centered(psf, img.shape).astype(np.complex64)
    got = host(deconvolution.wiener(dev(img), dev(padded), 0.002))
    ref = oracle.wiener(img, padded.astype(np.complex128), 0.002)
    assert rel(got, ref) <= 1e-5
    delta = np.zeros((3, 3))
    delta[1, 1] = 1
    ident = host(deconvolution.wiener(dev(img), dev(oracle.pad_psf_centered(delta, img.shape).astype(np.complex64)),
                                      1e-9))
    np.testing.assert_allclose(ident, img, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("d", [(1, -1), (-1, 0), (0, 2)])
def test_wiener_off_centre_tap_shifts_back(d):
    """A unit tap at offset d (5x5 PSF): with noise -> 0 the output is the
    observation circularly shifted by -d (a missing conj or a reversed
    centring would shift by +d)."""
    rng = np.random.default_rng(3)
    img = (rng.random((24, 30)) * 100).astype(np.float32)
    psf = np.zeros((5, 5))
    psf[2 + d[0], 2 + d[1]] = 1.0
    padded = oracle.pad_psf_centered(psf, img.shape).astype(np.complex64)
    got = host(deconvolution.wiener(dev(img), dev(padded), 1e-9))
    np.testing.assert_allclose(got, np.roll(img, (-d[0], -d[1]), axis=(0, 1)), rtol=1e-5, atol=1e-3)


def test_wiener_single_frequency_closed_form():
    """A two-tap asymmetric PSF on one spatial frequency: amplitude |G| and phase
    arg G with G = conj(H) / (|H|^2 + noise) at that frequency (|H|^2 ~ 0.04, so
    the noise term changes the amplitude by ~5%)."""
    shape, k, phase = (12, 16), (5, 7), 0.3
    ii, jj = np.indices(shape)
    theta = 2 * np.pi * (k[0] * ii / shape[0] + k[1] * jj / shape[1])
    img = (100 * np.cos(theta + phase)).astype(np.float32)
    psf = np.zeros((3, 3))
    psf[1, 1], psf[1, 2] = 0.5, 0.45
    h = 0.5 + 0.45 * np.exp(-2j * np.pi * k[1] / shape[1])
    g = np.conj(h) / (abs(h) ** 2 + 0.002)
    expect = 100 * abs(g) * np.cos(theta + phase + np.angle(g))
    padded = oracle.pad_psf_centered(psf, shape).astype(np.complex64)
    got = host(deconvolution.wiener(dev(img), dev(padded), 0.002))
    assert rel(got, expect) <= 1e-5
