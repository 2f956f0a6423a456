"""Pins the oracle's FFT, RL and Wiener restatements (CPU only).

FFT: brute-force FP64 DFT, the reference's FFT known answers, conventions
(negative exponent, unnormalised inverse, half spectrum on IANE axis 0).
RL: the reference's own RL loop (golden fixtures written through oracle/_ref),
the EM properties (non-decreasing log-likelihood, flux identity, fixed point).
Wiener: identity for a delta PSF as noise -> 0.
"""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

from oracle import oracle

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def test_fft_forward_matches_bruteforce():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 6)) + 1j * rng.standard_normal((5, 6))
    np.testing.assert_allclose(oracle.fft_forward(x), oracle.dft_bruteforce(x), atol=1e-10)


def test_fft_inverse_unnormalised_round_trip():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 4, 5)) + 0j
    np.testing.assert_allclose(oracle.fft_inverse(oracle.fft_forward(x)), x * x.size, atol=1e-9)


def test_fft_known_answer_delta():
    x = np.zeros((2, 4), complex)
    x[0, 1] = 1
    y = oracle.fft_forward(x)
    np.testing.assert_allclose(y[0].real, [1, 0, -1, 0], atol=1e-12)
    np.testing.assert_allclose(y[0].imag, [0, -1, 0, 1], atol=1e-12)


@pytest.mark.parametrize("shape", [(3, 4), (3, 5), (4, 6), (7,), (2, 3, 5)])
def test_rfft_half_spectrum_on_last_axis(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    spec = oracle.rfft(x)
    assert spec.shape == shape[:-1] + (shape[-1] // 2 + 1,)
    np.testing.assert_allclose(spec, oracle.dft_bruteforce(x)[..., : shape[-1] // 2 + 1], atol=1e-10)
    np.testing.assert_allclose(oracle.irfft(spec, shape), x * x.size, atol=1e-9)


def test_zero_border_fft_convolution_equals_direct():
    rng = np.random.default_rng(4)
    x = rng.random((9, 10, 11))
    k = rng.random((3, 5, 7))
    np.testing.assert_allclose(oracle.conv_zero_fft(x, k), oracle.convolve(x, k, np.float64, oracle.ZERO),
                               atol=1e-12)


# ---------------------------------------------------------------- RL -------
def _rl_case(shape=(12, 14), psf_shape=(5, 5), peak=200.0, seed=0):
    from iane_amd import synth
    truth = np.full(shape, 0.05)
    synth.add_blobs(truth, 3, (1.0, 2.0), (0.5, 1.0), seed, 1)
    psf = synth.gaussian_kernel([1.0] * len(shape), psf_shape)
    y = synth.poisson(np.maximum(oracle.conv_zero_fft(truth * peak, psf), 0), seed, 2).reshape(shape)
    return y, psf


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("rl_*.npz")), ids=lambda p: p.stem)
def test_rl_oracle_matches_reference_loop(path):
    """float64 restatement vs the reference loop in float32: agreement to float32 rounding."""
    data = np.load(path, allow_pickle=False)
    x, _ = oracle.rl(data["observed"], data["psf"].astype(np.float64), int(data["iterations"]),
                     float(data["epsilon"]), direct=True)
    ref = data["expected"].astype(np.float64)
    assert np.max(np.abs(x - ref)) <= 1e-4 * np.max(np.abs(ref))


def test_rl_fft_and_direct_agree():
    y, psf = _rl_case()
    a, lla = oracle.rl(y, psf, 5, direct=True)
    b, llb = oracle.rl(y, psf, 5, direct=False)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(lla, llb, rtol=1e-12)


def test_rl_log_likelihood_non_decreasing_and_flux():
    y, psf = _rl_case((16, 18), (5, 7))
    x, ll = oracle.rl(y, psf, 30)
    assert np.all(np.diff(ll) >= -1e-9 * np.abs(ll[:-1]))
    # flux identity of EM with exact sensitivity: sum(x_{k+1} * s) == sum(y) (no epsilon floor active)
    flipped = psf[::-1, ::-1]
    s = oracle.conv_zero_fft(np.ones_like(y), flipped)
    assert abs(np.sum(x * s) - np.sum(y)) <= 1e-8 * np.sum(y)


def test_rl_fixed_point():
    from iane_amd import synth
    psf = synth.gaussian_kernel((1.0, 1.5), (5, 7))
    truth = np.full((10, 12), 3.0) + np.arange(120).reshape(10, 12) / 50
    y = oracle.conv_zero_fft(truth, psf)
    x, _ = oracle.rl(y, psf, 3, init=truth)
    np.testing.assert_allclose(x, truth, rtol=1e-10)


def test_rl_dropping_a_term_fails():
    """A planted mistake (ratio without the epsilon floor replaced by y/p - 1) is caught."""
    y, psf = _rl_case()
    good, _ = oracle.rl(y, psf, 3)
    flipped = psf[::-1, ::-1]
    s = oracle.conv_zero_fft(np.ones_like(y), flipped)
    x = np.maximum(y, 1e-6)
    for _ in range(3):
        p = oracle.conv_zero_fft(x, psf)
        x = x * oracle.conv_zero_fft(y / p - 1, flipped) / s  # wrong multiplicative factor
    assert np.max(np.abs(x - good)) > 1e-3


def test_wiener_delta_psf_identity():
    rng = np.random.default_rng(5)
    img = rng.random((8, 10))
    psf = np.zeros((3, 3))
    psf[1, 1] = 1.0
    padded = oracle.pad_psf_centered(psf, img.shape)
    np.testing.assert_allclose(oracle.wiener(img, padded, 1e-12), img, atol=1e-9)
