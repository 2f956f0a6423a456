"""Pins the oracle's FFT, RL and Wiener restatements (CPU only).

FFT: brute-force FP64 DFT, the reference's FFT known answers, conventions
(negative exponent, unnormalised inverse, half spectrum on IANE axis 0).
RL: the reference's own RL loop (golden fixtures written through oracle/_ref),
the EM properties (non-decreasing log-likelihood, flux identity, fixed point).
Wiener: closed forms (an off-centre unit tap is a circular shift, a two-tap
PSF on one spatial frequency has a closed-form response) with planted mistakes.
Log-likelihood: scipy's Poisson log-pmf plus log(y!), and the epsilon floor.
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


# ------------------------------------------------------------- Wiener -----
# The oracle's Wiener restatement (examples/wiener/main.cu:21-32,44-83 with the
# PSF placement of :109-120) is pinned by closed forms that a plausible mistake
# breaks: an off-centre unit tap is a pure circular shift (a missing conj(H) or
# a reversed centring shifts the other way), and a two-tap PSF on a single
# spatial frequency has a closed-form response in which the sign of the noise
# term changes the amplitude.

def _wiener_no_conj(observed, padded, noise):
    y = np.fft.fftn(observed)
    h = np.fft.fftn(padded)
    return np.real(np.fft.ifftn(y * h / (np.abs(h) ** 2 + noise)))


def _wiener_noise_sign(observed, padded, noise):
    y = np.fft.fftn(observed)
    h = np.fft.fftn(padded)
    return np.real(np.fft.ifftn(y * np.conj(h) / (np.abs(h) ** 2 - noise)))


def _pad_reversed(psf, shape):
    out = np.zeros(shape)
    idx = np.indices(psf.shape).reshape(psf.ndim, -1)
    out[tuple((psf.shape[a] // 2 - idx[a]) % shape[a] for a in range(psf.ndim))] = psf.reshape(-1)
    return out


OFFSETS = [(1, -1), (-1, 0), (0, 1), (2, -1)]


def _unit_tap(d, psf_shape=(5, 3)):
    psf = np.zeros(psf_shape)
    psf[psf_shape[0] // 2 + d[0], psf_shape[1] // 2 + d[1]] = 1.0
    return psf


@pytest.mark.parametrize("d", OFFSETS)
def test_pad_psf_places_the_centre_at_the_origin(d):
    """wiener/main.cu:109-120: kernel tap x goes to x - extent/2 (mod size) per axis."""
    shape = (7, 9)
    padded = oracle.pad_psf_centered(_unit_tap(d), shape)
    expect = np.zeros(shape)
    expect[d[0] % shape[0], d[1] % shape[1]] = 1.0
    np.testing.assert_array_equal(padded, expect)
    assert not np.array_equal(_pad_reversed(_unit_tap(d), shape), expect)


@pytest.mark.parametrize("d", OFFSETS)
def test_wiener_undoes_the_shift_of_an_off_centre_tap(d):
    """A unit tap at offset d blurs by a shift: convolve gives dst[p] = src[p - d]
    (the reference's convention, checked on the interior with the C oracle); with
    noise -> 0 Wiener returns the observation shifted back by -d."""
    rng = np.random.default_rng(11)
    img = rng.random((10, 12))
    psf = _unit_tap(d)
    blurred_interior = oracle.convolve(img, psf, np.float64, oracle.ZERO)[3:-3, 3:-3]
    np.testing.assert_array_equal(blurred_interior, np.roll(img, d, axis=(0, 1))[3:-3, 3:-3])
    padded = oracle.pad_psf_centered(psf, img.shape)
    expect = np.roll(img, (-d[0], -d[1]), axis=(0, 1))
    np.testing.assert_allclose(oracle.wiener(img, padded, 1e-12), expect, atol=1e-9)
    np.testing.assert_allclose(oracle.wiener(np.roll(img, d, axis=(0, 1)), padded, 1e-12), img, atol=1e-9)
    # planted mistakes shift the other way
    assert np.max(np.abs(_wiener_no_conj(img, padded, 1e-12) - expect)) > 0.1
    assert np.max(np.abs(oracle.wiener(img, _pad_reversed(psf, img.shape), 1e-12) - expect)) > 0.1


def _single_frequency_case():
    # image: one real spatial frequency; PSF: two asymmetric taps (centre a, offset d b)
    shape = (12, 16)
    k = (5, 7)
    phase = 0.3
    ii, jj = np.indices(shape)
    theta = 2 * np.pi * (k[0] * ii / shape[0] + k[1] * jj / shape[1])
    img = np.cos(theta + phase)
    a, b, d = 0.5, 0.45, (0, 1)
    psf = np.zeros((3, 3))
    psf[1, 1] = a
    psf[1 + d[0], 1 + d[1]] = b
    # closed form: H(k) = a + b exp(-i w.d), G = conj(H) / (|H|^2 + noise); a real
    # cosine comes back with amplitude |G| and phase shifted by arg G
    w_d = 2 * np.pi * (k[0] * d[0] / shape[0] + k[1] * d[1] / shape[1])
    h = a + b * np.exp(-1j * w_d)
    return img, psf, theta, phase, h


def test_wiener_single_frequency_closed_form():
    img, psf, theta, phase, h = _single_frequency_case()
    noise = 0.002
    g = np.conj(h) / (abs(h) ** 2 + noise)
    expect = abs(g) * np.cos(theta + phase + np.angle(g))
    padded = oracle.pad_psf_centered(psf, img.shape)
    np.testing.assert_allclose(oracle.wiener(img, padded, noise), expect, atol=1e-9)
    # the noise term matters here (|H|^2 ~ 0.04): a wrong sign or a missing conj is caught
    assert abs(h) ** 2 < 0.05
    assert np.max(np.abs(_wiener_noise_sign(img, padded, noise) - expect)) > 1e-2 * np.max(np.abs(expect))
    assert np.max(np.abs(_wiener_no_conj(img, padded, noise) - expect)) > 1e-2 * np.max(np.abs(expect))


def test_wiener_delta_psf_identity():
    rng = np.random.default_rng(5)
    img = rng.random((8, 10))
    psf = np.zeros((3, 3))
    psf[1, 1] = 1.0
    padded = oracle.pad_psf_centered(psf, img.shape)
    np.testing.assert_allclose(oracle.wiener(img, padded, 1e-12), img, atol=1e-9)


# ---------------------------------------------------- log-likelihood -------
def test_poisson_ll_matches_scipy():
    """sum(y log p - p) equals the Poisson log-pmf summed by scipy plus sum(log y!)."""
    from scipy import special, stats
    rng = np.random.default_rng(7)
    p = rng.random(5000) * 50 + 1e-3
    y = rng.poisson(p).astype(np.float64)
    expect = float(np.sum(stats.poisson.logpmf(y, p)) + np.sum(special.gammaln(y + 1)))
    assert abs(oracle.poisson_ll(y, p) - expect) <= 1e-12 * abs(expect)
    # a likelihood without the -sum(p) term is caught
    dropped = float(np.sum(y * np.log(p)))
    assert abs(dropped - expect) > 1e-3 * abs(expect)


def test_poisson_ll_epsilon_floor():
    """Predictions below epsilon enter the log as epsilon and the linear term as themselves."""
    eps = 1e-6
    y = np.array([0.0, 3.0, 2.0, 0.0])
    p = np.array([0.0, 0.0, 2.5e-7, 4.0])
    expect = 0.0 + 3 * np.log(eps) + (2 * np.log(eps) - 2.5e-7) + (0 - 4.0)
    assert abs(oracle.poisson_ll(y, p, eps) - expect) <= 1e-12 * abs(expect)
