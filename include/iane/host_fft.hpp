// Host-view FFT behind iane::fft_forward / fft_inverse / rfft / irfft: the
// role of the reference's FFTW plans (include/iane/algorithms/fft/fft.hpp:80-99,
// 196-223; rfft.hpp:107-140) with the same conventions as the device path:
// negative exponent forward, unnormalised inverse, the packed half spectrum
// {n0/2 + 1, n1, ...} for real transforms, and C2R leaving its input untouched.
//
// Lines are transformed in double precision and rounded to float once. A line
// length is split into radix-4, radix-2 and prime radices up to 64 and run as
// a recursive mixed-radix decimation in time (twiddles from one table of N-th
// roots of unity); a length with a larger prime factor runs as Bluestein's
// chirp-z convolution on a power-of-two transform. Multi-dimensional
// transforms apply the line transform along each axis in turn.
#pragma once

#include <complex>
#include <cmath>
#include <cstddef>
#include <memory>
#include <numbers>
#include <stdexcept>
#include <vector>

namespace iane::detail {

using host_complex = std::complex<double>;

// Radices of aLength (fours first, then primes in increasing order); empty when
// a prime factor exceeds aLimit.
inline std::vector<int> host_fft_radices(int aLength, int aLimit)
{
	std::vector<int> radices;
	int rest = aLength;
	while (rest % 4 == 0) {
		radices.push_back(4);
		rest /= 4;
	}
	for (int p = 2; p * p <= rest; ++p) {
		while (rest % p == 0) {
			radices.push_back(p);
			rest /= p;
		}
	}
	if (rest > 1) {
		radices.push_back(rest);
	}
	for (int r : radices) {
		if (r > aLimit) {
			return {};
		}
	}
	return radices;
}

// One line length and direction: out[k] = sum_t in[t * stride] w^(t k),
// w = exp(-+2 pi i / n).
class host_fft_line {
public:
	static constexpr int kMaxRadix = 64;

	host_fft_line(int aLength, bool aInverse) : mLength(aLength), mInverse(aInverse)
	{
		if (aLength <= 0) {
			throw std::invalid_argument("FFT requires positive dimensions");
		}
		mRadices = host_fft_radices(aLength, kMaxRadix);
		if (aLength > 1 && mRadices.empty()) {
			setup_chirp();
		} else {
			mRoots = roots(aLength, aInverse);
		}
	}

	int length() const noexcept { return mLength; }

	void transform(const host_complex* aIn, std::ptrdiff_t aStride, host_complex* aOut) const
	{
		if (mChirp) {
			transform_chirp(aIn, aStride, aOut);
		} else {
			step(aIn, aStride, aOut, mLength, 0);
		}
	}

private:
	static std::vector<host_complex> roots(int aLength, bool aInverse)
	{
		std::vector<host_complex> table(static_cast<std::size_t>(aLength));
		const double sign = aInverse ? 1.0 : -1.0;
		for (int k = 0; k < aLength; ++k) {
			table[k] = std::polar(1.0, sign * 2.0 * std::numbers::pi * k / aLength);
		}
		return table;
	}

	// Decimation in time: with r the radix of this level and m = n / r, the
	// r sub-transforms of the inputs q, q + r, q + 2r, ... land at out[q m ..],
	// and X[k + m s] = sum_q (w_n^(q k) Y_q[k]) w_r^(q s).
	void step(const host_complex* aIn, std::ptrdiff_t aStride, host_complex* aOut, int aN, int aLevel) const
	{
		if (aN == 1) {
			aOut[0] = aIn[0];
			return;
		}
		const int r = mRadices[static_cast<std::size_t>(aLevel)];
		const int m = aN / r;
		for (int q = 0; q < r; ++q) {
			step(aIn + q * aStride, aStride * r, aOut + q * m, m, aLevel + 1);
		}
		const int twiddle_step = mLength / aN;  // w_n = w_N^(N / n)
		const int radix_step = mLength / r;     // w_r = w_N^(N / r)
		host_complex t[kMaxRadix];
		for (int k = 0; k < m; ++k) {
			for (int q = 0; q < r; ++q) {
				const host_complex y = aOut[q * m + k];
				t[q] = q == 0 ? y : y * mRoots[static_cast<std::size_t>(q * k * twiddle_step)];
			}
			if (r == 2) {
				aOut[k] = t[0] + t[1];
				aOut[k + m] = t[0] - t[1];
			} else if (r == 4) {
				const host_complex j = mRoots[static_cast<std::size_t>(radix_step)];  // w_4 = -+i
				const host_complex a = t[0] + t[2];
				const host_complex b = t[0] - t[2];
				const host_complex c = t[1] + t[3];
				const host_complex d = (t[1] - t[3]) * j;
				aOut[k] = a + c;
				aOut[k + m] = b + d;
				aOut[k + 2 * m] = a - c;
				aOut[k + 3 * m] = b - d;
			} else {
				for (int s = 0; s < r; ++s) {
					host_complex sum = t[0];
					for (int q = 1; q < r; ++q) {
						sum += t[q] * mRoots[static_cast<std::size_t>((q * s) % r * radix_step)];
					}
					aOut[k + s * m] = sum;
				}
			}
		}
	}

	// Bluestein: w^(t k) = c[t] c[k] conj(c[k - t]) with c[t] = exp(-+ i pi t^2 / n),
	// so X = c . (a (*) b) with a[t] = x[t] c[t] and b[j] = conj(c[j]), a cyclic
	// convolution of length M >= 2n - 1 done with power-of-two transforms.
	void setup_chirp()
	{
		mChirp = true;
		int m = 1;
		while (m < 2 * mLength - 1) {
			m *= 2;
		}
		mChirpLength = m;
		mChirpForward = std::make_unique<host_fft_line>(m, false);
		const double sign = mInverse ? 1.0 : -1.0;
		mChirpC.resize(static_cast<std::size_t>(mLength));
		for (int t = 0; t < mLength; ++t) {
			const long long t2 = static_cast<long long>(t) * t % (2LL * mLength);
			mChirpC[t] = std::polar(1.0, sign * std::numbers::pi * static_cast<double>(t2) / mLength);
		}
		std::vector<host_complex> b(static_cast<std::size_t>(m));
		b[0] = std::conj(mChirpC[0]);
		for (int j = 1; j < mLength; ++j) {
			b[j] = std::conj(mChirpC[j]);
			b[m - j] = std::conj(mChirpC[j]);
		}
		mChirpB.resize(static_cast<std::size_t>(m));
		mChirpForward->transform(b.data(), 1, mChirpB.data());
	}

	void transform_chirp(const host_complex* aIn, std::ptrdiff_t aStride, host_complex* aOut) const
	{
		const int m = mChirpLength;
		std::vector<host_complex> a(static_cast<std::size_t>(m));
		std::vector<host_complex> spectrum(static_cast<std::size_t>(m));
		for (int t = 0; t < mLength; ++t) {
			a[t] = aIn[t * aStride] * mChirpC[t];
		}
		mChirpForward->transform(a.data(), 1, spectrum.data());
		// inverse through the forward transform: ifft(P) = conj(fft(conj(P)))
		for (int i = 0; i < m; ++i) {
			a[i] = std::conj(spectrum[i] * mChirpB[i]);
		}
		mChirpForward->transform(a.data(), 1, spectrum.data());
		const double scale = 1.0 / m;
		for (int k = 0; k < mLength; ++k) {
			aOut[k] = mChirpC[k] * std::conj(spectrum[k]) * scale;
		}
	}

	int mLength;
	bool mInverse;
	std::vector<int> mRadices;
	std::vector<host_complex> mRoots;
	bool mChirp = false;
	int mChirpLength = 0;
	std::unique_ptr<host_fft_line> mChirpForward;
	std::vector<host_complex> mChirpC;
	std::vector<host_complex> mChirpB;
};

// Transform every line of a packed double array along aAxis (extents aExtent,
// aRank axes, axis 0 fastest) in place.
inline void host_fft_axis(std::vector<host_complex>& aData, const int* aExtent, int aRank, int aAxis,
	const host_fft_line& aLine)
{
	std::ptrdiff_t stride = 1;
	for (int i = 0; i < aAxis; ++i) {
		stride *= aExtent[i];
	}
	const std::ptrdiff_t n = aExtent[aAxis];
	std::ptrdiff_t outer = 1;
	for (int i = aAxis + 1; i < aRank; ++i) {
		outer *= aExtent[i];
	}
	std::vector<host_complex> line(static_cast<std::size_t>(n));
	for (std::ptrdiff_t o = 0; o < outer; ++o) {
		for (std::ptrdiff_t inner = 0; inner < stride; ++inner) {
			host_complex* base = aData.data() + o * stride * n + inner;
			aLine.transform(base, stride, line.data());
			for (std::ptrdiff_t k = 0; k < n; ++k) {
				base[k * stride] = line[static_cast<std::size_t>(k)];
			}
		}
	}
}

} // namespace iane::detail
