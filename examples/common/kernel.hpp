// Kernel and PSF text files of the example programs (format of the reference's
// examples/common/kernel.hpp:13-37 and deconvolution.hpp:10-30): the width and
// height, both positive and odd, then width x height finite coefficients in
// row-major order (x fastest) and nothing else. A PSF must be nonnegative with
// a positive sum; it is normalised in double precision and stored as float.
#pragma once

#include <iane/core.hpp>

#include <cmath>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>

namespace io {

inline iane::host_image<float, 2> read_kernel(const std::string& aPath)
{
	std::ifstream file(aPath);
	int width = 0, height = 0;
	if (!(file >> width >> height) || width <= 0 || height <= 0 || width % 2 == 0 || height % 2 == 0 ||
		width > std::numeric_limits<int>::max() / static_cast<int>(sizeof(float))) {
		throw std::invalid_argument("Expected positive odd kernel dimensions in " + aPath);
	}
	iane::host_image<float, 2> kernel(iane::math::vec2i{width, height});
	for (int i = 0; i < width * height; ++i) {
		float value;
		if (!(file >> value) || !std::isfinite(value)) {
			throw std::invalid_argument("Expected finite kernel coefficients in " + aPath);
		}
		kernel.pointer()[i] = value;
	}
	if (!(file >> std::ws).eof()) {
		throw std::invalid_argument("Unexpected data after kernel coefficients in " + aPath);
	}
	return kernel;
}

inline iane::host_image<float, 2> read_psf(const std::string& aPath)
{
	auto psf = read_kernel(aPath);
	const auto size = psf.dimensions();
	const auto count = static_cast<std::size_t>(size[0]) * size[1];
	double sum = 0.0;
	for (std::size_t i = 0; i < count; ++i) {
		const float coefficient = psf.pointer()[i];
		if (coefficient < 0.0f) {
			throw std::invalid_argument("PSF coefficients must be nonnegative");
		}
		sum += coefficient;
	}
	if (!(sum > 0.0)) {
		throw std::invalid_argument("PSF must have a positive coefficient sum");
	}
	for (std::size_t i = 0; i < count; ++i) {
		psf.pointer()[i] = static_cast<float>(psf.pointer()[i] / sum);
	}
	return psf;
}

/// Restored intensity in byte units -> grey pixel: 0 unless positive, 255 from
/// 255 up, else rounded half up (the reference's to_gray8, deconvolution.hpp:43-55).
inline std::uint8_t to_gray8(float aValue)
{
	if (!(aValue > 0.0f)) {
		return 0;
	}
	if (aValue >= 255.0f) {
		return 255;
	}
	const auto integer = static_cast<unsigned int>(aValue);
	return static_cast<std::uint8_t>(integer + (aValue - integer >= 0.5f));
}

} // namespace io
