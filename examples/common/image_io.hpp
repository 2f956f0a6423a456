// 8-bit image files of the example programs (the role of the reference's
// examples/common/io.{hpp,cpp}): binary PGM (P5, grey) and PPM (P6, RGB) with
// maxval <= 255, read and written. Other formats are refused with a message
// naming the two that are supported; convert with any image tool first.
#pragma once

#include <iane/core.hpp>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace io {

/// Packed RGB channels.
struct rgb8_pixel {
	std::uint8_t r;
	std::uint8_t g;
	std::uint8_t b;
};
static_assert(sizeof(rgb8_pixel) == 3);

using gray_image = iane::host_image<std::uint8_t, 2>;
using rgb_image = iane::host_image<rgb8_pixel, 2>;
using image8 = std::variant<gray_image, rgb_image>;

namespace detail {

using bytes = std::vector<std::uint8_t>;

inline bytes read_file(const std::string& aPath)
{
	std::ifstream file(aPath, std::ios::binary);
	if (!file) {
		throw std::runtime_error(aPath + ": unable to open image");
	}
	return bytes{std::istreambuf_iterator<char>(file), {}};
}

inline void write_file(const std::string& aPath, const bytes& aData)
{
	std::ofstream file(aPath, std::ios::binary);
	if (!file || !file.write(reinterpret_cast<const char*>(aData.data()), static_cast<std::streamsize>(aData.size()))) {
		throw std::runtime_error(aPath + ": unable to write image");
	}
}

// Interleaved 8-bit pixels with `channels` channels -> gray8 or RGB8 (alpha dropped).
inline image8 make_image(const bytes& aPixels, int aWidth, int aHeight, int aChannels)
{
	const std::size_t count = static_cast<std::size_t>(aWidth) * aHeight;
	if (aChannels <= 2) {
		gray_image out(iane::math::vec2i{aWidth, aHeight});
		for (std::size_t i = 0; i < count; ++i) {
			out.pointer()[i] = aPixels[i * aChannels];
		}
		return out;
	}
	rgb_image out(iane::math::vec2i{aWidth, aHeight});
	for (std::size_t i = 0; i < count; ++i) {
		out.pointer()[i] = {aPixels[i * aChannels], aPixels[i * aChannels + 1], aPixels[i * aChannels + 2]};
	}
	return out;
}

inline void check_size(long long aWidth, long long aHeight, const std::string& aPath)
{
	if (aWidth <= 0 || aHeight <= 0 || aWidth > (1 << 24) || aHeight > (1 << 24) || aWidth * aHeight > (1LL << 30)) {
		throw std::runtime_error(aPath + ": unsupported image dimensions");
	}
}

// ---- PGM / PPM (binary) ----
inline image8 decode_pnm(const bytes& aData, const std::string& aPath)
{
	const auto bad = [&](const char* aWhat) { return std::runtime_error(aPath + ": " + aWhat); };
	const int channels = aData[1] == '5' ? 1 : 3;
	std::size_t pos = 2;
	long values[3] = {0, 0, 0};
	for (long& v : values) {
		while (pos < aData.size() && (std::isspace(aData[pos]) || aData[pos] == '#')) {
			if (aData[pos] == '#') {
				while (pos < aData.size() && aData[pos] != '\n') {
					++pos;
				}
			} else {
				++pos;
			}
		}
		if (pos >= aData.size() || !std::isdigit(aData[pos])) {
			throw bad("invalid PNM header");
		}
		while (pos < aData.size() && std::isdigit(aData[pos])) {
			v = v * 10 + (aData[pos++] - '0');
			if (v > (1L << 30)) {
				throw bad("invalid PNM header");
			}
		}
	}
	++pos; // one whitespace byte before the samples
	check_size(values[0], values[1], aPath);
	if (values[2] <= 0 || values[2] > 255) {
		throw bad("PNM maxval must be 1..255");
	}
	const std::size_t size = static_cast<std::size_t>(values[0]) * values[1] * channels;
	if (pos + size > aData.size()) {
		throw bad("truncated PNM pixel data");
	}
	bytes pixels(aData.begin() + static_cast<std::ptrdiff_t>(pos), aData.begin() + static_cast<std::ptrdiff_t>(pos + size));
	if (values[2] != 255) {
		for (auto& p : pixels) {
			p = static_cast<std::uint8_t>((p * 255 + values[2] / 2) / values[2]);
		}
	}
	return make_image(pixels, static_cast<int>(values[0]), static_cast<int>(values[1]), channels);
}

inline bytes encode_pnm(const std::uint8_t* aPixels, int aWidth, int aHeight, int aChannels)
{
	const std::string header = std::string(aChannels == 1 ? "P5\n" : "P6\n") + std::to_string(aWidth) + " " +
		std::to_string(aHeight) + "\n255\n";
	bytes out(header.begin(), header.end());
	out.insert(out.end(), aPixels, aPixels + static_cast<std::size_t>(aWidth) * aHeight * aChannels);
	return out;
}

inline std::string lower_extension(const std::string& aPath)
{
	auto extension = std::filesystem::path(aPath).extension().string();
	std::transform(extension.begin(), extension.end(), extension.begin(),
		[](unsigned char c) { return static_cast<char>(std::tolower(c)); });
	return extension;
}

inline void write_image(const std::string& aPath, const std::uint8_t* aPixels, int aWidth, int aHeight, int aChannels)
{
	const auto extension = lower_extension(aPath);
	if (extension != ".pgm" && extension != ".ppm") {
		throw std::invalid_argument(aPath + ": output must be a binary .pgm (grey) or .ppm (colour) file");
	}
	if ((extension == ".pgm") != (aChannels == 1)) {
		throw std::invalid_argument(aPath + ": .pgm holds grey images and .ppm colour images");
	}
	write_file(aPath, encode_pnm(aPixels, aWidth, aHeight, aChannels));
}

} // namespace detail

/// Decode a binary PGM (grey) or PPM (RGB) file.
inline image8 read_image8(const std::string& aPath)
{
	const auto data = detail::read_file(aPath);
	if (data.size() >= 2 && data[0] == 'P' && (data[1] == '5' || data[1] == '6')) {
		return detail::decode_pnm(data, aPath);
	}
	throw std::runtime_error(aPath + ": unsupported image format (binary PGM or PPM expected)");
}

/// Decode a grey image.
inline gray_image read_gray8(const std::string& aPath)
{
	auto image = read_image8(aPath);
	if (auto* gray = std::get_if<gray_image>(&image)) {
		return std::move(*gray);
	}
	throw std::runtime_error(aPath + ": expected a grayscale image");
}

inline void write_gray8(const std::string& aPath, const gray_image& aImage)
{
	const auto size = aImage.dimensions();
	detail::write_image(aPath, aImage.pointer(), size[0], size[1], 1);
}

inline void write_rgb8(const std::string& aPath, const rgb_image& aImage)
{
	const auto size = aImage.dimensions();
	detail::write_image(aPath, reinterpret_cast<const std::uint8_t*>(aImage.pointer()), size[0], size[1], 3);
}

} // namespace io
