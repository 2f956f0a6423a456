// 8-bit image files of the example programs (the role of the reference's
// examples/common/io.{hpp,cpp}, written here without stb).
//
// read_image8 detects the format from the file contents and decodes to gray8
// (one or two channels) or RGB8 (three or four channels), discarding alpha:
//   PNG  every bit depth and colour type (palette, 16-bit keeps the high byte,
//        low bit depths scale to 0..255), non-interlaced; zlib inflates IDAT
//   BMP  uncompressed 8-bit palette, 24-bit, 32-bit
//   TGA  true-colour and grey, raw or run-length encoded, 8/24/32 bits
//   PGM/PPM binary (P5/P6) with maxval <= 255
// The output extension selects the encoder: .png (zlib, filter 0), .bmp
// (24-bit, grey stored as equal RGB channels), .tga (raw), .pgm/.ppm. JPEG is
// not available (no JPEG codec in this build): both directions raise.
#pragma once

#include <iane/core.hpp>

#include <zlib.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cstdint>
#include <cstring>
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

inline std::uint32_t be32(const std::uint8_t* p) { return (std::uint32_t(p[0]) << 24) | (p[1] << 16) | (p[2] << 8) | p[3]; }
inline std::uint32_t le32(const std::uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | (std::uint32_t(p[3]) << 24); }
inline std::uint16_t le16(const std::uint8_t* p) { return static_cast<std::uint16_t>(p[0] | (p[1] << 8)); }

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

// ---- PNG (ISO/IEC 15948) ----
inline std::uint8_t paeth(int a, int b, int c)
{
	const int p = a + b - c;
	const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
	return static_cast<std::uint8_t>(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}

inline image8 decode_png(const bytes& aData, const std::string& aPath)
{
	const auto bad = [&](const char* aWhat) { return std::runtime_error(aPath + ": " + aWhat); };
	std::size_t pos = 8;
	int width = 0, height = 0, depth = 0, color = -1, interlace = 0;
	bytes compressed, palette;
	bool ended = false;
	while (pos + 12 <= aData.size() && !ended) {
		const std::uint32_t length = be32(&aData[pos]);
		if (length > aData.size() - pos - 12) {
			throw bad("truncated PNG chunk");
		}
		const std::string type(reinterpret_cast<const char*>(&aData[pos + 4]), 4);
		const std::uint8_t* body = &aData[pos + 8];
		if (type == "IHDR") {
			if (length < 13) {
				throw bad("invalid PNG header");
			}
			width = static_cast<int>(be32(body));
			height = static_cast<int>(be32(body + 4));
			depth = body[8];
			color = body[9];
			interlace = body[12];
		} else if (type == "PLTE") {
			palette.assign(body, body + length);
		} else if (type == "IDAT") {
			compressed.insert(compressed.end(), body, body + length);
		} else if (type == "IEND") {
			ended = true;
		}
		pos += 12 + length;
	}
	check_size(width, height, aPath);
	static const int channels_of[] = {1, 0, 3, 1, 2, 0, 4};
	if (color < 0 || color > 6 || channels_of[color] == 0) {
		throw bad("unsupported PNG colour type");
	}
	const int channels = channels_of[color];
	const bool depth_ok = (depth == 8 || depth == 16) ? color != 3 || depth == 8 :
		((depth == 1 || depth == 2 || depth == 4) && (color == 0 || color == 3));
	if (!depth_ok) {
		throw bad("unsupported PNG bit depth");
	}
	if (interlace != 0) {
		throw bad("interlaced PNG is not supported");
	}
	const std::size_t bits_per_pixel = static_cast<std::size_t>(channels) * depth;
	const std::size_t row_bytes = (static_cast<std::size_t>(width) * bits_per_pixel + 7) / 8;
	const std::size_t bpp = std::max<std::size_t>(1, bits_per_pixel / 8);
	bytes raw((row_bytes + 1) * height);
	z_stream z{};
	if (inflateInit(&z) != Z_OK) {
		throw bad("zlib initialisation failed");
	}
	z.next_in = compressed.data();
	z.avail_in = static_cast<uInt>(compressed.size());
	z.next_out = raw.data();
	z.avail_out = static_cast<uInt>(raw.size());
	const int status = inflate(&z, Z_FINISH);
	const bool complete = z.avail_out == 0;
	inflateEnd(&z);
	if (!complete || (status != Z_STREAM_END && status != Z_OK && status != Z_BUF_ERROR)) {
		throw bad("corrupt PNG image data");
	}
	bytes rows(row_bytes * height);
	for (int y = 0; y < height; ++y) {
		const std::uint8_t filter = raw[y * (row_bytes + 1)];
		const std::uint8_t* in = &raw[y * (row_bytes + 1) + 1];
		std::uint8_t* out = &rows[y * row_bytes];
		const std::uint8_t* up = y > 0 ? &rows[(y - 1) * row_bytes] : nullptr;
		for (std::size_t i = 0; i < row_bytes; ++i) {
			const int a = i >= bpp ? out[i - bpp] : 0;
			const int b = up ? up[i] : 0;
			const int c = (up && i >= bpp) ? up[i - bpp] : 0;
			switch (filter) {
			case 0: out[i] = in[i]; break;
			case 1: out[i] = static_cast<std::uint8_t>(in[i] + a); break;
			case 2: out[i] = static_cast<std::uint8_t>(in[i] + b); break;
			case 3: out[i] = static_cast<std::uint8_t>(in[i] + ((a + b) >> 1)); break;
			case 4: out[i] = static_cast<std::uint8_t>(in[i] + paeth(a, b, c)); break;
			default: throw bad("invalid PNG filter");
			}
		}
	}
	// unpack to 8-bit interleaved samples
	const int out_channels = color == 3 ? 3 : channels;
	bytes pixels(static_cast<std::size_t>(width) * height * out_channels);
	for (int y = 0; y < height; ++y) {
		const std::uint8_t* row = &rows[y * row_bytes];
		for (int x = 0; x < width; ++x) {
			for (int ch = 0; ch < channels; ++ch) {
				const std::size_t sample = static_cast<std::size_t>(x) * channels + ch;
				unsigned v;
				if (depth == 16) {
					v = row[sample * 2]; // high byte
				} else if (depth == 8) {
					v = row[sample];
				} else {
					const std::size_t bit = sample * depth;
					v = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
					if (color == 0) {
						v = v * (255u / ((1u << depth) - 1)); // 1: x255, 2: x85, 4: x17
					}
				}
				std::uint8_t* dst = &pixels[(static_cast<std::size_t>(y) * width + x) * out_channels];
				if (color == 3) {
					if (3 * v + 2 >= palette.size()) {
						throw bad("PNG palette index out of range");
					}
					dst[0] = palette[3 * v];
					dst[1] = palette[3 * v + 1];
					dst[2] = palette[3 * v + 2];
				} else {
					dst[ch] = static_cast<std::uint8_t>(v);
				}
			}
		}
	}
	return make_image(pixels, width, height, out_channels);
}

inline void put_be32(bytes& aOut, std::uint32_t v)
{
	for (int s = 24; s >= 0; s -= 8) {
		aOut.push_back(static_cast<std::uint8_t>(v >> s));
	}
}

inline void png_chunk(bytes& aOut, const char* aType, const bytes& aBody)
{
	put_be32(aOut, static_cast<std::uint32_t>(aBody.size()));
	const std::size_t start = aOut.size();
	aOut.insert(aOut.end(), aType, aType + 4);
	aOut.insert(aOut.end(), aBody.begin(), aBody.end());
	const uLong crc = crc32(0L, &aOut[start], static_cast<uInt>(aOut.size() - start));
	put_be32(aOut, static_cast<std::uint32_t>(crc));
}

inline bytes encode_png(const std::uint8_t* aPixels, int aWidth, int aHeight, int aChannels)
{
	const std::size_t row_bytes = static_cast<std::size_t>(aWidth) * aChannels;
	bytes raw;
	raw.reserve((row_bytes + 1) * aHeight);
	for (int y = 0; y < aHeight; ++y) {
		raw.push_back(0);
		raw.insert(raw.end(), aPixels + y * row_bytes, aPixels + (y + 1) * row_bytes);
	}
	uLongf size = compressBound(static_cast<uLong>(raw.size()));
	bytes compressed(size);
	if (compress2(compressed.data(), &size, raw.data(), static_cast<uLong>(raw.size()), 6) != Z_OK) {
		throw std::runtime_error("PNG compression failed");
	}
	compressed.resize(size);
	bytes out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
	bytes header;
	put_be32(header, static_cast<std::uint32_t>(aWidth));
	put_be32(header, static_cast<std::uint32_t>(aHeight));
	header.push_back(8);
	header.push_back(aChannels == 1 ? 0 : 2);
	header.push_back(0);
	header.push_back(0);
	header.push_back(0);
	png_chunk(out, "IHDR", header);
	png_chunk(out, "IDAT", compressed);
	png_chunk(out, "IEND", {});
	return out;
}

// ---- BMP ----
inline image8 decode_bmp(const bytes& aData, const std::string& aPath)
{
	const auto bad = [&](const char* aWhat) { return std::runtime_error(aPath + ": " + aWhat); };
	if (aData.size() < 54) {
		throw bad("truncated BMP");
	}
	const std::uint32_t offset = le32(&aData[10]);
	const std::uint32_t dib = le32(&aData[14]);
	const int width = static_cast<int>(le32(&aData[18]));
	const int signed_height = static_cast<int>(le32(&aData[22]));
	const int bits = le16(&aData[28]);
	const std::uint32_t compression = le32(&aData[30]);
	const int height = signed_height < 0 ? -signed_height : signed_height;
	check_size(width, height, aPath);
	if (!(compression == 0 || (compression == 3 && bits == 32)) || !(bits == 8 || bits == 24 || bits == 32)) {
		throw bad("unsupported BMP format (uncompressed 8, 24 or 32 bits only)");
	}
	const std::size_t stride = (static_cast<std::size_t>(width) * bits / 8 + 3) & ~std::size_t(3);
	if (offset + stride * height > aData.size()) {
		throw bad("truncated BMP pixel data");
	}
	bytes pixels(static_cast<std::size_t>(width) * height * 3);
	for (int y = 0; y < height; ++y) {
		const int src_row = signed_height < 0 ? y : height - 1 - y;
		const std::uint8_t* row = &aData[offset + src_row * stride];
		for (int x = 0; x < width; ++x) {
			std::uint8_t* dst = &pixels[(static_cast<std::size_t>(y) * width + x) * 3];
			if (bits == 8) {
				const std::size_t entry = 14 + dib + 4 * static_cast<std::size_t>(row[x]);
				if (entry + 3 > aData.size()) {
					throw bad("BMP palette index out of range");
				}
				dst[0] = aData[entry + 2];
				dst[1] = aData[entry + 1];
				dst[2] = aData[entry];
			} else {
				const std::uint8_t* p = row + static_cast<std::size_t>(x) * (bits / 8);
				dst[0] = p[2];
				dst[1] = p[1];
				dst[2] = p[0];
			}
		}
	}
	return make_image(pixels, width, height, 3);
}

inline bytes encode_bmp(const std::uint8_t* aPixels, int aWidth, int aHeight, int aChannels)
{
	const std::size_t stride = (static_cast<std::size_t>(aWidth) * 3 + 3) & ~std::size_t(3);
	const std::uint32_t image_bytes = static_cast<std::uint32_t>(stride * aHeight);
	bytes out(54 + image_bytes, 0);
	const auto put32 = [&](std::size_t at, std::uint32_t v) {
		for (int i = 0; i < 4; ++i) {
			out[at + i] = static_cast<std::uint8_t>(v >> (8 * i));
		}
	};
	out[0] = 'B';
	out[1] = 'M';
	put32(2, static_cast<std::uint32_t>(out.size()));
	put32(10, 54);
	put32(14, 40);
	put32(18, static_cast<std::uint32_t>(aWidth));
	put32(22, static_cast<std::uint32_t>(aHeight));
	out[26] = 1;
	out[28] = 24;
	put32(34, image_bytes);
	for (int y = 0; y < aHeight; ++y) {
		std::uint8_t* row = &out[54 + (aHeight - 1 - y) * stride];
		for (int x = 0; x < aWidth; ++x) {
			const std::uint8_t* p = aPixels + (static_cast<std::size_t>(y) * aWidth + x) * aChannels;
			row[3 * x] = aChannels == 1 ? p[0] : p[2];
			row[3 * x + 1] = aChannels == 1 ? p[0] : p[1];
			row[3 * x + 2] = p[0];
		}
	}
	return out;
}

// ---- TGA ----
inline image8 decode_tga(const bytes& aData, const std::string& aPath)
{
	const auto bad = [&](const char* aWhat) { return std::runtime_error(aPath + ": " + aWhat); };
	if (aData.size() < 18) {
		throw bad("truncated TGA");
	}
	const int id_length = aData[0];
	const int map_type = aData[1];
	const int type = aData[2];
	const int width = le16(&aData[12]);
	const int height = le16(&aData[14]);
	const int bits = aData[16];
	const bool top_down = (aData[17] & 0x20) != 0;
	check_size(width, height, aPath);
	const bool rle = type == 10 || type == 11;
	const bool grey = type == 3 || type == 11;
	if (map_type != 0 || !(type == 2 || type == 3 || rle) || !(grey ? bits == 8 : (bits == 24 || bits == 32))) {
		throw bad("unsupported TGA format (grey 8-bit or true-colour 24/32-bit)");
	}
	const int in_channels = bits / 8;
	const std::size_t count = static_cast<std::size_t>(width) * height;
	bytes stored(count * in_channels);
	std::size_t pos = 18 + id_length;
	if (!rle) {
		if (pos + stored.size() > aData.size()) {
			throw bad("truncated TGA pixel data");
		}
		std::copy_n(&aData[pos], stored.size(), stored.begin());
	} else {
		std::size_t done = 0;
		while (done < count) {
			if (pos >= aData.size()) {
				throw bad("truncated TGA run");
			}
			const int header = aData[pos++];
			const std::size_t run = static_cast<std::size_t>(header & 0x7F) + 1;
			if (done + run > count) {
				throw bad("TGA run exceeds the image");
			}
			if (header & 0x80) {
				if (pos + in_channels > aData.size()) {
					throw bad("truncated TGA run");
				}
				for (std::size_t k = 0; k < run; ++k) {
					std::copy_n(&aData[pos], in_channels, &stored[(done + k) * in_channels]);
				}
				pos += in_channels;
			} else {
				if (pos + run * in_channels > aData.size()) {
					throw bad("truncated TGA run");
				}
				std::copy_n(&aData[pos], run * in_channels, &stored[done * in_channels]);
				pos += run * in_channels;
			}
			done += run;
		}
	}
	const int out_channels = grey ? 1 : 3;
	bytes pixels(count * out_channels);
	for (int y = 0; y < height; ++y) {
		const int src_row = top_down ? y : height - 1 - y;
		for (int x = 0; x < width; ++x) {
			const std::uint8_t* p = &stored[(static_cast<std::size_t>(src_row) * width + x) * in_channels];
			std::uint8_t* d = &pixels[(static_cast<std::size_t>(y) * width + x) * out_channels];
			if (grey) {
				d[0] = p[0];
			} else {
				d[0] = p[2];
				d[1] = p[1];
				d[2] = p[0];
			}
		}
	}
	return make_image(pixels, width, height, out_channels);
}

inline bytes encode_tga(const std::uint8_t* aPixels, int aWidth, int aHeight, int aChannels)
{
	bytes out(18, 0);
	out[2] = aChannels == 1 ? 3 : 2;
	out[12] = static_cast<std::uint8_t>(aWidth);
	out[13] = static_cast<std::uint8_t>(aWidth >> 8);
	out[14] = static_cast<std::uint8_t>(aHeight);
	out[15] = static_cast<std::uint8_t>(aHeight >> 8);
	out[16] = static_cast<std::uint8_t>(8 * aChannels);
	out[17] = 0x20; // top-left origin
	for (std::size_t i = 0; i < static_cast<std::size_t>(aWidth) * aHeight; ++i) {
		const std::uint8_t* p = aPixels + i * aChannels;
		if (aChannels == 1) {
			out.push_back(p[0]);
		} else {
			out.push_back(p[2]);
			out.push_back(p[1]);
			out.push_back(p[0]);
		}
	}
	return out;
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
	bytes data;
	if (extension == ".png") {
		data = encode_png(aPixels, aWidth, aHeight, aChannels);
	} else if (extension == ".bmp") {
		data = encode_bmp(aPixels, aWidth, aHeight, aChannels);
	} else if (extension == ".tga") {
		data = encode_tga(aPixels, aWidth, aHeight, aChannels);
	} else if (extension == ".pgm" || extension == ".ppm") {
		if ((extension == ".pgm") != (aChannels == 1)) {
			throw std::invalid_argument(aPath + ": .pgm holds grey images and .ppm colour images");
		}
		data = encode_pnm(aPixels, aWidth, aHeight, aChannels);
	} else if (extension == ".jpg" || extension == ".jpeg") {
		throw std::invalid_argument(aPath + ": JPEG output is not available in this build (no JPEG codec)");
	} else {
		throw std::invalid_argument(aPath + ": output extension must be .png, .bmp, .tga, .pgm or .ppm");
	}
	write_file(aPath, data);
}

} // namespace detail

/// Detect the format from the contents and decode, discarding alpha.
inline image8 read_image8(const std::string& aPath)
{
	const auto data = detail::read_file(aPath);
	const auto starts = [&](std::initializer_list<int> aMagic) {
		if (data.size() < aMagic.size()) {
			return false;
		}
		std::size_t i = 0;
		for (int m : aMagic) {
			if (data[i++] != m) {
				return false;
			}
		}
		return true;
	};
	if (starts({0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'})) {
		return detail::decode_png(data, aPath);
	}
	if (starts({'B', 'M'})) {
		return detail::decode_bmp(data, aPath);
	}
	if (starts({'P', '5'}) || starts({'P', '6'})) {
		return detail::decode_pnm(data, aPath);
	}
	if (starts({0xFF, 0xD8})) {
		throw std::runtime_error(aPath + ": JPEG decoding is not available in this build (no JPEG codec)");
	}
	if (data.size() >= 18 && data[1] == 0 && (data[2] == 2 || data[2] == 3 || data[2] == 10 || data[2] == 11)) {
		return detail::decode_tga(data, aPath);
	}
	throw std::runtime_error(aPath + ": unable to decode image");
}

/// Decode a grey image (one or two channels).
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
