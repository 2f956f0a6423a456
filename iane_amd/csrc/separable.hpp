// Shared by the separable convolution engines (conv_separable.hip: FP32 and
// vector engines, per-axis passes, dispatch; conv_separable_split.hip: the
// split-FP16 matrix engine). Each source is compiled with its own scheduling
// options, so the engines live in separate translation units.
#pragma once

#include "common.hpp"

namespace iane_hip {
namespace separable {

struct sep_args {
	const char* src;
	int src_type;
	int nx, ny, nz;
	long long sx, sy, sz;       // source byte strides
	char* dst;
	int dst_type;
	long long dx, dy, dz;       // destination byte strides
	int border;
	unsigned tiles_x, tiles_y, chunks_z;
	int chunk;                  // output planes per workgroup
	const char* taps[3];        // device tap vectors (null: the unit kernel, a 2-D image's z axis)
	long long tap_stride[3];
	int tap_type[3];
	int tap_len[3];             // odd tap counts, each at most the engine's compile-time length
	int src_span, dst_span;     // bytes of one source / destination plane that a launch addresses (matrix engine)
};

using iane_hip::plane_resource;

// Taps at window positions aIndex[0..N) of axis aAxis in an engine's window of
// KA taps (KA odd, at least the axis' odd tap count): the axis' taps centred in
// the window, zero around them (a shorter kernel is the longer one with zero
// outer taps, so every axis can use the engine's one compile-time length); a
// null tap pointer is the unit kernel (1 at the centre). One element-type
// dispatch per axis and no branch per tap: a position outside the axis' taps
// reads its first tap and selects zero, so all the axis' loads are in flight
// together (a switch around each load would put each in its own block and
// wait for it there).
template<int KA, int N>
__device__ __forceinline__ void sep_taps(const sep_args& a, int aAxis, const int (&aIndex)[N], float (&aOut)[N])
{
	const char* taps = a.taps[aAxis];
	if (taps == nullptr) {
#pragma unroll
		for (int i = 0; i < N; ++i) {
			aOut[i] = aIndex[i] == KA / 2 ? 1.0f : 0.0f;
		}
		return;
	}
	const int len = a.tap_len[aAxis];
	const int shift = (KA - len) / 2;
	const long long stride = a.tap_stride[aAxis];
	with_element_type(a.tap_type[aAxis], [&](auto aType) {
		using E = decltype(aType);
		E raw[N];
		bool ok[N];
#pragma unroll
		for (int i = 0; i < N; ++i) {
			const int t = aIndex[i] - shift;
			ok[i] = aIndex[i] >= 0 && aIndex[i] < KA && t >= 0 && t < len;
			raw[i] = *reinterpret_cast<const E*>(taps + (ok[i] ? t : 0) * stride);
		}
#pragma unroll
		for (int i = 0; i < N; ++i) {
			aOut[i] = ok[i] ? static_cast<float>(raw[i]) : 0.0f;
		}
	});
}


// Split-FP16 engine (conv_separable_split.hip): output chunks addressed with
// 32-bit offsets need aChunk planes to fit; the launcher takes the common
// tap length (odd, <= 15).
bool split_chunk_fits(const sep_args& a, int aChunk);
void launch_split(sep_args& a, int aTaps, hipStream_t aStream);

} // namespace separable
} // namespace iane_hip
