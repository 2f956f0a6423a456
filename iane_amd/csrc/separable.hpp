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
	const char* taps[3];        // device tap vectors
	long long tap_stride[3];
	int tap_type[3];
	int src_span, dst_span;     // bytes of one source / destination plane that a launch addresses (matrix engine)
};

// Raw buffer resource of one plane (GFX9 word 3: 32-bit data format). Loads
// and stores through it take a scalar base and a VGPR byte offset, and
// offsets at or past `aBytes` read 0 / are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_resource(const char* aBase, int aBytes)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(aBase), static_cast<short>(0), aBytes, 0x00020000);
}

// Split-FP16 engine (conv_separable_split.hip): output chunks addressed with
// 32-bit offsets need aChunk planes to fit; the launcher takes the common
// tap length (odd, <= 15).
bool split_chunk_fits(const sep_args& a, int aChunk);
void launch_split(sep_args& a, int aTaps, hipStream_t aStream);

} // namespace separable
} // namespace iane_hip
