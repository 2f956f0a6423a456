// Kernels of the learned RL warm start: a small 3-D convolutional network
// (no reference counterpart). Activations are channels-last ([z][y][x][c],
// c fastest); weights are [cout][tap][cin] with tap = (dz * 3 + dy) * 3 + dx.
//
// conv3d_fwd: 3x3x3 "same" cross-correlation with zero padding,
//   out[p][co] = bias[co] + sum_t sum_ci w[co][t][ci] * in[p + d(t) - 1][ci],
// as an implicit GEMM on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation): M = output channels, N = 16 consecutive voxels along x per
// wave, K = 27 * cin. Lane l feeds A[m = l % 16][k = (l / 16) * 8 + j] and
// B[k = (l / 16) * 8 + j][n = l % 16] and receives D[m = (l / 16) * 4 + i][n = l % 16].
// Layer = sum over input channels of the reference convolve with the flipped
// weight and zero border, which is what the oracle checks.
// The input gradient is the same kernel on transformed weights
// (w'[ci][t][co] = w[co][26 - t][ci]); the weight gradient is an FP32
// reduction over voxels.
#include "common.hpp"

namespace iane_hip {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kConvWaves = 4;

struct conv_geom {
	int nx, ny, nz;
	int cin, cout;
};

__device__ __forceinline__ float bf2f(__bf16 v) { return static_cast<float>(v); }

// The biases of output channels co0 .. co0 + 3 (0 for a missing bias or channel).
__device__ __forceinline__ f32x4 bias4(const float* bias, int cout, int co0)
{
	f32x4 b;
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		b[i] = (bias && co0 + i < cout) ? bias[co0 + i] : 0.0f;
	}
	return b;
}

// bias4 for co0 = 4 grp with cout <= 16, read as 16 uniform (scalar) loads and
// selected by lane group, so that no vector-memory wait is ever needed for it.
__device__ __forceinline__ f32x4 bias4_uniform(const float* __restrict__ bias, int cout, int grp)
{
	float all[16];
#pragma unroll
	for (int co = 0; co < 16; ++co) {
		all[co] = (bias && co < cout) ? bias[co] : 0.0f;
	}
	f32x4 b;
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		b[i] = grp == 0 ? all[i] : (grp == 1 ? all[4 + i] : (grp == 2 ? all[8 + i] : all[12 + i]));
	}
	return b;
}

// Write a lane's 4 consecutive output channels co0 .. co0 + 3 of one voxel
// (sum + bias, optional ReLU): one 8-byte (BF16) or 16-byte (FP32) store when
// all 4 exist and cout is a multiple of 4, so a wave writes whole voxels
// contiguously; channel by channel otherwise.
__device__ __forceinline__ void store_outputs(void* out, int out_bf16, int relu, f32x4 bias, long long voxel,
	int cout, int co0, f32x4 acc)
{
	float v[4];
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		const float t = acc[i] + bias[i];
		v[i] = relu ? (t > 0.0f ? t : 0.0f) : t;
	}
	const long long at = voxel * cout + co0;
	if ((cout & 3) == 0 && co0 + 3 < cout) { // aligned: cout and co0 are multiples of 4
		if (out_bf16) {
			typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
			bf16x4 packed;
#pragma unroll
			for (int i = 0; i < 4; ++i) {
				packed[i] = static_cast<__bf16>(v[i]);
			}
			*reinterpret_cast<bf16x4*>(static_cast<__bf16*>(out) + at) = packed;
		} else {
			*reinterpret_cast<f32x4*>(static_cast<float*>(out) + at) = f32x4{v[0], v[1], v[2], v[3]};
		}
		return;
	}
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		if (co0 + i < cout) {
			if (out_bf16) {
				static_cast<__bf16*>(out)[at + i] = static_cast<__bf16>(v[i]);
			} else {
				static_cast<float*>(out)[at + i] = v[i];
			}
		}
	}
}

// Layers with 16 input and at most 16 output channels (the full-resolution
// layers of the network): the same implicit GEMM, with K ordered so that a
// wave slides along y and loads each input row once. K = 3 dy x 2 chunks x
// 3 x shifts x 32: chunk A of kernel row dy holds the rows (dz = 0, dy) (lane
// groups 0, 1) and (dz = 1, dy) (groups 2, 3), chunk B the row (dz = 2, dy)
// (groups 0, 1; groups 2, 3 are zero padding), channels 8 (g % 2) .. + 7. An
// input row y' (its three dz planes) fills one slot of a four-slot register
// ring and serves output rows y' - 1, y', y' + 1 as dy = 2, 1, 0; each output
// row loads one new row (two 16-byte loads and their edges) ahead of its 18
// MFMAs instead of all nine (dz, dy) rows. A lane loads voxel x0 + n of its
// row; lane 0 of each group of 16 also loads voxel x0 - 1 and lane 15 voxel
// x0 + 16 (the "edge" value, zero elsewhere); the x - 1 and x + 1 operands are
// the centre values moved one lane within each group of 16 (DPP row shifts),
// the lane left without a source (0 or 15) taking its edge value. The weights
// of all 18 chunks stay in registers for the wave's kFwd16Rows rows.
__device__ __forceinline__ bf16x8 shift_lanes(bf16x8 aCentre, bf16x8 aEdge, int aDx)
{
	using i32x4 = __attribute__((__vector_size__(4 * sizeof(int)))) int;
	const i32x4 c = __builtin_bit_cast(i32x4, aCentre);
	const i32x4 e = __builtin_bit_cast(i32x4, aEdge);
	i32x4 r;
#pragma unroll
	for (int i = 0; i < 4; ++i) {
		// row_shr:1 (lane n <- lane n - 1) or row_shl:1 (lane n <- lane n + 1); the lane with
		// no source in its row keeps the edge value
		r[i] = aDx < 0 ? __builtin_amdgcn_update_dpp(e[i], c[i], 0x111, 0xF, 0xF, false)
		               : __builtin_amdgcn_update_dpp(e[i], c[i], 0x101, 0xF, 0xF, false);
	}
	return __builtin_bit_cast(bf16x8, r);
}

constexpr int kFwd16Rows = 8;
// A byte offset past the end of every buffer resource made here (the launches
// check that the buffers are smaller): a buffer load there returns 0.
constexpr int kOobOffset = 0x7ffffff0;

__global__ __launch_bounds__(64 * kConvWaves) void conv3d_fwd16_kernel(const __bf16* __restrict__ in,
	const __bf16* __restrict__ w, const float* __restrict__ bias, void* out, int out_bf16, int relu, conv_geom g)
{
	constexpr int kCin = 16;
	const int lane = threadIdx.x & 63;
	const long long wave = static_cast<long long>(blockIdx.x) * kConvWaves + (threadIdx.x >> 6);
	const int xt = (g.nx + 15) / 16;
	const int yb = (g.ny + kFwd16Rows - 1) / kFwd16Rows;
	if (wave >= static_cast<long long>(xt) * yb * g.nz) {
		return;
	}
	const int x0 = static_cast<int>(wave % xt) * 16;
	const int y0 = static_cast<int>((wave / xt) % yb) * kFwd16Rows;
	const int z = static_cast<int>(wave / (static_cast<long long>(xt) * yb));
	const int n = lane & 15;
	const int grp = lane >> 4;
	const int x = x0 + n;
	const int ch = 8 * (grp & 1);
	bf16x8 zero8;
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		zero8[j] = static_cast<__bf16>(0.0f);
	}
	// the input plane of this lane's chunk A and chunk B rows (B: none for groups 2, 3)
	const int dz_a = grp >> 1;
	const bool has_b = grp < 2;
	// A operands: output channel n, tap (dz, dy, dx) of the lane's chunk row
	bf16x8 wreg[3][2][3];
#pragma unroll
	for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			const int dz = c == 0 ? dz_a : 2;
			const bool ok = n < g.cout && (c == 0 || has_b);
#pragma unroll
			for (int d = 0; d < 3; ++d) {
				const long long at = ok ? (static_cast<long long>(n) * 27 + dz * 9 + dy * 3 + d) * kCin + ch : 0;
				wreg[dy][c][d] = *reinterpret_cast<const bf16x8*>(w + at);
				if (!ok) {
					wreg[dy][c][d] = zero8;
				}
			}
		}
	}
	// the whole input volume as one buffer resource (the launch checks that its
	// byte size fits the 32-bit offsets)
	const auto volume = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(in), static_cast<short>(0),
		g.nx * g.ny * g.nz * kCin * 2, 0x00020000);
	// the edge voxel of lanes 0 and 15 of each group (x0 - 1, x0 + 16), if inside
	const int edge_x = n == 0 ? x0 - 1 : (n == 15 ? x0 + 16 : -1);
	const bool edge_ok = edge_x >= 0 && edge_x < g.nx;
	const f32x4 bv = bias4(bias, g.cout, grp * 4);
	// ring slot of input row y0 - 1 + i: i % 4; chunk c's centre and edge
	bf16x8 centre[4][2], edge[4][2];
	const auto load_row = [&](int aI) {
		const int yi = y0 - 1 + aI;
#pragma unroll
		for (int c = 0; c < 2; ++c) {
			const int zi = z + (c == 0 ? dz_a : 2) - 1;
			const bool row_ok = (c == 0 || has_b) && yi >= 0 && yi < g.ny && zi >= 0 && zi < g.nz;
			// byte offsets into the volume's buffer resource; past its end reads 0
			const int row = ((zi * g.ny + yi) * g.nx) * (kCin * 2) + ch * 2;
			const int centre_at = row_ok && x < g.nx ? row + x * (kCin * 2) : kOobOffset;
			const int edge_at = row_ok && edge_ok ? row + edge_x * (kCin * 2) : kOobOffset;
			centre[aI % 4][c] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(volume, centre_at, 0, 0));
			edge[aI % 4][c] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(volume, edge_at, 0, 0));
		}
	};
	load_row(0);
	load_row(1);
	load_row(2);
#pragma unroll
	for (int j = 0; j < kFwd16Rows; ++j) {
		const int y = y0 + j;
		// the row two ahead into the free slot, then this row's products
		// (each MFMA waits only for its own operands)
		if (j + 3 < kFwd16Rows + 2) {
			load_row(j + 3);
		}
		__builtin_amdgcn_sched_barrier(0);
		f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
		for (int dy = 0; dy < 3; ++dy) {
			const int slot = (j + dy) % 4;
#pragma unroll
			for (int c = 0; c < 2; ++c) {
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[dy][c][0], shift_lanes(centre[slot][c], edge[slot][c], -1),
					acc, 0, 0, 0);
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[dy][c][1], centre[slot][c], acc, 0, 0, 0);
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wreg[dy][c][2], shift_lanes(centre[slot][c], edge[slot][c], 1),
					acc, 0, 0, 0);
			}
		}
		if (x < g.nx && y < g.ny) {
			const long long voxel = (static_cast<long long>(z) * g.ny + y) * g.nx + x;
			store_outputs(out, out_bf16, relu, bv, voxel, g.cout, grp * 4, acc);
		}
	}
}

// The input layer (one input channel, at most 16 output channels) on the same
// MFMA: K = the 27 taps (zero-padded to 32), so lane group g holds taps
// 8g .. 8g + 7 of its voxel x0 + n and the weights of output channel n for
// those taps; one MFMA per row gives all output channels of 16 voxels. A wave
// computes kCin1Rows consecutive y rows and issues the samples of all of them
// (one 2-byte buffer load per tap, zero outside the volume) before its first
// MFMA.
constexpr int kCin1Rows = 8;

__global__ __launch_bounds__(64 * kConvWaves) void conv3d_cin1_kernel(const __bf16* __restrict__ in,
	const __bf16* __restrict__ w, const float* __restrict__ bias, void* out, int out_bf16, int relu, conv_geom g)
{
	const int lane = threadIdx.x & 63;
	const long long wave = static_cast<long long>(blockIdx.x) * kConvWaves + (threadIdx.x >> 6);
	const int xt = (g.nx + 15) / 16;
	const int yb = (g.ny + kCin1Rows - 1) / kCin1Rows;
	if (wave >= static_cast<long long>(xt) * yb * g.nz) {
		return;
	}
	const int x0 = static_cast<int>(wave % xt) * 16;
	const int y0 = static_cast<int>((wave / xt) % yb) * kCin1Rows;
	const int z = static_cast<int>(wave / (static_cast<long long>(xt) * yb));
	const int n = lane & 15;
	const int grp = lane >> 4;
	const int x = x0 + n;
	const f32x4 bv = bias4_uniform(bias, g.cout, grp);
	bf16x8 a;
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		const int t = grp * 8 + j;
		a[j] = n < g.cout && t < 27 ? w[n * 27 + t] : static_cast<__bf16>(0.0f);
	}
	const auto volume = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(in), static_cast<short>(0),
		g.nx * g.ny * g.nz * 2, 0x00020000);
	// samples[row][j]: tap 8 grp + j of voxel (x, y0 + row, z), one per 32-bit register
	unsigned samples[kCin1Rows][8];
#pragma unroll
	for (int r = 0; r < kCin1Rows; ++r) {
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const int t = grp * 8 + j;
			const int xi = x + t % 3 - 1, yi = y0 + r + (t / 3) % 3 - 1, zi = z + t / 9 - 1;
			const bool ok = t < 27 && x < g.nx && y0 + r < g.ny && xi >= 0 && xi < g.nx && yi >= 0 && yi < g.ny
				&& zi >= 0 && zi < g.nz;
			const int at = ok ? ((zi * g.ny + yi) * g.nx + xi) * 2 : kOobOffset;
			samples[r][j] = __builtin_amdgcn_raw_buffer_load_b16(volume, at, 0, 0);
		}
	}
	// keep the loads ahead of the MFMAs
	__builtin_amdgcn_sched_barrier(0);
#pragma unroll
	for (int r = 0; r < kCin1Rows; ++r) {
		bf16x8 b;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			b[j] = __builtin_bit_cast(__bf16, static_cast<unsigned short>(samples[r][j]));
		}
		const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
		if (x < g.nx && y0 + r < g.ny) {
			const long long voxel = (static_cast<long long>(z) * g.ny + y0 + r) * g.nx + x;
			store_outputs(out, out_bf16, relu, bv, voxel, g.cout, grp * 4, acc);
		}
	}
}

// One wave: 16 voxels along x in kFwdRows consecutive rows y0 .. y0 + kFwdRows - 1
// at z, all output channels (<= 64); the rows share each K chunk's weight loads.
// When cin is a multiple of 8 and the input and weights fit 32-bit byte
// offsets (grouped != 0), K is walked kFwdGroup<MT> chunks at a time with all
// of a group's loads issued before its MFMAs (buffer loads: an offset past the
// end reads 0, so padding and out-of-volume taps need no branch); the chunks
// left over, and other cin, take one chunk at a time.
constexpr int kFwdRows = 4;

template<int MT>
constexpr int kFwdGroup = MT == 1 ? 5 : (MT == 2 ? 4 : 3);

template<int MT>
__global__ __launch_bounds__(64 * kConvWaves) void conv3d_fwd_kernel(const __bf16* __restrict__ in,
	const __bf16* __restrict__ w, const float* __restrict__ bias, void* out, int out_bf16, int relu, conv_geom g,
	int grouped)
{
	const int lane = threadIdx.x & 63;
	const long long wave = static_cast<long long>(blockIdx.x) * kConvWaves + (threadIdx.x >> 6);
	const int xt = (g.nx + 15) / 16;
	const int yb = (g.ny + kFwdRows - 1) / kFwdRows;
	if (wave >= static_cast<long long>(xt) * yb * g.nz) {
		return;
	}
	const int x0 = static_cast<int>(wave % xt) * 16;
	const int y0 = static_cast<int>((wave / xt) % yb) * kFwdRows;
	const int z = static_cast<int>(wave / (static_cast<long long>(xt) * yb));
	const int n = lane & 15;
	const int grp = lane >> 4;
	const int x = x0 + n;
	const int K = 27 * g.cin;
	const int chunks = (K + 31) / 32;
	f32x4 acc[kFwdRows][MT];
#pragma unroll
	for (int r = 0; r < kFwdRows; ++r) {
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			acc[r][mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
		}
	}
	const bool vec = (g.cin & 7) == 0;
	int kc = 0;
	if (grouped) {
		constexpr int U = kFwdGroup<MT>;
		const auto volume = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(in), static_cast<short>(0),
			g.nx * g.ny * g.nz * g.cin * 2, 0x00020000);
		const auto weights = __builtin_amdgcn_make_buffer_rsrc(const_cast<__bf16*>(w), static_cast<short>(0),
			g.cout * K * 2, 0x00020000);
		for (; kc + U <= chunks; kc += U) {
			bf16x8 b[U][kFwdRows];
			bf16x8 a[U][MT];
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const int k0 = (kc + u) * 32 + grp * 8;
				// 8 consecutive input channels of one tap (cin is a multiple of 8)
				const int t = k0 / g.cin;
				const int c = k0 - t * g.cin;
				const int xi = x + t % 3 - 1, zi = z + t / 9 - 1;
				const bool ok_xz = t < 27 && x < g.nx && xi >= 0 && xi < g.nx && zi >= 0 && zi < g.nz;
#pragma unroll
				for (int r = 0; r < kFwdRows; ++r) {
					const int yi = y0 + r + (t / 3) % 3 - 1;
					const bool ok = ok_xz && yi >= 0 && yi < g.ny;
					const int at = ok ? (((zi * g.ny + yi) * g.nx + xi) * g.cin + c) * 2 : kOobOffset;
					b[u][r] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(volume, at, 0, 0));
				}
#pragma unroll
				for (int mt = 0; mt < MT; ++mt) {
					const int m = mt * 16 + n;
					const int wat = m < g.cout && t < 27 ? (m * K + k0) * 2 : kOobOffset;
					a[u][mt] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(weights, wat, 0, 0));
				}
			}
			// keep the group's loads ahead of its MFMAs
			__builtin_amdgcn_sched_barrier(0);
#pragma unroll
			for (int u = 0; u < U; ++u) {
#pragma unroll
				for (int r = 0; r < kFwdRows; ++r) {
#pragma unroll
					for (int mt = 0; mt < MT; ++mt) {
						acc[r][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u][mt], b[u][r], acc[r][mt], 0, 0, 0);
					}
				}
			}
		}
	}
	for (; kc < chunks; ++kc) {
		const int k0 = kc * 32 + grp * 8;
		bf16x8 b[kFwdRows];
#pragma unroll
		for (int r = 0; r < kFwdRows; ++r) {
			const int y = y0 + r; // a row past the volume computes values that are never stored
			if (vec) {
				// 8 consecutive input channels of one tap
				const int t = k0 / g.cin;
				const int c = k0 - t * g.cin;
				const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
				const int xi = x + dx, yi = y + dy, zi = z + dz;
				const bool ok = t < 27 && x < g.nx && xi >= 0 && xi < g.nx && yi >= 0 && yi < g.ny && zi >= 0 && zi < g.nz;
				if (ok) {
					const long long voxel = (static_cast<long long>(zi) * g.ny + yi) * g.nx + xi;
					b[r] = *reinterpret_cast<const bf16x8*>(in + voxel * g.cin + c);
				} else {
					for (int j = 0; j < 8; ++j) {
						b[r][j] = static_cast<__bf16>(0.0f);
					}
				}
			} else {
#pragma unroll
				for (int j = 0; j < 8; ++j) {
					const int k = k0 + j;
					const int t = k / g.cin;
					const int c = k - t * g.cin;
					const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
					const int xi = x + dx, yi = y + dy, zi = z + dz;
					const bool ok = k < K && x < g.nx && xi >= 0 && xi < g.nx && yi >= 0 && yi < g.ny && zi >= 0 && zi < g.nz;
					const long long voxel = (static_cast<long long>(zi) * g.ny + yi) * g.nx + xi;
					b[r][j] = ok ? in[voxel * g.cin + c] : static_cast<__bf16>(0.0f);
				}
			}
		}
#pragma unroll
		for (int mt = 0; mt < MT; ++mt) {
			const int m = mt * 16 + n;
			bf16x8 a;
			if (m < g.cout && vec) {
				const int t = k0 / g.cin;
				if (t < 27) {
					a = *reinterpret_cast<const bf16x8*>(w + static_cast<long long>(m) * K + k0);
				} else {
					for (int j = 0; j < 8; ++j) {
						a[j] = static_cast<__bf16>(0.0f);
					}
				}
			} else {
#pragma unroll
				for (int j = 0; j < 8; ++j) {
					const int k = k0 + j;
					a[j] = (m < g.cout && k < K) ? w[static_cast<long long>(m) * K + k] : static_cast<__bf16>(0.0f);
				}
			}
#pragma unroll
			for (int r = 0; r < kFwdRows; ++r) {
				acc[r][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[r], acc[r][mt], 0, 0, 0);
			}
		}
	}
	if (x >= g.nx) {
		return;
	}
#pragma unroll
	for (int r = 0; r < kFwdRows; ++r) {
		if (y0 + r < g.ny) {
			const long long voxel = (static_cast<long long>(z) * g.ny + y0 + r) * g.nx + x;
#pragma unroll
			for (int mt = 0; mt < MT; ++mt) {
				store_outputs(out, out_bf16, relu, bias4(bias, g.cout, mt * 16 + grp * 4), voxel, g.cout,
					mt * 16 + grp * 4, acc[r][mt]);
			}
		}
	}
}

// w'[ci][t][co] = w[co][26 - t][ci]: the input gradient as a forward correlation.
__global__ void conv3d_transpose_weights_kernel(const __bf16* w, __bf16* wt, int cin, int cout)
{
	const int total = cin * cout * 27;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
		const int co = i % cout;
		const int t = (i / cout) % 27;
		const int ci = i / (cout * 27);
		wt[i] = w[(static_cast<long long>(co) * 27 + (26 - t)) * cin + ci];
	}
}

// dw[co][t][ci] += sum_p dy[p][co] * in[p + d(t) - 1][ci]; dbias[co] += sum_p dy[p][co].
// Block = one tap and a chunk of voxels; threads own (co, ci) pairs; FP32 partials
// are added to the FP32 gradient with atomics.
__global__ __launch_bounds__(256) void conv3d_wgrad_kernel(const __bf16* in, const float* dy, float* dw, float* dbias,
	conv_geom g, int chunk)
{
	extern __shared__ float sh[];
	float* s_dy = sh;                          // [chunk_tile][cout]
	float* s_in = sh + 64 * g.cout;            // [chunk_tile][cin]
	const int t = blockIdx.y;
	const int dx = t % 3 - 1, dyy = (t / 3) % 3 - 1, dz = t / 9 - 1;
	const long long nvox = static_cast<long long>(g.nx) * g.ny * g.nz;
	const long long p0 = static_cast<long long>(blockIdx.x) * chunk;
	const long long p1 = p0 + chunk < nvox ? p0 + chunk : nvox;
	const int pairs = g.cout * g.cin;
	float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	float bacc = 0.0f;
	for (long long base = p0; base < p1; base += 64) {
		const int tile = static_cast<int>(p1 - base < 64 ? p1 - base : 64);
		__syncthreads();
		for (int i = threadIdx.x; i < tile * g.cout; i += blockDim.x) {
			s_dy[i] = dy[base * g.cout + i];
		}
		for (int i = threadIdx.x; i < tile * g.cin; i += blockDim.x) {
			const int v = i / g.cin;
			const int c = i - v * g.cin;
			const long long p = base + v;
			const int x = static_cast<int>(p % g.nx);
			const int y = static_cast<int>((p / g.nx) % g.ny);
			const int z = static_cast<int>(p / (static_cast<long long>(g.nx) * g.ny));
			const int xi = x + dx, yi = y + dyy, zi = z + dz;
			float val = 0.0f;
			if (xi >= 0 && xi < g.nx && yi >= 0 && yi < g.ny && zi >= 0 && zi < g.nz) {
				val = bf2f(in[((static_cast<long long>(zi) * g.ny + yi) * g.nx + xi) * g.cin + c]);
			}
			s_in[i] = val;
		}
		__syncthreads();
		for (int r = 0; r < 8; ++r) {
			const int pair = threadIdx.x + r * 256;
			if (pair < pairs) {
				const int co = pair / g.cin;
				const int ci = pair - co * g.cin;
				float s = acc[r];
				for (int v = 0; v < tile; ++v) {
					s = fmaf(s_dy[v * g.cout + co], s_in[v * g.cin + ci], s);
				}
				acc[r] = s;
			}
		}
		if (t == 13 && static_cast<int>(threadIdx.x) < g.cout) {
			for (int v = 0; v < tile; ++v) {
				bacc += s_dy[v * g.cout + threadIdx.x];
			}
		}
	}
	for (int r = 0; r < 8; ++r) {
		const int pair = threadIdx.x + r * 256;
		if (pair < pairs) {
			const int co = pair / g.cin;
			const int ci = pair - co * g.cin;
			atomicAdd(&dw[(static_cast<long long>(co) * 27 + t) * g.cin + ci], acc[r]);
		}
	}
	if (t == 13 && static_cast<int>(threadIdx.x) < g.cout && dbias) {
		atomicAdd(&dbias[threadIdx.x], bacc);
	}
}

// dy *= (y > 0) for post-ReLU activations y (bf16), dy f32
__global__ void relu_backward_kernel(const __bf16* y, float* dy, long long n)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		if (!(bf2f(y[i]) > 0.0f)) {
			dy[i] = 0.0f;
		}
	}
}

// 2x2x2 average pooling, channels-last bf16 -> bf16 (extents must be even)
__global__ void pool2_kernel(const __bf16* in, __bf16* out, int nx, int ny, int nz, int c)
{
	const int ox = nx / 2, oy = ny / 2, oz = nz / 2;
	const long long total = static_cast<long long>(ox) * oy * oz * c;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int ch = static_cast<int>(i % c);
		const long long v = i / c;
		const int x = static_cast<int>(v % ox), y = static_cast<int>((v / ox) % oy), z = static_cast<int>(v / (static_cast<long long>(ox) * oy));
		float s = 0.0f;
		for (int d = 0; d < 8; ++d) {
			const int xi = 2 * x + (d & 1), yi = 2 * y + ((d >> 1) & 1), zi = 2 * z + (d >> 2);
			s += bf2f(in[((static_cast<long long>(zi) * ny + yi) * nx + xi) * c + ch]);
		}
		out[i] = static_cast<__bf16>(s * 0.125f);
	}
}

// gradient of pool2: dx[fine] = dy[coarse] / 8 (f32)
__global__ void pool2_backward_kernel(const float* dy, float* dx, int nx, int ny, int nz, int c, int accumulate)
{
	const long long total = static_cast<long long>(nx) * ny * nz * c;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int ch = static_cast<int>(i % c);
		const long long v = i / c;
		const int x = static_cast<int>(v % nx), y = static_cast<int>((v / nx) % ny), z = static_cast<int>(v / (static_cast<long long>(nx) * ny));
		const float g = dy[((static_cast<long long>(z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2) * c + ch] * 0.125f;
		dx[i] = accumulate ? dx[i] + g : g;
	}
}

// out[fine] = up2(coarse)[fine] + skip[fine] (nearest upsampling), bf16
// Channel counts that are multiples of 8: one thread per 8 channels (16 bytes)
// of a voxel, y and z from the grid (no 64-bit division per element); the
// same arithmetic, in the same order, as the kernels below.
__global__ void up2_add8_kernel(const __bf16* coarse, const __bf16* skip, __bf16* out, int nx, int ny, int c)
{
	const int c8 = c / 8;
	const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= nx * c8) {
		return;
	}
	const int x = t / c8;
	const int g = t - x * c8;
	const int y = static_cast<int>(blockIdx.y);
	const int z = static_cast<int>(blockIdx.z);
	const long long fine = ((static_cast<long long>(z) * ny + y) * nx + x) * c + 8 * g;
	const long long crs = ((static_cast<long long>(z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2) * c + 8 * g;
	const bf16x8 u = *reinterpret_cast<const bf16x8*>(coarse + crs);
	const bf16x8 v = *reinterpret_cast<const bf16x8*>(skip + fine);
	bf16x8 o;
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		o[j] = static_cast<__bf16>(bf2f(u[j]) + bf2f(v[j]));
	}
	*reinterpret_cast<bf16x8*>(out + fine) = o;
}

// FINE extents nx, ny; grid over the coarse voxels
__global__ void pool2_8_kernel(const __bf16* in, __bf16* out, int nx, int ny, int c)
{
	const int ox = nx / 2, oy = ny / 2;
	const int c8 = c / 8;
	const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= ox * c8) {
		return;
	}
	const int x = t / c8;
	const int g = t - x * c8;
	const int y = static_cast<int>(blockIdx.y);
	const int z = static_cast<int>(blockIdx.z);
	float sum[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
	for (int d = 0; d < 8; ++d) {
		const int xi = 2 * x + (d & 1), yi = 2 * y + ((d >> 1) & 1), zi = 2 * z + (d >> 2);
		const bf16x8 v = *reinterpret_cast<const bf16x8*>(in + ((static_cast<long long>(zi) * ny + yi) * nx + xi) * c + 8 * g);
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			sum[j] += bf2f(v[j]);
		}
	}
	bf16x8 o;
#pragma unroll
	for (int j = 0; j < 8; ++j) {
		o[j] = static_cast<__bf16>(sum[j] * 0.125f);
	}
	*reinterpret_cast<bf16x8*>(out + ((static_cast<long long>(z) * oy + y) * ox + x) * c + 8 * g) = o;
}

__global__ void up2_add_kernel(const __bf16* coarse, const __bf16* skip, __bf16* out, int nx, int ny, int nz, int c)
{
	const long long total = static_cast<long long>(nx) * ny * nz * c;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int ch = static_cast<int>(i % c);
		const long long v = i / c;
		const int x = static_cast<int>(v % nx), y = static_cast<int>((v / nx) % ny), z = static_cast<int>(v / (static_cast<long long>(nx) * ny));
		const float u = bf2f(coarse[((static_cast<long long>(z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2) * c + ch]);
		out[i] = static_cast<__bf16>(u + bf2f(skip[i]));
	}
}

// gradient of up2: dcoarse = sum of the 8 fine gradients
__global__ void up2_backward_kernel(const float* dfine, float* dcoarse, int nx, int ny, int nz, int c)
{
	const int ox = nx / 2, oy = ny / 2, oz = nz / 2;
	const long long total = static_cast<long long>(ox) * oy * oz * c;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int ch = static_cast<int>(i % c);
		const long long v = i / c;
		const int x = static_cast<int>(v % ox), y = static_cast<int>((v / ox) % oy), z = static_cast<int>(v / (static_cast<long long>(ox) * oy));
		float s = 0.0f;
		for (int d = 0; d < 8; ++d) {
			const int xi = 2 * x + (d & 1), yi = 2 * y + ((d >> 1) & 1), zi = 2 * z + (d >> 2);
			s += dfine[((static_cast<long long>(zi) * ny + yi) * nx + xi) * c + ch];
		}
		dcoarse[i] = s;
	}
}

// f32 -> bf16 with a scale (input normalisation and weight copies)
__global__ void to_bf16_kernel(const float* in, __bf16* out, long long n, float scale)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		out[i] = static_cast<__bf16>(in[i] * scale);
	}
}

// Adam on FP32 master weights, then the BF16 copy the forward kernel reads.
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, __bf16* wb, long long n, float lr, float b1,
	float b2, float eps, float c1, float c2)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float gi = g[i];
		const float mi = b1 * m[i] + (1.0f - b1) * gi;
		const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
		m[i] = mi;
		v[i] = vi;
		const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
		w[i] = wi;
		if (wb) {
			wb[i] = static_cast<__bf16>(wi);
		}
	}
}

// residual head: pred = u + r; grad = 2 (pred - target) / n; loss partial per block
__global__ void residual_mse_kernel(const float* u, const float* r, const float* target, float* pred, float* grad,
	long long n, double* partials)
{
	__shared__ double shared[4];
	double loss = 0.0;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float p = u[i] + r[i];
		if (pred) {
			pred[i] = p;
		}
		const float d = p - (target ? target[i] : 0.0f);
		if (grad) {
			grad[i] = 2.0f * d / static_cast<float>(n);
		}
		loss += static_cast<double>(d) * d;
	}
	for (int off = 32; off > 0; off >>= 1) {
		loss += __shfl_down(loss, off, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		shared[threadIdx.x >> 6] = loss;
	}
	__syncthreads();
	if (threadIdx.x == 0 && partials) {
		partials[blockIdx.x] = shared[0] + shared[1] + shared[2] + shared[3];
	}
}

// warm start x0 = max(scale * (u + r), eps)
__global__ void warm_start_kernel(const float* u, const float* r, float* x, long long n, float scale, float eps)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float v = scale * (u[i] + r[i]);
		x[i] = v > eps ? v : eps;
	}
}

// ---------------------------------------------------- training-data synthesis --
// volume[p] = background + sum_b amp_b exp(-|p + 0.5 - c_b|^2 / (2 sigma_b^2)); blobs: (cz, cy, cx, sigma, amp)
__global__ void blobs_kernel(float* out, int nx, int ny, int nz, float background, const float* blobs, int count)
{
	const long long total = static_cast<long long>(nx) * ny * nz;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float x = static_cast<float>(i % nx) + 0.5f;
		const float y = static_cast<float>((i / nx) % ny) + 0.5f;
		const float z = static_cast<float>(i / (static_cast<long long>(nx) * ny)) + 0.5f;
		float s = background;
		for (int b = 0; b < count; ++b) {
			const float* q = blobs + 5 * b;
			const float dz = z - q[0], dy = y - q[1], dx = x - q[2];
			const float r2 = dx * dx + dy * dy + dz * dz;
			const float inv = 1.0f / (2.0f * q[3] * q[3]);
			if (r2 * inv < 16.0f) {
				s += q[4] * expf(-r2 * inv);
			}
		}
		out[i] = s;
	}
}

// Philox4x32-10 (Salmon et al., SC'11), key (seed, stream), counter (lo, hi, 0, 0)
__device__ inline void philox(unsigned seed, unsigned stream, unsigned long long counter, unsigned (&r)[4])
{
	unsigned c0 = static_cast<unsigned>(counter), c1 = static_cast<unsigned>(counter >> 32), c2 = 0, c3 = 0;
	unsigned k0 = seed, k1 = stream;
	for (int i = 0; i < 10; ++i) {
		const unsigned long long p0 = 0xD2511F53ull * c0;
		const unsigned long long p1 = 0xCD9E8D57ull * c2;
		const unsigned hi0 = static_cast<unsigned>(p0 >> 32), lo0 = static_cast<unsigned>(p0);
		const unsigned hi1 = static_cast<unsigned>(p1 >> 32), lo1 = static_cast<unsigned>(p1);
		const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
		c0 = n0;
		c1 = lo1;
		c2 = n2;
		c3 = lo0;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	r[0] = c0;
	r[1] = c1;
	r[2] = c2;
	r[3] = c3;
}

__device__ inline double unit(unsigned v) { return static_cast<double>(v >> 8) * (1.0 / 16777216.0); }

// Poisson variates: inversion for lambda < 10, PTRS (Hörmann 1993) otherwise;
// sample i uses Philox counter block (attempt * n + i).
__global__ void poisson_kernel(const float* lam, float* out, long long n, unsigned seed)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const double l = lam[i] > 0.0f ? lam[i] : 0.0f;
		unsigned r[4];
		double k = 0.0;
		if (l < 10.0) {
			philox(seed, 0, static_cast<unsigned long long>(i), r);
			const double u = unit(r[0]);
			double p = exp(-l), cdf = p;
			while (u > cdf && k < 1000.0) {
				k += 1.0;
				p *= l / k;
				cdf += p;
			}
		} else {
			const double slam = sqrt(l), loglam = log(l);
			const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
			const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4), v_r = 0.9277 - 3.6224 / (b - 2.0);
			for (int attempt = 0; attempt < 200; ++attempt) {
				philox(seed, 1, static_cast<unsigned long long>(attempt) * n + i, r);
				const double u = unit(r[0]) - 0.5, v = unit(r[1]);
				const double us = 0.5 - fabs(u);
				k = floor((2.0 * a / us + b) * u + l + 0.43);
				if (us >= 0.07 && v <= v_r) {
					break;
				}
				if (k < 0.0 || (us < 0.013 && v > us)) {
					continue;
				}
				if (log(v * inv_alpha / (a / (us * us) + b)) <= -l + k * loglam - lgamma(k + 1.0)) {
					break;
				}
			}
		}
		out[i] = static_cast<float>(k);
	}
}

__global__ void max_kernel(const float* in, long long n, float* partial)
{
	__shared__ float s[256];
	float m = -3.402823466e38f;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		m = fmaxf(m, in[i]);
	}
	s[threadIdx.x] = m;
	__syncthreads();
	for (unsigned off = 128; off > 0; off >>= 1) {
		if (threadIdx.x < off) {
			s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + off]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		partial[blockIdx.x] = s[0];
	}
}

__global__ void scale_kernel(float* x, long long n, float scale)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		x[i] *= scale;
	}
}

unsigned grid_for(long long n)
{
	return static_cast<unsigned>(std::min<long long>(4096, std::max<long long>(1, (n + 255) / 256)));
}

} // namespace
} // namespace iane_hip

using namespace iane_hip;

extern "C" {

// out: f32 (out_bf16 = 0) or bf16 channels-last [z][y][x][cout]; bias may be NULL.
int iane_hip_cnn_conv3d(const void* in, const void* weights, const float* bias, void* out, int out_bf16, int relu,
	int nx, int ny, int nz, int cin, int cout, void* stream)
{
	return guarded([&] {
		if (!in || !weights || !out) {
			invalid("conv3d buffers are null");
		}
		if (nx <= 0 || ny <= 0 || nz <= 0 || cin <= 0 || cout <= 0 || cout > 64 || cin > 256) {
			invalid("conv3d extents or channel counts out of range (cout <= 64, cin <= 256)");
		}
		const conv_geom g{nx, ny, nz, cin, cout};
		const long long waves = static_cast<long long>((nx + 15) / 16) * ((ny + kFwdRows - 1) / kFwdRows) * nz;
		const unsigned blocks = static_cast<unsigned>((waves + kConvWaves - 1) / kConvWaves);
		auto s = static_cast<hipStream_t>(stream);
		const auto* x = static_cast<const __bf16*>(in);
		const auto* w = static_cast<const __bf16*>(weights);
		if (cin == 1 && cout <= 16 && static_cast<long long>(nx) * ny * nz * 2 < kOobOffset - 64) {
			const long long waves1 = static_cast<long long>((nx + 15) / 16) * ((ny + kCin1Rows - 1) / kCin1Rows) * nz;
			const unsigned blocks1 = static_cast<unsigned>((waves1 + kConvWaves - 1) / kConvWaves);
			conv3d_cin1_kernel<<<blocks1, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g);
			check_hip(hipGetLastError(), "conv3d launch");
			return;
		}
		if (cin == 16 && cout <= 16 && static_cast<long long>(nx) * ny * nz * 32 < kOobOffset - 64) {
			const long long waves16 = static_cast<long long>((nx + 15) / 16) * ((ny + kFwd16Rows - 1) / kFwd16Rows) * nz;
			const unsigned blocks16 = static_cast<unsigned>((waves16 + kConvWaves - 1) / kConvWaves);
			conv3d_fwd16_kernel<<<blocks16, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g);
			check_hip(hipGetLastError(), "conv3d launch");
			return;
		}
		// grouped buffer loads need cin % 8 == 0 and byte sizes below the out-of-range offset
		const int grouped = cin % 8 == 0 && static_cast<long long>(nx) * ny * nz * cin * 2 < kOobOffset - 64 ? 1 : 0;
		switch ((cout + 15) / 16) {
		case 1: conv3d_fwd_kernel<1><<<blocks, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g, grouped); break;
		case 2: conv3d_fwd_kernel<2><<<blocks, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g, grouped); break;
		case 3: conv3d_fwd_kernel<3><<<blocks, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g, grouped); break;
		default: conv3d_fwd_kernel<4><<<blocks, 64 * kConvWaves, 0, s>>>(x, w, bias, out, out_bf16, relu, g, grouped); break;
		}
		check_hip(hipGetLastError(), "conv3d launch");
	});
}

int iane_hip_cnn_transpose_weights(const void* w, void* wt, int cin, int cout, void* stream)
{
	return guarded([&] {
		conv3d_transpose_weights_kernel<<<grid_for(static_cast<long long>(cin) * cout * 27), 256, 0,
			static_cast<hipStream_t>(stream)>>>(static_cast<const __bf16*>(w), static_cast<__bf16*>(wt), cin, cout);
		check_hip(hipGetLastError(), "weight transpose launch");
	});
}

// dw (f32 [cout][27][cin]) and dbias (f32 [cout], may be NULL) are ACCUMULATED: clear them first.
int iane_hip_cnn_wgrad(const void* in, const float* dy, float* dw, float* dbias, int nx, int ny, int nz, int cin,
	int cout, void* stream)
{
	return guarded([&] {
		if (cin * cout > 2048 || cout > 256) {
			invalid("wgrad supports cin * cout <= 2048");
		}
		const conv_geom g{nx, ny, nz, cin, cout};
		const long long nvox = static_cast<long long>(nx) * ny * nz;
		const int chunk = 4096;
		const dim3 grid(static_cast<unsigned>((nvox + chunk - 1) / chunk), 27);
		const size_t lds = sizeof(float) * 64 * (cin + cout);
		conv3d_wgrad_kernel<<<grid, 256, lds, static_cast<hipStream_t>(stream)>>>(static_cast<const __bf16*>(in), dy, dw,
			dbias, g, chunk);
		check_hip(hipGetLastError(), "wgrad launch");
	});
}

int iane_hip_cnn_relu_backward(const void* y, float* dy, int64_t n, void* stream)
{
	return guarded([&] {
		relu_backward_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const __bf16*>(y), dy, n);
		check_hip(hipGetLastError(), "relu backward launch");
	});
}

int iane_hip_cnn_pool2(const void* in, void* out, int nx, int ny, int nz, int c, void* stream)
{
	return guarded([&] {
		if (nx % 2 || ny % 2 || nz % 2) {
			invalid("pool2 needs even extents");
		}
		if (c % 8 == 0 && ny / 2 <= 65535 && nz / 2 <= 65535 && static_cast<long long>(nx / 2) * (c / 8) < (1LL << 30)) {
			const dim3 grid(static_cast<unsigned>((nx / 2 * (c / 8) + 255) / 256), static_cast<unsigned>(ny / 2),
				static_cast<unsigned>(nz / 2));
			if (grid.x && grid.y && grid.z) {
				pool2_8_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const __bf16*>(in),
					static_cast<__bf16*>(out), nx, ny, c);
			}
		} else {
			pool2_kernel<<<grid_for(static_cast<long long>(nx) * ny * nz * c / 8), 256, 0, static_cast<hipStream_t>(stream)>>>(
				static_cast<const __bf16*>(in), static_cast<__bf16*>(out), nx, ny, nz, c);
		}
		check_hip(hipGetLastError(), "pool2 launch");
	});
}

// extents are the FINE extents
int iane_hip_cnn_pool2_backward(const float* dy, float* dx, int nx, int ny, int nz, int c, int accumulate, void* stream)
{
	return guarded([&] {
		pool2_backward_kernel<<<grid_for(static_cast<long long>(nx) * ny * nz * c), 256, 0, static_cast<hipStream_t>(stream)>>>(
			dy, dx, nx, ny, nz, c, accumulate);
		check_hip(hipGetLastError(), "pool2 backward launch");
	});
}

// extents are the FINE extents
int iane_hip_cnn_up2_add(const void* coarse, const void* skip, void* out, int nx, int ny, int nz, int c, void* stream)
{
	return guarded([&] {
		if (c % 8 == 0 && ny <= 65535 && nz <= 65535 && static_cast<long long>(nx) * (c / 8) < (1LL << 30)) {
			const dim3 grid(static_cast<unsigned>((nx * (c / 8) + 255) / 256), static_cast<unsigned>(ny),
				static_cast<unsigned>(nz));
			if (grid.x && grid.y && grid.z) {
				up2_add8_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(static_cast<const __bf16*>(coarse),
					static_cast<const __bf16*>(skip), static_cast<__bf16*>(out), nx, ny, c);
			}
		} else {
			up2_add_kernel<<<grid_for(static_cast<long long>(nx) * ny * nz * c), 256, 0, static_cast<hipStream_t>(stream)>>>(
				static_cast<const __bf16*>(coarse), static_cast<const __bf16*>(skip), static_cast<__bf16*>(out), nx, ny, nz, c);
		}
		check_hip(hipGetLastError(), "up2 launch");
	});
}

// extents are the FINE extents
int iane_hip_cnn_up2_backward(const float* dfine, float* dcoarse, int nx, int ny, int nz, int c, void* stream)
{
	return guarded([&] {
		up2_backward_kernel<<<grid_for(static_cast<long long>(nx) * ny * nz * c / 8), 256, 0, static_cast<hipStream_t>(stream)>>>(
			dfine, dcoarse, nx, ny, nz, c);
		check_hip(hipGetLastError(), "up2 backward launch");
	});
}

int iane_hip_cnn_to_bf16(const float* in, void* out, int64_t n, float scale, void* stream)
{
	return guarded([&] {
		to_bf16_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(in, static_cast<__bf16*>(out), n, scale);
		check_hip(hipGetLastError(), "to bf16 launch");
	});
}

int iane_hip_cnn_adam(float* w, const float* g, float* m, float* v, void* w_bf16, int64_t n, float lr, float beta1,
	float beta2, float eps, int step, void* stream)
{
	return guarded([&] {
		const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
		const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
		adam_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(w, g, m, v, static_cast<__bf16*>(w_bf16), n, lr,
			beta1, beta2, eps, c1, c2);
		check_hip(hipGetLastError(), "adam launch");
	});
}

// pred = u + r, grad = 2 (pred - target) / n, loss partials (<= 1024 blocks) into `partials`
int iane_hip_cnn_residual_mse(const float* u, const float* r, const float* target, float* pred, float* grad, int64_t n,
	double* partials, int* blocks_used, void* stream)
{
	return guarded([&] {
		const unsigned blocks = static_cast<unsigned>(std::min<long long>(1024, std::max<long long>(1, (n + 255) / 256)));
		residual_mse_kernel<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(u, r, target, pred, grad, n, partials);
		check_hip(hipGetLastError(), "mse launch");
		if (blocks_used) {
			*blocks_used = static_cast<int>(blocks);
		}
	});
}

int iane_hip_synth_blobs(float* out, int nx, int ny, int nz, float background, const float* blobs, int count,
	void* stream)
{
	return guarded([&] {
		blobs_kernel<<<grid_for(static_cast<long long>(nx) * ny * nz), 256, 0, static_cast<hipStream_t>(stream)>>>(out, nx,
			ny, nz, background, blobs, count);
		check_hip(hipGetLastError(), "blobs launch");
	});
}

int iane_hip_synth_poisson(const float* lambda, float* out, int64_t n, unsigned seed, void* stream)
{
	return guarded([&] {
		poisson_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(lambda, out, n, seed);
		check_hip(hipGetLastError(), "poisson launch");
	});
}

// device max of n floats -> *out (device float); partial needs 4096 floats of scratch
int iane_hip_max(const float* in, int64_t n, float* partial, float* out, void* stream)
{
	return guarded([&] {
		auto s = static_cast<hipStream_t>(stream);
		const unsigned blocks = grid_for(n);
		max_kernel<<<blocks, 256, 0, s>>>(in, n, partial);
		max_kernel<<<1, 256, 0, s>>>(partial, blocks, out);
		check_hip(hipGetLastError(), "max launch");
	});
}

int iane_hip_scale(float* x, int64_t n, float scale, void* stream)
{
	return guarded([&] {
		scale_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(x, n, scale);
		check_hip(hipGetLastError(), "scale launch");
	});
}

int iane_hip_cnn_warm_start(const float* u, const float* r, float* x, int64_t n, float scale, float eps, void* stream)
{
	return guarded([&] {
		warm_start_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(u, r, x, n, scale, eps);
		check_hip(hipGetLastError(), "warm start launch");
	});
}

} // extern "C"
