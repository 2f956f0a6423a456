// Split-precision MFMA direct convolution (IANE_PREC_SPLIT_FP16 / _BF16).
//
// The reference sums taps in FP64 (include/iane/algorithms/convolution.hpp:36-49);
// this path computes the same sum on the matrix cores with FP32 accumulation.
// Every FP32 operand v is split into 16-bit parts. FP16: v ~ hi + lo * 2^-L,
// 11+11 significant bits (power-of-two scales keep the values inside FP16
// range), and a product is formed from three MFMAs: hi*hi into one accumulator
// and hi*lo + lo*hi into a second one scaled by 2^-L at the end (lo*lo is below
// FP32 resolution). BF16 (FP32 exponent range, 8 significant bits per part):
// v ~ hi + mid + lo, 24 bits, and six MFMAs: hi*hi into the first accumulator,
// hi*mid + mid*hi + hi*lo + lo*hi + mid*mid (the terms down to 2^-16 of the
// leading one) into the second; mid*lo and smaller are below FP32 resolution.
//
// GEMM form. Along x the convolution is a Toeplitz product: for a window of
// 16 outputs starting at x0, out[x0 + m] = sum_k A[m][k] * in[x0 - hx + k] with
// A[m][k] = kernel[qz][qy][m + 2 hx - k] (zero outside the kernel). One
// v_mfma_f32_16x16x32 takes that A (16 x 32) and, as its 16 columns, the
// windows of 16 consecutive x-tiles of ONE input row: a wave covers a
// 256-wide output strip of a row with one MFMA per (input row, qz, qy). A wave
// owns the accumulators of 16 output rows (y .. y+15 at one z) and streams the
// input rows past them; the A matrices of a chunk of qy values stay in
// registers, so each row load feeds up to 16 x 3 MFMAs.
//
// Border handling is done once, up front: the source is written into a padded
// buffer (hx, hy, hz margins filled by the border rule, values already split
// into hi and lo planes), so the hot loop has no bounds checks and every load
// is 16-byte aligned.
#include "common.hpp"

#include <type_traits>

namespace iane_hip {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 b16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 16;         // output rows per wave
constexpr int kStrip = 256;       // output x per wave (16 tiles of 16)
constexpr int kWaves = 4;         // waves per workgroup
constexpr int kLoShiftFp16 = 12;  // lo = (v - hi) * 2^12 for FP16

struct split_geom {
	int nx, ny, nz;       // image extents
	int hx, hy, hz;       // kernel half extents
	int kz;               // kernel extent along z
	int qy_padded;        // ky rounded up to the qy chunk
	int px, py, pz;       // padded buffer extents (halves per row, rows, planes)
	int strips, yblocks;  // wave tasks per plane
	long long plane;      // halves per padded plane
	long long part;       // halves per hi (or lo) buffer
};

template<class T16>
struct split16;

template<>
struct split16<_Float16> {
	using vec = f16x8;
	static constexpr int parts = 2;
	static constexpr int lo_shift = kLoShiftFp16;
	__device__ static _Float16 to(float v) { return static_cast<_Float16>(v); }
	__device__ static float from(_Float16 v) { return static_cast<float>(v); }
	__device__ static f32x4 mfma(vec a, vec b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};

template<>
struct split16<__bf16> {
	using vec = b16x8;
	static constexpr int parts = 3;
	static constexpr int lo_shift = 0;
	__device__ static __bf16 to(float v) { return static_cast<__bf16>(v); }
	__device__ static float from(__bf16 v) { return static_cast<float>(v); }
	__device__ static f32x4 mfma(vec a, vec b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// ------------------------------------------------------------------ scales --
// Per-block maximum of |element| over a strided rank<=3 view.
__global__ void absmax_partial_kernel(view4 v, float* partial)
{
	__shared__ float shared[4];
	const long long count = static_cast<long long>(v.extent[0]) * v.extent[1] * v.extent[2];
	float m = 0.0f;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < count;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const long long x = i % v.extent[0];
		const long long y = (i / v.extent[0]) % v.extent[1];
		const long long z = i / (static_cast<long long>(v.extent[0]) * v.extent[1]);
		const char* p = v.ptr + x * v.stride[0] + y * v.stride[1] + z * v.stride[2];
		m = fmaxf(m, fabsf(static_cast<float>(load_as_double(p, v.type))));
	}
	for (int off = 32; off > 0; off >>= 1) {
		m = fmaxf(m, __shfl_down(m, off, 64));
	}
	if ((threadIdx.x & 63) == 0) {
		shared[threadIdx.x >> 6] = m;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float r = shared[0];
		for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
			r = fmaxf(r, shared[w]);
		}
		partial[blockIdx.x] = r;
	}
}

// *scale = 2^(14 - e) with max < 2^e (1 for a zero maximum, or when `aUnit`).
__global__ void scale_from_partials_kernel(const float* partial, int count, float* scale, int unit)
{
	__shared__ float shared[4];
	float m = 0.0f;
	for (int i = threadIdx.x; i < count; i += blockDim.x) {
		m = fmaxf(m, partial[i]);
	}
	for (int off = 32; off > 0; off >>= 1) {
		m = fmaxf(m, __shfl_down(m, off, 64));
	}
	if ((threadIdx.x & 63) == 0) {
		shared[threadIdx.x >> 6] = m;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		float r = shared[0];
		for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
			r = fmaxf(r, shared[w]);
		}
		int e = 0;
		frexpf(r, &e);
		*scale = (unit || r == 0.0f) ? 1.0f : ldexpf(1.0f, 14 - e);
	}
}

// Write the 16-bit parts of v (scaled by 2^L after the first for FP16) to
// aOut[0], aOut[aStride], ...
template<class T16>
__device__ __forceinline__ void store_parts(float v, T16* aOut, long long aStride)
{
	using S = split16<T16>;
	const float lo_scale = static_cast<float>(1 << S::lo_shift);
	const T16 hi = S::to(v);
	aOut[0] = hi;
	const float rest = (v - S::from(hi)) * lo_scale;
	if constexpr (S::parts == 2) {
		aOut[aStride] = S::to(rest);
	} else {
		const T16 mid = S::to(rest);
		aOut[aStride] = mid;
		aOut[2 * aStride] = S::to(rest - S::from(mid));
	}
}

// ------------------------------------------------------------ padded split --
// Part planes of the padded source: padded (xp, yp, zp) holds the sample at
// (xp - hx, yp - hy, zp - hz) under the border rule.
template<class T16>
__global__ void split_pad_kernel(view4 src, int border, split_geom g, const float* src_scale, T16* out)
{
	const float sd = *src_scale;
	const long long total = g.part;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int xp = static_cast<int>(i % g.px);
		const int yp = static_cast<int>((i / g.px) % g.py);
		const int zp = static_cast<int>(i / (static_cast<long long>(g.px) * g.py));
		int x = xp - g.hx, y = yp - g.hy, z = zp - g.hz;
		float v = 0.0f;
		const bool inside = x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
		if (inside || border == IANE_BORDER_CLAMP) {
			x = min(max(x, 0), g.nx - 1);
			y = min(max(y, 0), g.ny - 1);
			z = min(max(z, 0), g.nz - 1);
			v = load_as_float(src.ptr + x * src.stride[0] + y * src.stride[1] + z * src.stride[2], src.type) * sd;
		}
		store_parts(v, out + i, total);
	}
}

// ------------------------------------------------------------------ A table --
// table[((qz * qy_padded + qy) * steps + s) * parts + part][lane][j]: the A
// operand (one part) of the Toeplitz matrix of kernel row (qz, qy), K-step s,
// in the MFMA lane layout (lane l holds A[l % 16][32 s + 8 (l / 16) + j]).
template<class T16>
__global__ void toeplitz_table_kernel(view4 ker, int hx, int qy_padded, int steps, const float* ker_scale, T16* table)
{
	using S = split16<T16>;
	const float sk = *ker_scale;
	const int kx = ker.extent[0], ky = ker.extent[1], kz = ker.extent[2];
	const long long entries = static_cast<long long>(kz) * qy_padded * steps * 64 * 8;
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < entries;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int j = static_cast<int>(i % 8);
		const int lane = static_cast<int>((i / 8) % 64);
		const long long row = i / (8 * 64);  // (qz * qy_padded + qy) * steps + s
		const int s = static_cast<int>(row % steps);
		const int qy = static_cast<int>((row / steps) % qy_padded);
		const int qz = static_cast<int>(row / (static_cast<long long>(steps) * qy_padded));
		const int m = lane % 16;
		const int k = 32 * s + 8 * (lane / 16) + j;
		const int qx = m + 2 * hx - k;
		float v = 0.0f;
		if (qy < ky && qx >= 0 && qx < kx) {
			v = static_cast<float>(load_as_double(ker.ptr + qx * ker.stride[0] + qy * ker.stride[1] + qz * ker.stride[2],
				ker.type)) * sk;
		}
		store_parts(v, table + (row * S::parts) * 64 * 8 + lane * 8 + j, 64 * 8);
	}
}

// ------------------------------------------------------------- main kernel --
// A pass is one (qz, qy chunk): its QC Toeplitz operands (every part and
// K-step) sit in registers while the U = kRows + QC - 1 input rows it needs
// stream past them; input row u meets operand j for output row
// r = u - (QC - 1) + j. Loads run ahead of the matrix work, with one wave per
// SIMD and nothing else to hide their latency:
//  - input rows D rows ahead through a ring of D row buffers; the ring runs
//    on into the next pass, so its first rows arrive while this pass's last
//    rows (which meet few operands) are multiplied;
//  - operand j of the next pass is loaded right after its last use in this
//    one (output row kRows - 1), QC rows before its first use.
// Rows and operands come through buffer resources with the row offset in a
// scalar register. The order of the products is the plain (pass, row, j)
// order, so the sums do not depend on D.
template<class T16, int QC, int STEPS, int D>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(1, 1)))
void conv_split_kernel(const T16* __restrict__ padded, const T16* __restrict__ table,
	const float* __restrict__ src_scale, const float* __restrict__ ker_scale, view4 dst, split_geom g, unsigned tasks)
{
	using S = split16<T16>;
	using vec = typename S::vec;
	constexpr int P = S::parts;
	constexpr int U = kRows + QC - 1;        // input rows per pass
	constexpr int UP = (U + D - 1) / D * D;  // rows per pass including idle ring slots
	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	// The 4 waves of a workgroup take 4 consecutive z planes of one (strip,
	// y block), and consecutive workgroups continue along z: workgroups that
	// run together on one XCD then read overlapping input planes from its L2.
	// wave-uniform (the buffer resources below are scalar)
	const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const unsigned zgroups = (static_cast<unsigned>(g.nz) + kWaves - 1) / kWaves;
	const unsigned zgroup = block % zgroups;
	const unsigned column = block / zgroups;
	const int z = static_cast<int>(zgroup * kWaves + wave);
	if (block >= tasks || z >= g.nz) {
		return;
	}
	const int strip = static_cast<int>(column % g.strips);
	const int y0 = static_cast<int>(column / g.strips) * kRows;
	const int x0 = strip * kStrip;
	const int lane = threadIdx.x & 63;
	const int n = lane & 15;
	const int grp = lane >> 4;

	f32x4 acc_hi[kRows];
	f32x4 acc_lo[kRows];
#pragma unroll
	for (int r = 0; r < kRows; ++r) {
		acc_hi[r] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
		acc_lo[r] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	}
	const int chunks = g.qy_padded / QC;
	const int passes = g.kz * chunks;
	const char* parts = reinterpret_cast<const char*>(padded);
	const long long part_bytes = 2 * g.part;
	const long long plane_bytes = 2 * g.plane;
	const int row_bytes = 2 * g.px;
	const int lane_bytes = 2 * (x0 + 16 * n + 8 * grp);  // this lane's first B element in a padded row
	const auto tab = plane_resource(reinterpret_cast<const char*>(table),
		static_cast<int>(static_cast<long long>(g.kz) * g.qy_padded * STEPS * P * 1024));

	vec a[P][QC][STEPS];
	vec b[D][P][STEPS];
	// operand j of pass t (the last pass again past the end)
	const auto load_operands = [&](int aPass, int aJ) {
		const int t = min(aPass, passes - 1);
		const int qz = t / chunks;
		const int qy = (t - qz * chunks) * QC + aJ;
#pragma unroll
		for (int s = 0; s < STEPS; ++s) {
			const int row = (qz * g.qy_padded + qy) * STEPS + s;
#pragma unroll
			for (int p = 0; p < P; ++p) {
				a[p][aJ][s] = __builtin_bit_cast(vec,
					__builtin_amdgcn_raw_buffer_load_b128(tab, 16 * lane, (row * P + p) * 1024, 0));
			}
		}
	};
	// input row u of pass t into ring slot aSlot
	const auto load_row = [&](int aPass, int aU, int aSlot) {
		const int t = min(aPass, passes - 1);
		const int qz = t / chunks;
		const int q0 = (t - qz * chunks) * QC;
		// rows < 0 only meet zero operands (padded qy); rows past U are idle slots
		const int yp = min(max(y0 + 2 * g.hy - q0 - (QC - 1) + aU, 0), g.py - 1);
		const char* plane = parts + static_cast<long long>(z + 2 * g.hz - qz) * plane_bytes;
#pragma unroll
		for (int p = 0; p < P; ++p) {
			const auto rows = plane_resource(plane + p * part_bytes, static_cast<int>(plane_bytes));
#pragma unroll
			for (int s = 0; s < STEPS; ++s) {
				b[aSlot][p][s] = __builtin_bit_cast(vec,
					__builtin_amdgcn_raw_buffer_load_b128(rows, lane_bytes + 64 * s, yp * row_bytes, 0));
			}
		}
	};
#pragma unroll
	for (int j = 0; j < QC; ++j) {
		load_operands(0, j);
	}
#pragma unroll
	for (int d = 0; d < D; ++d) {
		load_row(0, d, d);
	}
	for (int t = 0; t < passes; ++t) {
#pragma unroll
		for (int u = 0; u < UP; ++u) {
			const int slot = u % D;
#pragma unroll
			for (int j = 0; j < QC; ++j) {
				const int r = u - (QC - 1) + j;
				if (u < U && r >= 0 && r < kRows) {
#pragma unroll
					for (int s = 0; s < STEPS; ++s) {
						acc_hi[r] = S::mfma(a[0][j][s], b[slot][0][s], acc_hi[r]);
						acc_lo[r] = S::mfma(a[0][j][s], b[slot][1][s], acc_lo[r]);
						acc_lo[r] = S::mfma(a[1][j][s], b[slot][0][s], acc_lo[r]);
						if constexpr (P == 3) {
							acc_lo[r] = S::mfma(a[0][j][s], b[slot][2][s], acc_lo[r]);
							acc_lo[r] = S::mfma(a[2][j][s], b[slot][0][s], acc_lo[r]);
							acc_lo[r] = S::mfma(a[1][j][s], b[slot][1][s], acc_lo[r]);
						}
					}
				}
			}
			// the slot is consumed: row u + D, of this pass or the next
			if (u + D < UP) {
				load_row(t, u + D, slot);
			} else {
				load_row(t + 1, u + D - UP, slot);
			}
			// operands whose last use was this row: the next pass's
#pragma unroll
			for (int j = 0; j < QC; ++j) {
				if (u == kRows + QC - 2 - j) {
					load_operands(t + 1, j);
				}
			}
			// keep each row's loads behind its products: hoisted above them
			// they would need a second set of row registers
			__builtin_amdgcn_sched_barrier(0);
		}
	}

	const float unscale = 1.0f / (*src_scale * *ker_scale);
	const float lo_unscale = 1.0f / static_cast<float>(1 << S::lo_shift);
	const int xs = x0 + 16 * n + 4 * grp;
#pragma unroll
	for (int r = 0; r < kRows; ++r) {
		const int y = y0 + r;
		if (y >= g.ny) {
			break;
		}
		char* row = dst.ptr + static_cast<long long>(y) * dst.stride[1] + static_cast<long long>(z) * dst.stride[2];
#pragma unroll
		for (int i = 0; i < 4; ++i) {
			const int x = xs + i;
			if (x < g.nx) {
				const float v = (acc_hi[r][i] + acc_lo[r][i] * lo_unscale) * unscale;
				char* p = row + static_cast<long long>(x) * dst.stride[0];
				if (dst.type == IANE_F32) {
					*reinterpret_cast<float*>(p) = v;
				} else {
					*reinterpret_cast<double*>(p) = static_cast<double>(v);
				}
			}
		}
	}
}

// Tile variants: QC rows of A kept in registers, STEPS K-steps of 32 along x.
enum split_variant { SPLIT_NONE = 0, SPLIT_QC8_S1, SPLIT_QC15_S1, SPLIT_QC8_S2 };

split_variant pick_variant(const view4& aKer)
{
	const int kx = aKer.extent[0], ky = aKer.extent[1];
	const int steps = (15 + kx + 31) / 32;
	if (steps == 1) {
		return ky <= 8 ? SPLIT_QC8_S1 : SPLIT_QC15_S1;
	}
	return steps == 2 ? SPLIT_QC8_S2 : SPLIT_NONE;
}

int variant_qc(split_variant aV) { return aV == SPLIT_QC15_S1 ? 15 : 8; }
int variant_steps(split_variant aV) { return aV == SPLIT_QC8_S2 ? 2 : 1; }

split_geom make_geom(const view4& aSrc, const view4& aKer, split_variant aV)
{
	const int qc = variant_qc(aV), steps = variant_steps(aV);
	split_geom g{};
	g.nx = aSrc.extent[0];
	g.ny = aSrc.extent[1];
	g.nz = aSrc.extent[2];
	g.hx = aKer.extent[0] / 2;
	g.hy = aKer.extent[1] / 2;
	g.hz = aKer.extent[2] / 2;
	g.kz = aKer.extent[2];
	g.qy_padded = ((aKer.extent[1] + qc - 1) / qc) * qc;
	g.strips = (g.nx + kStrip - 1) / kStrip;
	g.yblocks = (g.ny + kRows - 1) / kRows;
	g.px = g.strips * kStrip + 32 * steps;
	g.py = g.yblocks * kRows + 2 * g.hy;
	g.pz = g.nz + 2 * g.hz;
	g.plane = static_cast<long long>(g.px) * g.py;
	g.part = g.plane * g.pz;
	return g;
}

constexpr int kPartialBlocks = 1024;
constexpr size_t kScaleBytes = 256;                                  // one float, padded
constexpr size_t kPartialBytes = sizeof(float) * kPartialBlocks;      // block maxima

size_t round_up(size_t aBytes) { return (aBytes + 255) / 256 * 256; }

// 16-bit parts per operand: hi + lo (FP16), hi + mid + lo (BF16).
int split_parts(int aPrecision) { return aPrecision == IANE_PREC_SPLIT_BF16 ? 3 : 2; }

// Halves of the Toeplitz table of one kernel (every part, every K-step).
long long table_halves(const view4& aKer, split_variant aV, int aParts)
{
	const int qc = variant_qc(aV);
	const long long qy_padded = (aKer.extent[1] + qc - 1) / qc * qc;
	return static_cast<long long>(aKer.extent[2]) * qy_padded * variant_steps(aV) * aParts * 64 * 8;
}

template<class T16>
void build_kernel_part(const view4& aKer, split_variant aV, int aUnit, float* aKerScale, float* aPartial,
	T16* aTable, hipStream_t aStream)
{
	const int qc = variant_qc(aV);
	const int qy_padded = (aKer.extent[1] + qc - 1) / qc * qc;
	absmax_partial_kernel<<<64, 256, 0, aStream>>>(aKer, aPartial);
	scale_from_partials_kernel<<<1, 256, 0, aStream>>>(aPartial, 64, aKerScale, aUnit);
	toeplitz_table_kernel<T16><<<ceil_div(table_halves(aKer, aV, split16<T16>::parts) / split16<T16>::parts, 256), 256,
		0, aStream>>>(aKer, aKer.extent[0] / 2,
		qy_padded, variant_steps(aV), aKerScale, aTable);
	check_hip(hipGetLastError(), "split convolution kernel table");
}

template<class T16, int QC, int STEPS>
void run_source_part(const view4& aSrc, const split_geom& aGeom, const view4& aDst, int aBorder, int aUnit,
	char* aWork, const T16* aTable, const float* aKerScale, hipStream_t aStream)
{
	auto* src_scale = reinterpret_cast<float*>(aWork);
	auto* partial = reinterpret_cast<float*>(aWork + kScaleBytes);
	auto* padded = reinterpret_cast<T16*>(aWork + kScaleBytes + kPartialBytes);
	absmax_partial_kernel<<<kPartialBlocks, 256, 0, aStream>>>(aSrc, partial);
	scale_from_partials_kernel<<<1, 256, 0, aStream>>>(partial, kPartialBlocks, src_scale, aUnit);
	split_pad_kernel<T16><<<ceil_div(aGeom.part, 256 * 8), 256, 0, aStream>>>(aSrc, aBorder, aGeom, src_scale, padded);
	const unsigned tasks = static_cast<unsigned>(aGeom.strips * aGeom.yblocks) * ceil_div(aGeom.nz, kWaves);
	// input rows in flight: 6 (a 15-row chunk's pass boundary then has about
	// 400 cycles of matrix work between a row's load and its use); fewer for
	// the three-part BF16 operands, within the register file
	constexpr int D = split16<T16>::parts == 3 ? (STEPS == 2 ? 4 : 3) : 6;
	conv_split_kernel<T16, QC, STEPS, D><<<tasks, 64 * kWaves, 0, aStream>>>(padded, aTable, src_scale, aKerScale,
		aDst, aGeom, tasks);
	check_hip(hipGetLastError(), "split convolution launch");
}

template<class T16>
void run_source_part(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, split_variant aV,
	char* aWork, const void* aTable, const float* aKerScale, hipStream_t aStream)
{
	const int unit = std::is_same_v<T16, __bf16> ? 1 : 0;
	const split_geom g = make_geom(aSrc, aKer, aV);
	const auto* table = static_cast<const T16*>(aTable);
	switch (aV) {
	case SPLIT_QC8_S1: run_source_part<T16, 8, 1>(aSrc, g, aDst, aBorder, unit, aWork, table, aKerScale, aStream); break;
	case SPLIT_QC15_S1: run_source_part<T16, 15, 1>(aSrc, g, aDst, aBorder, unit, aWork, table, aKerScale, aStream); break;
	default: run_source_part<T16, 8, 2>(aSrc, g, aDst, aBorder, unit, aWork, table, aKerScale, aStream); break;
	}
}

} // namespace

// ------------------------------------------------------------ entry points --
// The MFMA path tiles rank <= 3 images with float destinations and kernels up
// to 49 wide along x; other shapes run the FP32 direct kernel instead.
bool split_applicable(const view4& aSrc, const view4& aKer, const view4& aDst)
{
	if (!(aSrc.extent[3] == 1 && aKer.extent[3] == 1 && (aDst.type == IANE_F32 || aDst.type == IANE_F64) &&
		aSrc.extent[0] < (1 << 24) && aSrc.extent[1] < (1 << 24) && pick_variant(aKer) != SPLIT_NONE)) {
		return false;
	}
	// the kernel addresses a padded plane and the operand table through 32-bit
	// buffer offsets
	const split_variant v = pick_variant(aKer);
	const split_geom g = make_geom(aSrc, aKer, v);
	return 2 * g.plane < (1LL << 31) && 2 * table_halves(aKer, v, 3) < (1LL << 31);
}

size_t split_kernel_bytes(const view4& aKer, int aPrecision)
{
	// a kernel scale, the block maxima of the kernel, and the table (2-byte halves)
	return kScaleBytes + kPartialBytes +
		round_up(2 * static_cast<size_t>(table_halves(aKer, pick_variant(aKer), split_parts(aPrecision))));
}

size_t split_source_bytes(const view4& aSrc, const view4& aKer, int aPrecision)
{
	const split_geom g = make_geom(aSrc, aKer, pick_variant(aKer));
	return kScaleBytes + kPartialBytes +
		round_up(split_parts(aPrecision) * sizeof(uint16_t) * static_cast<size_t>(g.part));
}

// Build the kernel-side data (scale and Toeplitz table) into `aKernelPart`
// (split_kernel_bytes of memory); it stays valid while the kernel values do.
void split_prepare_kernel(const view4& aKer, int aPrecision, void* aKernelPart, hipStream_t aStream)
{
	char* base = static_cast<char*>(aKernelPart);
	auto* ker_scale = reinterpret_cast<float*>(base);
	auto* partial = reinterpret_cast<float*>(base + kScaleBytes);
	void* table = base + kScaleBytes + kPartialBytes;
	const split_variant v = pick_variant(aKer);
	if (aPrecision == IANE_PREC_SPLIT_FP16) {
		build_kernel_part<_Float16>(aKer, v, 0, ker_scale, partial, static_cast<_Float16*>(table), aStream);
	} else {
		build_kernel_part<__bf16>(aKer, v, 1, ker_scale, partial, static_cast<__bf16*>(table), aStream);
	}
}

// Convolve with a kernel prepared by split_prepare_kernel; `aSourcePart` is
// split_source_bytes of scratch (the source's scale and padded part planes).
void split_convolve_prepared(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, int aPrecision,
	const void* aKernelPart, void* aSourcePart, hipStream_t aStream)
{
	const char* base = static_cast<const char*>(aKernelPart);
	const auto* ker_scale = reinterpret_cast<const float*>(base);
	const void* table = base + kScaleBytes + kPartialBytes;
	const split_variant v = pick_variant(aKer);
	char* work = static_cast<char*>(aSourcePart);
	if (aPrecision == IANE_PREC_SPLIT_FP16) {
		run_source_part<_Float16>(aSrc, aKer, aDst, aBorder, v, work, table, ker_scale, aStream);
	} else {
		run_source_part<__bf16>(aSrc, aKer, aDst, aBorder, v, work, table, ker_scale, aStream);
	}
}

} // namespace iane_hip
