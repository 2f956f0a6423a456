// Separable convolution, engine 3 (precision split_fp16): the split-FP16
// matrix engine behind iane_hip_convolve_separable for rank-3 FP32 images with
// zero borders and 16-byte sample groups. Dispatch, the FP32 and vector
// engines and the per-axis passes are in conv_separable.hip; this kernel is a
// translation unit of its own so it can be compiled with its own scheduling
// options.
#include "separable.hpp"

namespace iane_hip {
namespace separable {
namespace {

constexpr int kSxTileX = 64;  // a workgroup's 4 waves side by side in x
constexpr int kSxTileY = 16;
typedef float sx_f32x4 __attribute__((ext_vector_type(4)));

// Split-FP16 matrix engine (engine 3, precision split_fp16): the x and y
// filters (and, in z mode 0, the z filter) are Toeplitz products on
// v_mfma_f32_16x16x32_f16, whose rate is 16 times that of the FP32 matrix
// instruction. Every operand v is carried as two FP16
// parts after an exact power-of-two scale S: hi = f16(v S) rounded toward
// zero, lo = f16(v S - hi) rounded to nearest (21-22 significant bits), and a
// product is three MFMAs (lo*hi, hi*lo, hi*hi) into one FP32 accumulator.
//
// Scales. A plane's x operands take S = 2^(141 - e), e the biased exponent of
// the largest |sample| of the plane region a wave reads (one wave max), so
// they lie below 2^15. The taps of axis a are scaled by 2^b_a (largest tap in
// [2^14, 2^15)), and each filter's output is brought back below 2^15 for the
// next filter by the constant 2^-(b_a + s_a), 2^s_a > sum |taps|. Planes with
// different e meet in the z filter: the z taps that output plane o applies to
// plane i are multiplied by 2^(e_i - E_o), E_o the largest e of the K planes
// output o reads, which gives all products of an output the same scale; the
// output is rescaled with ldexp. Accuracy is therefore relative to the largest
// sample of a wave's 32 x 32 plane region and of a K-plane band, not only to
// the largest sample of the image.
//
// Layout. Each wave owns a 16 x 16 (x, y) output tile and streams a z chunk.
//   x  A = 32 rows x 32-wide window of one plane (lane (g, m): row m, window
//      columns 4g..4g+3 and 16+4g..16+4g+3, two 16-byte loads), B = taps;
//      two row tiles cover the 16 + K - 1 rows the y filter needs;
//   y  A = taps in the K order the x accumulators leave the rows in
//      (lane group g holds rows 4g..4g+3 of both row tiles), B = x outputs;
//   z  mode 1 (default): the y outputs (lane (g, n): rows 4g..4g+3 at column
//      n), rescaled to their true values, enter a register ring of 16 planes
//      and the z taps run as FP32 FMAs; no LDS, so two waves per SIMD fit.
//      Mode 0: the y outputs go, split, into a per-wave LDS ring of 32 planes,
//      column-major (16 y + x, 32 FP16 slots and 4 pad halves per column:
//      8-byte reads of 8 slots hit distinct banks); every 16 planes a window
//      of 32 ring planes gives 16 output planes: A = compensated taps, B = 8
//      consecutive ring slots of a column per lane group, one MFMA triple per
//      tile row (148 KB of LDS per workgroup: one wave per SIMD).
// No barriers: each wave's LDS region is its own. Zero border only: loads of
// 16-byte groups wholly outside the image read 0 through the buffer
// resource's range check.
typedef _Float16 sx_f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned sx_u32x4 __attribute__((ext_vector_type(4)));

constexpr int kSxColDwords = 18;                        // 32 FP16 slots + 4 pad halves per column
constexpr int kSxPartDwords = 256 * kSxColDwords;       // one part (hi or lo) of a wave's ring
constexpr int kSxWaveDwords = 2 * kSxPartDwords;

// Split of a pair of scaled values into packed FP16 parts, in plain vector
// instructions (the compiler sees them, so it keeps the wait states that
// matrix instructions reading them need): hi = RTZ(v s) packed, which for |v s|
// in the FP16 normal range equals v s with its low 13 mantissa bits cleared,
// so the remainder v s - hi is one AND and one subtraction and exact; lo is
// the remainder rounded to nearest.
typedef _Float16 sx_f16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float sx_rest(float aV)
{
	return aV - __uint_as_float(__float_as_uint(aV) & 0xffffe000u);
}

__device__ __forceinline__ void sx_pair(float aA, float aB, unsigned& aHi, unsigned& aLo)
{
	aHi = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(aA, aB));
	aLo = __builtin_bit_cast(unsigned, sx_f16x2{static_cast<_Float16>(sx_rest(aA)), static_cast<_Float16>(sx_rest(aB))});
}

// One MFMA operand (8 values, times aS) split into packed hi and lo parts.
__device__ __forceinline__ void sx_split8(const float* aV, float aS, sx_u32x4& aHi, sx_u32x4& aLo)
{
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		unsigned h, l;
		sx_pair(aV[2 * q] * aS, aV[2 * q + 1] * aS, h, l);
		aHi[q] = h;
		aLo[q] = l;
	}
}

__device__ __forceinline__ sx_f32x4 sx_mfma(sx_u32x4 aA, sx_u32x4 aB, sx_f32x4 aC)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(sx_f16x8, aA), __builtin_bit_cast(sx_f16x8, aB),
		aC, 0, 0, 0);
}

// Split product: lo*hi + hi*lo first (the small terms), then hi*hi.
__device__ __forceinline__ sx_f32x4 sx_product(sx_u32x4 aAh, sx_u32x4 aAl, sx_u32x4 aBh, sx_u32x4 aBl, sx_f32x4 aC)
{
	aC = sx_mfma(aAl, aBh, aC);
	aC = sx_mfma(aAh, aBl, aC);
	return sx_mfma(aAh, aBh, aC);
}

// Largest value of the wave (bit patterns of non-negative floats compare as
// unsigned integers): row shifts, then the row broadcasts; lane 63 holds it.
__device__ __forceinline__ unsigned sx_wave_max(unsigned aValue)
{
	unsigned v = aValue;
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, true)));
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, true)));
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, true)));
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, true)));
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
	v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
	return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// aValue < aLimit ? aIfBelow : aOtherwise for uniform values, as a scalar
// select the optimizer cannot turn into a branch around the caller's store.
__device__ __forceinline__ int sx_select_below(int aValue, int aLimit, int aIfBelow, int aOtherwise)
{
	int r;
	asm("s_cmp_lt_i32 %1, %2\n\ts_cselect_b32 %0, %3, %4"
	    : "=s"(r) : "s"(aValue), "s"(aLimit), "s"(aIfBelow), "s"(aOtherwise) : "scc");
	return r;
}

// The same for an unsigned comparison (a negative value as unsigned is past
// any limit below 2^31).
__device__ __forceinline__ int sx_select_below_u(unsigned aValue, unsigned aLimit, int aIfBelow, int aOtherwise)
{
	int r;
	asm("s_cmp_lt_u32 %1, %2\n\ts_cselect_b32 %0, %3, %4"
	    : "=s"(r) : "s"(aValue), "s"(aLimit), "s"(aIfBelow), "s"(aOtherwise) : "scc");
	return r;
}

// Uniform integer -> scalar register.
__device__ __forceinline__ int sx_uniform(int aValue)
{
	return __builtin_amdgcn_readfirstlane(aValue);
}

// Exponent E with |v| < 2^E for the largest |v| (0 for 0).
__device__ __forceinline__ int sx_exponent(float aValue)
{
	int e = 0;
	(void)frexpf(aValue, &e);
	return aValue == 0.0f ? 0 : e;
}

// Z: where the z filter runs. 0: split FP16 MFMA windows over the per-wave LDS
// ring (above). 1: a register ring of 16 y-filtered planes rescaled to their
// true values, z taps as FP32 FMAs on the vector ALUs; no LDS, so W waves per
// SIMD can share a CU.
template<int K, int Z, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void sep3_split_kernel(sep_args a)
{
	constexpr int C = K / 2;
	constexpr int P = 4;  // planes of loads in flight (divides the 16-plane block)
	__shared__ unsigned lds[Z == 0 ? 4 * kSxWaveDwords : 1];

	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const unsigned tile_x = block % a.tiles_x;
	const unsigned rest = block / a.tiles_x;
	const unsigned tile_y = rest % a.tiles_y;
	const unsigned chunk = rest / a.tiles_y;
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const int g = lane >> 4;
	const int n = lane & 15;
	const int x0 = static_cast<int>(tile_x) * kSxTileX + 16 * wave;
	const int y0 = static_cast<int>(tile_y) * kSxTileY;
	const int z0 = static_cast<int>(chunk) * a.chunk;
	const int zend = min(z0 + a.chunk, a.nz);
	unsigned* ring_hi = lds + wave * kSxWaveDwords;
	unsigned* ring_lo = ring_hi + kSxPartDwords;
	// plane exponents: plane i's in lane i mod 64 (a lane select writes it,
	// ds_bpermute reads it across lanes; no branch splits the unrolled body)
	int sigv = 14;

	// tap scales per axis: largest tap -> [2^14, 2^15), and 2^s > sum |taps|
	int b[3], s[3];
	int all[K];
#pragma unroll
	for (int t = 0; t < K; ++t) {
		all[t] = t;
	}
#pragma unroll
	for (int axis = 0; axis < 3; ++axis) {
		float w[K];
		sep_taps<K>(a, axis, all, w);
		float wmax = 0.0f, wsum = 0.0f;
#pragma unroll
		for (int t = 0; t < K; ++t) {
			wmax = fmaxf(wmax, fabsf(w[t]));
			wsum += fabsf(w[t]);
		}
		b[axis] = sx_uniform(wmax == 0.0f ? 0 : 15 - sx_exponent(wmax));
		s[axis] = sx_uniform(sx_exponent(wsum));
	}
	// the taps at N window positions of one axis, scaled by 2^b
	const auto taps_at = [&](int aAxis, const auto& aAt, auto& aOut) {
		sep_taps<K>(a, aAxis, aAt, aOut);
#pragma unroll
		for (auto& v : aOut) {
			v = ldexpf(v, b[aAxis]);
		}
	};
	const float c1 = __int_as_float(sx_uniform(__float_as_int(ldexpf(1.0f, -b[0] - s[0]))));
	const float c2 = __int_as_float(sx_uniform(__float_as_int(ldexpf(1.0f, -b[1] - s[1]))));
	const int rescale = b[2] - s[0] - s[1] + 141;  // output = D * 2^(E - rescale)

	// window position of operand element j for lane group g (x columns, y rows)
	const auto window = [&](int aJ) { return 16 * (aJ >> 2) + 4 * g + (aJ & 3); };
	sx_u32x4 bxh, bxl, ayh, ayl;
	float wz[8];  // z taps of this lane's A elements (row n, planes 8g + j of a window), scaled
	{
		float xv[8], yv[8];
		int at[8], zat[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			at[j] = n + 8 + C - window(j);
			zat[j] = n - 8 * g - j + 2 * C;
		}
		taps_at(0, at, xv);
		taps_at(1, at, yv);
		taps_at(2, zat, wz);
		sx_split8(xv, 1.0f, bxh, bxl);
		sx_split8(yv, 1.0f, ayh, ayl);
	}

	// the lane's four 16-byte groups per plane: row tile mt, half h
	unsigned offset[2][2];
#pragma unroll
	for (int mt = 0; mt < 2; ++mt) {
		const int yi = y0 - 8 + 16 * mt + n;
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const int xi = x0 - 8 + 16 * h + 4 * g;
			const bool valid = yi >= 0 && yi < a.ny && xi >= 0 && xi + 3 < a.nx;
			offset[mt][h] = valid ? static_cast<unsigned>(xi * a.sx + yi * a.sy) : static_cast<unsigned>(a.src_span);
		}
	}
	// output: one resource over the chunk's planes; the lane part of the offset
	// carries x and the lane group's plane, the scalar part y and the window
	const long long dst_bytes = static_cast<long long>(zend - z0 - 1) * a.dz + a.dst_span;
	const auto out_res = plane_resource(a.dst + static_cast<long long>(z0) * a.dz, static_cast<int>(dst_bytes));
	const unsigned out_lane = (x0 + n < a.nx) ? static_cast<unsigned>((x0 + n) * a.dx + 4 * g * a.dz)
	                                          : static_cast<unsigned>(dst_bytes);

	const int inputs = (zend - z0) + 2 * C;  // input planes z0 - C .. zend - 1 + C
	const int windows = (zend - z0 + 15) / 16;
	sx_f32x4 pre[P][2][2];
	auto fetch = [&](int aPlane, int aSlot) {
		const int zi = z0 - C + aPlane;
		const bool in = aPlane < inputs && zi >= 0 && zi < a.nz;
		const auto plane = plane_resource(a.src + (in ? zi : 0) * a.sz, in ? a.src_span : 0);
#pragma unroll
		for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				pre[aSlot][mt][h] = __builtin_bit_cast(sx_f32x4,
					__builtin_amdgcn_raw_buffer_load_b128(plane, static_cast<int>(offset[mt][h]), 0, 0));
			}
		}
	};
	const sx_f32x4 zero4 = sx_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	// x filter of the plane in slot aSlot (writes dx); records the plane's exponent
	auto x_stage = [&](int aPlane, int aSlot, sx_f32x4* aDx) {
		float m = 0.0f;
#pragma unroll
		for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
#pragma unroll
				for (int e = 0; e < 4; ++e) {
					m = fmaxf(m, fabsf(pre[aSlot][mt][h][e]));
				}
			}
		}
		const int e = sx_uniform(min(max(static_cast<int>(sx_wave_max(__float_as_uint(m)) >> 23), 14), 254));
		const float sc = __uint_as_float(static_cast<unsigned>(268 - e) << 23);
		sx_u32x4 ah[2], al[2];
#pragma unroll
		for (int mt = 0; mt < 2; ++mt) {
			float v[8];
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				v[j] = pre[aSlot][mt][j >> 2][j & 3];
			}
			sx_split8(v, sc, ah[mt], al[mt]);
		}
		// this slot's operands are consumed: load plane aPlane + P into it
		fetch(aPlane + P, aSlot);
#pragma unroll
		for (int mt = 0; mt < 2; ++mt) {
			aDx[mt] = sx_product(ah[mt], al[mt], bxh, bxl, zero4);
		}
		sigv = lane == (aPlane & 63) ? e : sigv;
	};
	auto y_stage = [&](const sx_f32x4* aDx) {
		sx_u32x4 bh, bl;
		float v[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			v[j] = aDx[j >> 2][j & 3];
		}
		sx_split8(v, c1, bh, bl);
		return sx_product(ayh, ayl, bh, bl, zero4);
	};
	// ring slots of planes aPlane - 1 (even) and aPlane: one packed pair per column
	auto z_stage = [&](int aPlane, const sx_f32x4& aEven, const sx_f32x4& aOdd) {
		const int pair = ((aPlane - 1) & 31) >> 1;
#pragma unroll
		for (int jj = 0; jj < 4; ++jj) {
			unsigned hp, lp;
			sx_pair(aEven[jj] * c2, aOdd[jj] * c2, hp, lp);
			const int col = 16 * (4 * g + jj) + n;
			ring_hi[col * kSxColDwords + pair] = hp;
			ring_lo[col * kSxColDwords + pair] = lp;
		}
	};
	// z filter of window aWin: ring planes 16 aWin .. 16 aWin + 31 -> output
	// planes 16 aWin .. 16 aWin + 15 of the chunk. A row n (output plane
	// 16 aWin + n) sums ring planes 16 aWin + n .. 16 aWin + n + 2C; its taps
	// for plane i are multiplied by 2^(e_i - E_n), E_n the largest exponent of
	// those planes, so the row's products share the scale 2^(141 - E_n).
	auto z_window = [&](int aWin) {
		const auto plane_exponent = [&](int aPlane) { return __builtin_amdgcn_ds_bpermute(4 * (aPlane & 63), sigv); };
		int erow = 14;
#pragma unroll
		for (int t = 0; t <= 2 * C; ++t) {
			erow = max(erow, plane_exponent(16 * aWin + n + t));
		}
		sx_u32x4 azh, azl;
		float w[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			w[j] = ldexpf(wz[j], plane_exponent(16 * aWin + 8 * g + j) - erow);
		}
		sx_split8(w, 1.0f, azh, azl);
		// this lane's outputs are rows 4g .. 4g + 3: their exponents from the
		// lanes 4g + jj, which hold those A rows
		int shift[4];
#pragma unroll
		for (int jj = 0; jj < 4; ++jj) {
			shift[jj] = __builtin_amdgcn_ds_bpermute(4 * (4 * g + jj), erow) - rescale;
		}
		const int pair = ((16 * aWin + 8 * g) & 31) >> 1;
#pragma unroll
		for (int c = 0; c < 16; ++c) {
			const int at = (16 * c + n) * kSxColDwords + pair;
			const uint2 h01 = *reinterpret_cast<const uint2*>(ring_hi + at);
			const uint2 h23 = *reinterpret_cast<const uint2*>(ring_hi + at + 2);
			const uint2 l01 = *reinterpret_cast<const uint2*>(ring_lo + at);
			const uint2 l23 = *reinterpret_cast<const uint2*>(ring_lo + at + 2);
			const sx_u32x4 bh = sx_u32x4{h01.x, h01.y, h23.x, h23.y};
			const sx_u32x4 bl = sx_u32x4{l01.x, l01.y, l23.x, l23.y};
			const sx_f32x4 d = sx_product(azh, azl, bh, bl, zero4);
#pragma unroll
			for (int jj = 0; jj < 4; ++jj) {
				// rows past the image go past the resource's end; lane part, row part
				// and chunk stay below 2^30 each, so the sum does not wrap
				const int soff = sx_select_below(y0 + c, a.ny,
					(y0 + c) * static_cast<int>(a.dy) + (16 * aWin + jj) * static_cast<int>(a.dz),
					static_cast<int>(dst_bytes));
				__builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ldexpf(d[jj], shift[jj])), out_res,
					static_cast<int>(out_lane), soff, 0);
			}
		}
	};

#pragma unroll
	for (int q = 0; q < P; ++q) {
		fetch(q, q);
	}
	// pipeline: iteration p runs z_stage of plane p (odd p: the pair p - 1, p),
	// the y filter of plane p + 1 and the x filter of plane p + 2
	sx_f32x4 dx[2];
	x_stage(0, 0, dx);
	sx_f32x4 dy = y_stage(dx);
	x_stage(1, 1, dx);
	sx_f32x4 dy_even = zero4;
	// register ring (Z == 1): plane p's y output at its true scale in slot p mod 16
	float ring[16][4];
	float kz[K];
	unsigned ring_lane[4];
	if constexpr (Z == 1) {
#pragma unroll
		for (int j = 0; j < 16; ++j) {
#pragma unroll
			for (int jj = 0; jj < 4; ++jj) {
				ring[j][jj] = 0.0f;
			}
		}
		{
			int all[K];
#pragma unroll
			for (int t = 0; t < K; ++t) {
				all[t] = t;
			}
			sep_taps<K>(a, 2, all, kz);
		}
#pragma unroll
		for (int t = 0; t < K; ++t) {
			kz[t] = __int_as_float(sx_uniform(__float_as_int(kz[t])));
		}
		// this lane's outputs: rows y0 + 4g + jj at column x0 + n
#pragma unroll
		for (int jj = 0; jj < 4; ++jj) {
			const bool ok = x0 + n < a.nx && y0 + 4 * g + jj < a.ny;
			ring_lane[jj] = ok ? static_cast<unsigned>((x0 + n) * a.dx + (4 * g + jj) * a.dy) : static_cast<unsigned>(dst_bytes);
		}
	}
	// y output of plane aPlane -> true scale (2^(e - 141 + s_x - b_y)), into the
	// ring; the z filter of output plane aPlane - 2C (newest plane with tap 0)
	auto ring_stage = [&](int aPlane, int aSlot, const sx_f32x4& aDy) {
		const int e = __builtin_amdgcn_readlane(sigv, aPlane & 63);
		const int shift = e - 141 + s[0] - b[1];
#pragma unroll
		for (int jj = 0; jj < 4; ++jj) {
			ring[aSlot][jj] = ldexpf(aDy[jj], shift);
		}
		float zout[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
		for (int t = K - 1; t >= 0; --t) {
#pragma unroll
			for (int jj = 0; jj < 4; ++jj) {
				zout[jj] = fmaf(ring[(aSlot - t + 16) % 16][jj], kz[t], zout[jj]);
			}
		}
		const int o = aPlane - 2 * C;  // output plane of the chunk
		const int soff = sx_select_below_u(static_cast<unsigned>(o), static_cast<unsigned>(zend - z0),
			o * static_cast<int>(a.dz) + y0 * static_cast<int>(a.dy), static_cast<int>(dst_bytes));
#pragma unroll
		for (int jj = 0; jj < 4; ++jj) {
			__builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(zout[jj]), out_res, static_cast<int>(ring_lane[jj]),
				soff, 0);
		}
	};
	for (int k = 0; k <= windows; ++k) {
#pragma unroll
		for (int q = 0; q < 16; ++q) {
			const int p = 16 * k + q;
			sx_f32x4 dx_next[2];
			x_stage(p + 2, (q + 2) % P, dx_next);
			const sx_f32x4 dy_next = y_stage(dx);
			if constexpr (Z == 0) {
				if (q & 1) {
					z_stage(p, dy_even, dy);
				} else {
					dy_even = dy;
				}
			} else {
				ring_stage(p, q, dy);
			}
			dy = dy_next;
			dx[0] = dx_next[0];
			dx[1] = dx_next[1];
		}
		if constexpr (Z == 0) {
			if (k > 0) {
				z_window(k - 1);
			}
		}
	}
}

template<int K>
void launch_split_k(sep_args& a, hipStream_t aStream)
{
	a.tiles_x = ceil_div(a.nx, kSxTileX);
	a.tiles_y = ceil_div(a.ny, kSxTileY);
	const long long tiles = static_cast<long long>(a.tiles_x) * a.tiles_y;
	// z on the register ring by default (53 vs 57 us on the 15^3 headline: the
	// LDS windows' store bursts and single wave per SIMD cost more than the z FMAs)
	const int zmode = tune_get("sep3.split_z", 1) == 0 ? 0 : 1;
	const int waves = zmode == 1 && tune_get("sep3.waves", 1) == 2 ? 2 : 1;
	int chunk = tune_get("sep3.chunk", 0);
	if (chunk <= 0) {
		// one workgroup per CU and wave slot (the LDS ring takes 148 KB): long chunks
		const long long chunks = std::max<long long>(1, (256 * waves + tiles - 1) / tiles);
		chunk = static_cast<int>(ceil_div(a.nz, chunks));
	}
	chunk = std::max(16, (chunk + 15) / 16 * 16);  // whole 16-plane windows
	while (chunk > 16 && !split_chunk_fits(a, chunk)) {
		chunk -= 16;
	}
	a.chunk = chunk;
	a.chunks_z = ceil_div(a.nz, chunk);
	const unsigned nb = static_cast<unsigned>(tiles * a.chunks_z);
	if (zmode == 0) {
		sep3_split_kernel<K, 0, 1><<<nb, 256, 0, aStream>>>(a);
	} else if (waves == 2) {
		sep3_split_kernel<K, 1, 2><<<nb, 256, 0, aStream>>>(a);
	} else {
		sep3_split_kernel<K, 1, 1><<<nb, 256, 0, aStream>>>(a);
	}
}

} // namespace

// Output planes of a split-engine chunk are addressed with 32-bit offsets from
// the chunk's first plane (lane and scalar parts each below 2^30).
bool split_chunk_fits(const sep_args& a, int aChunk)
{
	return static_cast<long long>(aChunk) * a.dz + a.dst_span < (1LL << 30);
}

void launch_split(sep_args& a, int aTaps, hipStream_t aStream)
{
	switch (aTaps) {
	case 1: launch_split_k<1>(a, aStream); break;
	case 3: launch_split_k<3>(a, aStream); break;
	case 5: launch_split_k<5>(a, aStream); break;
	case 7: launch_split_k<7>(a, aStream); break;
	case 9: launch_split_k<9>(a, aStream); break;
	case 11: launch_split_k<11>(a, aStream); break;
	case 13: launch_split_k<13>(a, aStream); break;
	default: launch_split_k<15>(a, aStream); break;
	}
}

} // namespace separable
} // namespace iane_hip
