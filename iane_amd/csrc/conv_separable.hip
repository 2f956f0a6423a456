// Separable convolution behind iane_hip_convolve_separable: the kernel is the
// outer product of one odd-length tap vector per axis. With either border rule
// this equals the reference convolve with the dense outer-product kernel
// (include/iane/algorithms/convolution.hpp:23-117): the clamp rule clamps each
// axis independently and the zero rule zeroes a sample when any axis is out of
// range, and both factor over axes.
//
// Fast path (rank 2 or 3, F32 images, FP32 or SPLIT_FP16 arithmetic, odd tap
// lengths up to 15, each centred in the window K of the longest; a rank-2
// image is one plane with the unit kernel along z): one fused kernel. A workgroup owns a 64 x 16 (x, y) output tile and streams a
// chunk of z planes. For every input plane it stages the plane region with a
// K-1 halo in x and y into LDS, filters it along x and y, and pushes the
// filtered plane into a per-thread register ring of K planes from which the z
// filter produces one output plane per input plane. Input is read once per
// chunk (plus the z halo), output written once. Three engines share this plan:
// sep3_direct_kernel (default for zero borders and 4-aligned rows) and
// sep3_mfma_kernel run the x and y filters as FP32 Toeplitz products on the
// matrix cores, the first with per-wave operand loads and no LDS, the second
// with the plane staged in LDS; sep3_fused_kernel uses the vector ALUs only.
//
// Engine 3 (split-FP16 operands, every filter on the matrix cores) is in
// conv_separable_split.hip.
//
// General path: one 1-D pass per axis through temporaries in the caller's workspace.
#include "separable.hpp"

namespace iane_hip {

void validate_convolution(const view4& aSrc, const view4& aKer, const view4& aDst);

namespace {

using separable::plane_resource;
using separable::sep_args;

constexpr int kSepTileX = 64;

__device__ __forceinline__ bool clamp_or_zero(int& aCoord, int aExtent, int aBorder)
{
	if (aCoord >= 0 && aCoord < aExtent) {
		return true;
	}
	if (aBorder == IANE_BORDER_ZERO) {
		return false;
	}
	aCoord = aCoord < 0 ? 0 : aExtent - 1;
	return true;
}

// Uniform value -> scalar register (the taps are the same for every lane).
__device__ __forceinline__ float uniform_float(float aValue)
{
	return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(aValue)));
}

template<int K, int kSepThreads, int kSepTileY>
__global__ __launch_bounds__(kSepThreads) void sep3_fused_kernel(sep_args a)
{
	constexpr int kSepColsPerThread = kSepTileX * kSepTileY / kSepThreads; // consecutive x columns per thread
	constexpr int kRowThreads = kSepTileX / kSepColsPerThread;
	constexpr int C = K / 2;
	constexpr int cols = kSepTileX + K - 1;   // staged x extent
	constexpr int rows = kSepTileY + K - 1;   // staged y extent
	constexpr int pitch = cols + 1;
	constexpr int per_thread = (rows * cols + kSepThreads - 1) / kSepThreads;
	__shared__ float in_s[rows * pitch];
	__shared__ float yr_s[kSepTileY * pitch]; // y-filtered, still x-halo wide

	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const unsigned tile_x = block % a.tiles_x;
	const unsigned rest = block / a.tiles_x;
	const unsigned tile_y = rest % a.tiles_y;
	const unsigned chunk = rest / a.tiles_y;
	const int x0 = static_cast<int>(tile_x) * kSepTileX;
	const int y0 = static_cast<int>(tile_y) * kSepTileY;
	const int z0 = static_cast<int>(chunk) * a.chunk;
	const int zend = min(z0 + a.chunk, a.nz);
	const int tid = threadIdx.x;

	float kx[K], ky[K], kz[K];
	{
		int at[K];
#pragma unroll
		for (int i = 0; i < K; ++i) {
			at[i] = i;
		}
		separable::sep_taps<K>(a, 0, at, kx);
		separable::sep_taps<K>(a, 1, at, ky);
		separable::sep_taps<K>(a, 2, at, kz);
#pragma unroll
		for (int i = 0; i < K; ++i) {
			kx[i] = uniform_float(kx[i]);
			ky[i] = uniform_float(ky[i]);
			kz[i] = uniform_float(kz[i]);
		}
	}

	// Staging offsets of this thread's elements (same for every plane).
	int offset[per_thread]; // byte offset within a plane (host checks it fits in int)
	bool valid[per_thread];
#pragma unroll
	for (int e = 0; e < per_thread; ++e) {
		const int i = tid + e * kSepThreads;
		const int ly = i / cols;
		const int lx = i - ly * cols;
		int xi = x0 - C + lx;
		int yi = y0 - C + ly;
		valid[e] = i < rows * cols && clamp_or_zero(xi, a.nx, a.border) && clamp_or_zero(yi, a.ny, a.border);
		offset[e] = valid[e] ? static_cast<int>(xi * a.sx + yi * a.sy) : 0;
	}

	// Thread ownership for the x filter and the z ring: row y = tid / kRowThreads,
	// columns x = (tid % kRowThreads) * kSepColsPerThread + r (a sliding window along x).
	const int ox = (tid % kRowThreads) * kSepColsPerThread;
	const int oy = tid / kRowThreads;

	float ring[K][kSepColsPerThread];
#pragma unroll
	for (int j = 0; j < K; ++j) {
#pragma unroll
		for (int r = 0; r < kSepColsPerThread; ++r) {
			ring[j][r] = 0.0f;
		}
	}

	const int planes = (zend - z0) + K - 1; // input planes z0 - C .. zend - 1 + C
	// Register prefetch P planes deep: the loads of planes p+1 .. p+P are in
	// flight while plane p is filtered (one plane of filtering is shorter than
	// the memory latency under load). Slot of plane p = p mod P, a compile-time
	// index because the plane loop is unrolled by K and P divides K.
	constexpr int P = (K % 3 == 0) ? 3 : 1;
	float pending[kSepColsPerThread];
	int pending_z = -1;
	auto flush = [&]() {
		const int py = y0 + oy;
		if (pending_z >= 0 && py < a.ny) {
#pragma unroll
			for (int r = 0; r < kSepColsPerThread; ++r) {
				const int px = x0 + ox + r;
				if (px < a.nx) {
					*reinterpret_cast<float*>(a.dst + px * a.dx + py * a.dy + pending_z * a.dz) = pending[r];
				}
			}
		}
		pending_z = -1;
	};
	float pre[P][per_thread];
	bool pre_in[P];
	auto fetch = [&](int aPlane, int aSlot) {
		int zi = z0 - C + aPlane;
		const bool in = clamp_or_zero(zi, a.nz, a.border);
		// loads are unconditional (invalid lanes read plane element 0) and
		// selected afterwards, so no load sits behind a branch
		// (the fused path takes f32 sources only: a per-element type switch
		// would put each load behind a branch and serialise them)
		// The raw values stay in `pre` and are masked only when written to LDS,
		// so nothing consumes them until their plane is staged.
		const char* base = a.src + (in ? zi : 0) * a.sz;
#pragma unroll
		for (int e = 0; e < per_thread; ++e) {
			pre[aSlot][e] = *reinterpret_cast<const float*>(base + offset[e]);
		}
		pre_in[aSlot] = in;
	};
#pragma unroll
	for (int q = 0; q < P; ++q) {
		if (q < planes) {
			fetch(q, q);
		}
	}
	for (int base = 0; base < planes; base += K) {
#pragma unroll
		for (int j = 0; j < K; ++j) {
			const int p = base + j;
			if (p >= planes) {
				break;
			}
			const int slot = j % P;
			int zi = z0 - C + p;
			const bool plane_in = clamp_or_zero(zi, a.nz, a.border);
			float xy[kSepColsPerThread];
			if (plane_in) {
#pragma unroll
				for (int e = 0; e < per_thread; ++e) {
					const int i = tid + e * kSepThreads;
					if (i < rows * cols) {
						const int ly = i / cols;
						in_s[ly * pitch + (i - ly * cols)] = (pre_in[slot] && valid[e]) ? pre[slot][e] : 0.0f;
					}
				}
			}
			if (p + P < planes) {
				fetch(p + P, slot);
			}
			// the previous plane's outputs go out now and complete during this plane's filters
			flush();
			if (plane_in) {
				lds_barrier(); // staged plane visible; previous x filter done with yr_s
				// y filter over the halo-wide region: kSepTileY x cols outputs, 4 rows per item.
				for (int i = tid; i < (kSepTileY / 4) * cols; i += kSepThreads) {
					const int lx = i % cols;
					const int gy = (i / cols) * 4;
					float w[K + 3];
#pragma unroll
					for (int t = 0; t < K + 3; ++t) {
						w[t] = in_s[(gy + t) * pitch + lx];
					}
#pragma unroll
					for (int r = 0; r < 4; ++r) {
						float s = 0.0f;
#pragma unroll
						for (int q = 0; q < K; ++q) {
							s = fmaf(w[r + K - 1 - q], ky[q], s);
						}
						yr_s[(gy + r) * pitch + lx] = s;
					}
				}
				lds_barrier(); // yr_s complete; in_s free for the next plane
				{
					const float* row = yr_s + oy * pitch + ox;
					float w[kSepColsPerThread + K - 1];
#pragma unroll
					for (int t = 0; t < kSepColsPerThread + K - 1; ++t) {
						w[t] = row[t];
					}
#pragma unroll
					for (int r = 0; r < kSepColsPerThread; ++r) {
						float s = 0.0f;
#pragma unroll
						for (int q = 0; q < K; ++q) {
							s = fmaf(w[r + K - 1 - q], kx[q], s);
						}
						xy[r] = s;
					}
				}
			} else {
#pragma unroll
				for (int r = 0; r < kSepColsPerThread; ++r) {
					xy[r] = 0.0f;
				}
			}
			// push into the ring: slot (p mod K) == j because base is a multiple of K
#pragma unroll
			for (int r = 0; r < kSepColsPerThread; ++r) {
				ring[j][r] = xy[r];
			}
			if (p >= K - 1) {
				const int zo = z0 + p - (K - 1);
				float out[kSepColsPerThread];
#pragma unroll
				for (int r = 0; r < kSepColsPerThread; ++r) {
					out[r] = 0.0f;
				}
#pragma unroll
				for (int q = 0; q < K; ++q) {
					const int slot = ((j - q) % K + K) % K;
#pragma unroll
					for (int r = 0; r < kSepColsPerThread; ++r) {
						out[r] = fmaf(ring[slot][r], kz[q], out[r]);
					}
				}
#pragma unroll
				for (int r = 0; r < kSepColsPerThread; ++r) {
					pending[r] = out[r];
				}
				pending_z = zo;
			}
		}
	}
	flush();
}

// ---------------------------------------------------------------------------
// Matrix-core variant of the fused kernel (FP32 MFMA, full FP32 products).
//
// The x and y filters of a 64 x 16 plane tile are Toeplitz products on
// v_mfma_f32_16x16x4_f32. Wave w owns output columns 16w .. 16w+15:
//   x pass  T (32 staged rows x 16 columns) = S (rows x 32-wide window) * X,
//           X[k][n] = kx[n + 2C - k]: two row tiles, 8 K-steps each;
//   y pass  U (16 rows x 16 columns) = Y * T, K = the 32 staged rows taken in
//           the order in which the x pass leaves them in the accumulators
//           (K-step s, lane group g <-> staged row 16 (s / 4) + 4 g + s % 4),
//           so T feeds the y pass from registers without an LDS round trip;
//           Y[m][row] = ky[m + 2C - row] in that permuted order.
// The z filter stays a register ring on the vector ALUs (4 outputs per lane).
// The staged plane is double-buffered in LDS: one barrier per plane.
// MFMA lane layout (16x16x4 f32): lane l holds A[l % 16][l / 16],
// B[l / 16][l % 16] and D[4 (l / 16) + i][l % 16], i = 0..3.
constexpr int kMxTileX = 64;
constexpr int kMxTileY = 16;
constexpr int kMxRows = 32;   // staged rows: 16 + K - 1 <= 30, padded to two 16-row MFMA tiles
constexpr int kMxCols = 80;   // staged columns: the last tile's 32-wide window ends at column 79
constexpr int kMxPitch = 84;  // 16-byte aligned rows; 84 = 20 mod 32 and mod 64: a wave's 128-bit reads of 16 rows hit disjoint banks
constexpr int kMxPerThread = kMxRows * kMxCols / 256;

// LDS position of staged column c: columns are permuted inside blocks of 16 so
// that the 4 A operands a lane needs from one block (columns 16b + 4a + g,
// a = 0..3) are adjacent and load with one 128-bit read.
__device__ __forceinline__ int mx_lds_column(int aCol)
{
	return (aCol & ~15) + 4 * (aCol & 3) + ((aCol >> 2) & 3);
}

typedef float mx_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mx_f32x4 mfma_f32_16x16x4(float a, float b, mx_f32x4 c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// W: waves per SIMD the register allocation is fitted to (2 or 3).
template<int K, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void sep3_mfma_kernel(sep_args a)
{
	constexpr int C = K / 2;
	constexpr int P = (K % 3 == 0) ? 3 : 1; // prefetch depth in planes (P divides K: compile-time slots)
	__shared__ float stage[2][kMxRows * kMxPitch];

	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const unsigned tile_x = block % a.tiles_x;
	const unsigned rest = block / a.tiles_x;
	const unsigned tile_y = rest % a.tiles_y;
	const unsigned chunk = rest / a.tiles_y;
	const int x0 = static_cast<int>(tile_x) * kMxTileX;
	const int y0 = static_cast<int>(tile_y) * kMxTileY;
	const int z0 = static_cast<int>(chunk) * a.chunk;
	const int zend = min(z0 + a.chunk, a.nz);
	const int tid = threadIdx.x;
	const int lane = tid & 63;
	const int wave = tid >> 6;
	const int g = lane >> 4;
	const int n = lane & 15;

	// MFMA operands built from the taps (per lane, constant over the run)
	float bx[8], ay[8], kz[K];
	{
		int xat[8], yat[8], zat[K];
#pragma unroll
		for (int s = 0; s < 8; ++s) {
			xat[s] = n + 2 * C - (4 * s + g);
			yat[s] = n + 2 * C - (16 * (s / 4) + 4 * g + s % 4);
		}
#pragma unroll
		for (int i = 0; i < K; ++i) {
			zat[i] = i;
		}
		separable::sep_taps<K>(a, 0, xat, bx);
		separable::sep_taps<K>(a, 1, yat, ay);
		separable::sep_taps<K>(a, 2, zat, kz);
#pragma unroll
		for (int i = 0; i < K; ++i) {
			kz[i] = uniform_float(kz[i]);
		}
	}

	// The plane loop below has no branch but its exit: every condition is an
	// address instead. Loads and stores go through a per-plane buffer resource
	// (scalar registers) and fixed 32-bit lane offsets; an element outside the
	// image (zero border) or a padding element gets an offset past the plane
	// and reads 0, a plane outside the volume (zero border) or past the chunk
	// gets an empty resource, and a store that must not happen goes to an
	// empty resource. (A branchy loop makes the compiler wait for all memory
	// operations at every join, which serialises the prefetch.)
	unsigned offset[kMxPerThread];
#pragma unroll
	for (int e = 0; e < kMxPerThread; ++e) {
		const int i = tid + e * 256;
		const int row = i / kMxCols;
		const int col = i - row * kMxCols;
		int xi = x0 - C + col;
		int yi = y0 - C + row;
		const bool valid = row < kMxTileY + K - 1 && col < kMxTileX + K - 1 && clamp_or_zero(xi, a.nx, a.border) &&
			clamp_or_zero(yi, a.ny, a.border);
		offset[e] = valid ? static_cast<unsigned>(xi * a.sx + yi * a.sy) : static_cast<unsigned>(a.src_span);
	}
	int lds_index[kMxPerThread];
#pragma unroll
	for (int e = 0; e < kMxPerThread; ++e) {
		const int i = tid + e * 256;
		const int row = i / kMxCols;
		lds_index[e] = row * kMxPitch + mx_lds_column(i - row * kMxCols);
	}
	const int ox = x0 + 16 * wave + n;  // output column of this lane
	const int oy = y0 + 4 * g;          // first of its 4 output rows
	unsigned store_offset[4];
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		store_offset[r] = (ox < a.nx && oy + r < a.ny) ? static_cast<unsigned>(ox * a.dx + (oy + r) * a.dy)
		                                               : static_cast<unsigned>(a.dst_span);
	}

	const int planes = (zend - z0) + K - 1; // input planes z0 - C .. zend - 1 + C
	const auto source_plane = [&](int aPlane) {
		int zi = z0 - C + aPlane;
		const bool in = aPlane < planes && clamp_or_zero(zi, a.nz, a.border);
		return plane_resource(a.src + (in ? zi : 0) * a.sz, in ? a.src_span : 0);
	};
	float ring[K][4];
#pragma unroll
	for (int j = 0; j < K; ++j) {
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			ring[j][r] = 0.0f;
		}
	}
	float pre[P][kMxPerThread];
	auto fetch = [&](int aPlane, int aSlot) {
		const auto plane = source_plane(aPlane);
#pragma unroll
		for (int e = 0; e < kMxPerThread; ++e) {
			pre[aSlot][e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(plane, static_cast<int>(offset[e]), 0, 0));
		}
	};
	auto stage_plane = [&](int aPlane, int aSlot) {
		float* S = stage[aPlane & 1];
#pragma unroll
		for (int e = 0; e < kMxPerThread; ++e) {
			S[lds_index[e]] = pre[aSlot][e];
		}
	};
	// store the z-filtered output plane z0 + aPlane - (K - 1) (dropped unless in the chunk)
	auto z_store = [&](int aPlane, const float* aOut) {
		const int z = z0 + aPlane - (K - 1);
		const bool ok = z >= z0 && z < zend;
		// integer selects (not a branch: the z filter above must stay in this block)
		const auto plane = plane_resource(a.dst + (ok ? z : 0) * a.dz, ok ? a.dst_span : 0);
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			// the b32 builtins move raw 32-bit words: bit casts, not conversions
			__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, aOut[r]), plane,
				static_cast<int>(store_offset[r]), 0, 0);
		}
	};
#pragma unroll
	for (int q = 0; q < P; ++q) {
		fetch(q, q);
	}
	stage_plane(0, 0);
	fetch(P, 0);
	// Software pipeline: iteration p runs the MFMA filters of plane p while the
	// vector ALUs finish plane p - 1 (z filter over ring slots, store) and
	// stage plane p + 1. Ring slot of plane q is q mod K.
	// Iteration `planes` filters an empty plane and completes the last output.
	float prev[4] = {0.0f, 0.0f, 0.0f, 0.0f};
	for (int base = 0; base <= planes; base += K) {
#pragma unroll
		for (int j = 0; j < K; ++j) {
			const int p = base + j;
			if (p > planes) {
				break;
			}
			// buffer p & 1 is staged, and every wave is past plane p - 1, whose
			// buffer (p + 1) & 1 is restaged next
			lds_barrier();
			const float* S = stage[p & 1];
			// A operands of both row tiles first (four 128-bit reads), then the MFMA chains
			mx_f32x4 av[2][2];
#pragma unroll
			for (int mt = 0; mt < 2; ++mt) {
				const float* rowp = S + (16 * mt + n) * kMxPitch + 16 * wave + 4 * g;
				av[mt][0] = *reinterpret_cast<const mx_f32x4*>(rowp);       // K-steps 0..3 (block w)
				av[mt][1] = *reinterpret_cast<const mx_f32x4*>(rowp + 16);  // K-steps 4..7 (block w + 1)
			}
			// buffer (p + 1) & 1 was last read in iteration p - 1, which every wave
			// has left: stage plane p + 1 and fetch plane p + 1 + P before the MFMAs
			const int next_slot = (j + 1) % P;
			stage_plane(p + 1, next_slot);
			fetch(p + 1 + P, next_slot);
			// A wave issues in order, so the z filter of plane p - 1 (vector ALU)
			// is written between the MFMAs of plane p: two taps per x-pass step,
			// while the matrix pipe works on the previous MFMA.
			const int zr = (j + K - 1) % K;  // ring slot of plane p - 1
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				ring[zr][r] = prev[r];
			}
			float zout[4] = {0.0f, 0.0f, 0.0f, 0.0f};
			mx_f32x4 t[2];
#pragma unroll
			for (int mt = 0; mt < 2; ++mt) {
				t[mt] = mx_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
			}
#pragma unroll
			for (int s = 0; s < 8; ++s) {
#pragma unroll
				for (int mt = 0; mt < 2; ++mt) {
					t[mt] = mfma_f32_16x16x4(av[mt][s / 4][s % 4], bx[s], t[mt]);
				}
#pragma unroll
				for (int q = 2 * s; q < 2 * s + 2 && q < K; ++q) {
					const int ring_slot = ((zr - q) % K + K) % K;
#pragma unroll
					for (int r = 0; r < 4; ++r) {
						zout[r] = fmaf(ring[ring_slot][r], kz[q], zout[r]);
					}
				}
			}
			// y filter: two accumulation chains (row tiles), summed at the end
			mx_f32x4 u0 = mx_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
			mx_f32x4 u1 = mx_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
			for (int s = 0; s < 4; ++s) {
				u0 = mfma_f32_16x16x4(ay[s], t[0][s], u0);
				u1 = mfma_f32_16x16x4(ay[s + 4], t[1][s], u1);
			}
			z_store(p - 1, zout);
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				prev[r] = u0[r] + u1[r];
			}
		}
	}
}

// ---------------------------------------------------------------------------
// Barrier-free matrix-core variant (engine 2) for zero borders and widths that
// are multiples of 4: no LDS staging. Each wave owns a 16 x 16 output tile of
// every plane of its z chunk and loads its own x-pass operands from memory:
// lane (g, m) of row tile mt needs, for K-step s, window column
// c(s, g) = 16 (s / 4) + 4 g + s % 4 of staged row 16 mt + m (the window starts
// 8 columns and 8 rows before the tile), i.e. two 16-byte groups per row tile,
// which arrive as two aligned buffer_load_dwordx4 (a group lies wholly inside
// or wholly outside the image, and outside reads 0). The Toeplitz operands use
// the same column order, so no sample moves between lanes.
//
// The loop is software-pipelined over planes so that four independent MFMA
// chains interleave: iteration p runs the x filter of plane p (two row tiles)
// and the y filter of plane p - 1 (two partial sums), while the vector ALUs run
// the z filter that completes output plane p - 2 - (K - 1) from the ring.
template<int K, int W, int KZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void sep3_direct_kernel(sep_args a)
{
	constexpr int C = K / 2;     // x and y taps (window of K)
	constexpr int CZ = KZ / 2;   // z taps: K, or 1 for a single plane
	constexpr int R = KZ + 1;                                    // ring slots = iterations per unrolled group
	constexpr int P = (R % 4 == 0) ? 4 : ((R % 3 == 0) ? 3 : 2);  // planes of loads in flight (P divides R)
	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const unsigned tile_x = block % a.tiles_x;
	const unsigned rest = block / a.tiles_x;
	const unsigned tile_y = rest % a.tiles_y;
	const unsigned chunk = rest / a.tiles_y;
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const int g = lane >> 4;
	const int n = lane & 15;
	const int x0 = static_cast<int>(tile_x) * kMxTileX + 16 * wave;  // this wave's 16 output columns
	const int y0 = static_cast<int>(tile_y) * kMxTileY;
	const int z0 = static_cast<int>(chunk) * a.chunk;
	const int zend = min(z0 + a.chunk, a.nz);

	// window position of K-step s for lane group g (columns for x, rows for y)
	const auto window = [&](int aStep) { return 16 * (aStep / 4) + 4 * g + aStep % 4; };
	float bx[8], ay[8], kz[KZ];
	int at[8];
#pragma unroll
	for (int s = 0; s < 8; ++s) {
		at[s] = n + 8 + C - window(s);
	}
	separable::sep_taps<K>(a, 0, at, bx);
	separable::sep_taps<K>(a, 1, at, ay);
	int zat[KZ];
#pragma unroll
	for (int i = 0; i < KZ; ++i) {
		zat[i] = i;
	}
	separable::sep_taps<KZ>(a, 2, zat, kz);
#pragma unroll
	for (int i = 0; i < KZ; ++i) {
		kz[i] = uniform_float(kz[i]);
	}
	// the lane's four 16-byte groups per plane: row tile mt, half h (columns 16 h + 4 g .. + 3)
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
	const int ox = x0 + n;
	const int oy = y0 + 4 * g;
	unsigned store_offset[4];
#pragma unroll
	for (int r = 0; r < 4; ++r) {
		store_offset[r] = (ox < a.nx && oy + r < a.ny) ? static_cast<unsigned>(ox * a.dx + (oy + r) * a.dy)
		                                               : static_cast<unsigned>(a.dst_span);
	}

	const int planes = (zend - z0) + KZ - 1;  // input planes z0 - CZ .. zend - 1 + CZ
	mx_f32x4 pre[P][2][2];
	auto fetch = [&](int aPlane, int aSlot) {
		const int zi = z0 - CZ + aPlane;
		const bool in = aPlane < planes && zi >= 0 && zi < a.nz;
		const auto plane = plane_resource(a.src + (in ? zi : 0) * a.sz, in ? a.src_span : 0);
#pragma unroll
		for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				pre[aSlot][mt][h] = __builtin_bit_cast(mx_f32x4,
					__builtin_amdgcn_raw_buffer_load_b128(plane, static_cast<int>(offset[mt][h]), 0, 0));
			}
		}
	};
	auto z_store = [&](int aPlane, const float* aOut) {
		const int z = z0 + aPlane - (KZ - 1);
		const bool ok = z >= z0 && z < zend;
		const auto plane = plane_resource(a.dst + (ok ? z : 0) * a.dz, ok ? a.dst_span : 0);
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, aOut[r]), plane,
				static_cast<int>(store_offset[r]), 0, 0);
		}
	};
	float ring[R][4];  // y-filtered planes by plane index mod R (KZ of them feed an output)
#pragma unroll
	for (int j = 0; j < R; ++j) {
#pragma unroll
		for (int r = 0; r < 4; ++r) {
			ring[j][r] = 0.0f;
		}
	}
#pragma unroll
	for (int q = 0; q < P; ++q) {
		fetch(q, q);
	}
	const mx_f32x4 zero4 = mx_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
	mx_f32x4 t_prev[2] = {zero4, zero4};  // x filter of plane p - 1
	float u_prev[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // y filter of plane p - 2
	// iteration p: x of plane p, y of plane p - 1, z completing plane p - 2
	// Whole groups of R iterations: the group is fully unrolled (prefetch and
	// ring slots are compile-time indices), which an early exit inside it would
	// prevent. Iterations past planes + 1 load empty planes and store nothing;
	// the host picks z chunks that are multiples of R, for which there are none
	// (a chunk of c planes runs c + R iterations).
	for (int base = 0; base <= planes + 1; base += R) {
#pragma unroll
		for (int j = 0; j < R; ++j) {
			const int p = base + j;
			const int slot = j % P;
			const int zr = (j + R - 2) % R;  // ring slot of plane p - 2
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				ring[zr][r] = u_prev[r];
			}
			mx_f32x4 t[2] = {zero4, zero4};
			mx_f32x4 u0 = zero4;
			mx_f32x4 u1 = zero4;
			float zout[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
			for (int s = 0; s < 8; ++s) {
#pragma unroll
				for (int mt = 0; mt < 2; ++mt) {
					t[mt] = mfma_f32_16x16x4(pre[slot][mt][s / 4][s % 4], bx[s], t[mt]);
				}
				if (s < 4) {
					u0 = mfma_f32_16x16x4(ay[s], t_prev[0][s], u0);
				} else {
					u1 = mfma_f32_16x16x4(ay[s], t_prev[1][s - 4], u1);
				}
#pragma unroll
				for (int q = 2 * s; q < 2 * s + 2 && q < KZ; ++q) {
					const int ring_slot = ((zr - q) % R + R) % R;
#pragma unroll
					for (int r = 0; r < 4; ++r) {
						zout[r] = fmaf(ring[ring_slot][r], kz[q], zout[r]);
					}
				}
			}
			// this slot's operands are consumed: load plane p + P into it
			fetch(p + P, slot);
			z_store(p - 2, zout);
			t_prev[0] = t[0];
			t_prev[1] = t[1];
#pragma unroll
			for (int r = 0; r < 4; ++r) {
				u_prev[r] = u0[r] + u1[r];
			}
		}
	}
}

// One 1-D pass along `axis` of a rank-4-padded view; Acc-typed temporaries in
// between passes, final pass converts to the destination type.
template<class Acc>
__global__ __launch_bounds__(256) void axis_pass_kernel(view4 aIn, view4 aOut, const char* aTaps, int aTapType,
	long long aTapStride, int aTaps_n, int aAxis, int aBorder, long long aCount)
{
	const long long index = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (index >= aCount) {
		return;
	}
	int p[4];
	long long rest = index;
	for (int i = 0; i < 4; ++i) {
		p[i] = static_cast<int>(rest % aOut.extent[i]);
		rest /= aOut.extent[i];
	}
	long long base = 0;
	long long obase = 0;
	for (int i = 0; i < 4; ++i) {
		if (i != aAxis) {
			base += static_cast<long long>(p[i]) * aIn.stride[i];
		}
		obase += static_cast<long long>(p[i]) * aOut.stride[i];
	}
	const int c = aTaps_n / 2;
	const int extent = aIn.extent[aAxis];
	Acc sum = Acc(0);
	for (int q = 0; q < aTaps_n; ++q) {
		int coord = p[aAxis] + c - q;
		Acc sample = Acc(0);
		if (clamp_or_zero(coord, extent, aBorder)) {
			sample = load_as<Acc>(aIn.ptr + base + coord * aIn.stride[aAxis], aIn.type);
		}
		sum = fma(sample, load_as<Acc>(aTaps + q * aTapStride, aTapType), sum);
	}
	store_converted(aOut.ptr + obase, aOut.type, sum);
}

view4 packed_temp(const view4& aLike, int aType, void* aPtr)
{
	view4 v = aLike;
	v.ptr = static_cast<char*>(aPtr);
	v.type = aType;
	long long stride = dtype_size(aType);
	for (int i = 0; i < 4; ++i) {
		v.stride[i] = stride;
		stride *= v.extent[i];
	}
	return v;
}

// Tiles and z chunks of a fused launch; returns the workgroup count.
unsigned plan_tiles(sep_args& a, int aTileX, int aTileY)
{
	a.tiles_x = ceil_div(a.nx, aTileX);
	a.tiles_y = ceil_div(a.ny, aTileY);
	// chunk so that the grid covers the chip at least twice: 256 CUs
	const long long xy = static_cast<long long>(a.tiles_x) * a.tiles_y;
	int chunk = tune_get("sep3.chunk", 0);
	if (chunk <= 0) {
		const long long want = tune_get("sep3.min_blocks", 512);
		chunk = a.nz;
		while (chunk > 32 && xy * ceil_div(a.nz, chunk) < want) {
			chunk = (chunk + 1) / 2;
		}
	}
	a.chunk = chunk;
	a.chunks_z = ceil_div(a.nz, chunk);
	return static_cast<unsigned>(xy * a.chunks_z);
}

template<int K>
void launch_mfma(sep_args& a, hipStream_t aStream)
{
	const unsigned nb = plan_tiles(a, kMxTileX, kMxTileY);
	if (tune_get("sep3.waves", 2) == 3) {
		sep3_mfma_kernel<K, 3><<<nb, 256, 0, aStream>>>(a);
	} else {
		sep3_mfma_kernel<K, 2><<<nb, 256, 0, aStream>>>(a);
	}
}

template<int K, int KZ>
void launch_direct(sep_args& a, hipStream_t aStream)
{
	unsigned nb = plan_tiles(a, kMxTileX, kMxTileY);
	if (tune_get("sep3.chunk", 0) <= 0) {
		// Long z chunks cut the halo (each chunk filters KZ - 1 extra planes in x
		// and y); one workgroup per CU (256 of them, one wave per SIMD) is
		// enough for this kernel. The chunk is a multiple of KZ + 1, so the
		// unrolled groups of KZ + 1 iterations have no idle iterations.
		constexpr int R = KZ + 1;
		const long long tiles = static_cast<long long>(a.tiles_x) * a.tiles_y;
		const long long chunks = std::max<long long>(1, (256 + tiles - 1) / tiles);
		int chunk = static_cast<int>(ceil_div(a.nz, chunks));
		chunk = (chunk + R - 1) / R * R;
		a.chunk = chunk;
		a.chunks_z = ceil_div(a.nz, chunk);
		nb = static_cast<unsigned>(tiles * a.chunks_z);
	}
	if (tune_get("sep3.waves", 1) == 2) {
		sep3_direct_kernel<K, 2, KZ><<<nb, 256, 0, aStream>>>(a);
	} else {
		sep3_direct_kernel<K, 1, KZ><<<nb, 256, 0, aStream>>>(a);
	}
}

// A single-plane image (rank 2) runs with a one-tap z filter.
template<int K>
void launch_direct_rank(sep_args& a, bool aPlanar, hipStream_t aStream)
{
	if (aPlanar) {
		launch_direct<K, 1>(a, aStream);
	} else {
		launch_direct<K, K>(a, aStream);
	}
}

template<int K>
void launch_fused(sep_args& a, hipStream_t aStream, int aThreads)
{
	const int tile_y = tune_get("sep3.tile_y", 16) == 32 ? 32 : 16;
	const unsigned nb = plan_tiles(a, kSepTileX, tile_y);
	if (aThreads == 512) {
		if (tile_y == 32) {
			sep3_fused_kernel<K, 512, 32><<<nb, 512, 0, aStream>>>(a);
		} else {
			sep3_fused_kernel<K, 512, 16><<<nb, 512, 0, aStream>>>(a);
		}
	} else {
		if (tile_y == 32) {
			sep3_fused_kernel<K, 256, 32><<<nb, 256, 0, aStream>>>(a);
		} else {
			sep3_fused_kernel<K, 256, 16><<<nb, 256, 0, aStream>>>(a);
		}
	}
}

} // namespace

// The single-kernel path's tap window: the longest tap vector (all are odd).
int separable_window(const view4* aTaps, int aN)
{
	int k = 1;
	for (int i = 0; i < aN; ++i) {
		k = std::max(k, aTaps[i].extent[0]);
	}
	return k;
}

// The one-kernel path: rank 2 or 3, F32 images, FP32 or SPLIT_FP16 arithmetic,
// tap lengths up to 15, staged offsets within a plane kept in 32-bit registers.
bool separable_fused(const view4& aSrc, const view4* aTaps, int aN, const view4& aDst, int aPrecision)
{
	const bool fused = (aSrc.rank == 3 || aSrc.rank == 2) &&
		(aPrecision == IANE_PREC_FP32 || aPrecision == IANE_PREC_SPLIT_FP16) && aSrc.type == IANE_F32 &&
		aDst.type == IANE_F32 && separable_window(aTaps, aN) <= 15;
	return fused && (static_cast<long long>(aSrc.extent[0]) + 64) * std::llabs(aSrc.stride[0]) +
		(static_cast<long long>(aSrc.extent[1]) + 64) * std::llabs(aSrc.stride[1]) < (1LL << 31);
}

// Scratch of the per-axis path: up to two packed intermediate images (FP64 in
// REFERENCE precision, FP32 otherwise), each on a 256-byte boundary.
size_t separable_workspace(const view4& aSrc, const view4* aTaps, int aN, const view4& aDst, int aPrecision,
	bool aAllowFused)
{
	if (element_count(aDst) == 0 || (aAllowFused && separable_fused(aSrc, aTaps, aN, aDst, aPrecision))) {
		return 0;
	}
	const int temps = std::min(2, aSrc.rank - 1);
	const size_t one = static_cast<size_t>(element_count(aDst)) * (aPrecision == IANE_PREC_REFERENCE ? 8 : 4);
	return static_cast<size_t>(temps) * ((one + 255) / 256 * 256);
}

void convolve_separable(const view4& aSrc, const view4* aTaps, int aN, const view4& aDst, int aBorder,
	int aPrecision, void* aWorkspace, hipStream_t aStream, bool aAllowFused)
{
	const int rank = aSrc.rank;
	const int k = separable_window(aTaps, aN);
	const bool fused = aAllowFused && separable_fused(aSrc, aTaps, aN, aDst, aPrecision);
	if (fused) {
		// taps of unequal (odd) lengths are centred in the window of the longest;
		// a rank-2 image is one plane with the unit kernel along z
		const bool planar = rank == 2;
		sep_args a{};
		for (int axis = 0; axis < 3; ++axis) {
			const bool has = axis < aN;
			a.taps[axis] = has ? aTaps[axis].ptr : nullptr;
			a.tap_stride[axis] = has ? aTaps[axis].stride[0] : 0;
			a.tap_type[axis] = has ? aTaps[axis].type : IANE_F32;
			a.tap_len[axis] = has ? aTaps[axis].extent[0] : 1;
		}
		a.src = aSrc.ptr;
		a.src_type = aSrc.type;
		a.nx = aSrc.extent[0];
		a.ny = aSrc.extent[1];
		a.nz = aSrc.extent[2];
		a.sx = aSrc.stride[0];
		a.sy = aSrc.stride[1];
		a.sz = aSrc.stride[2];
		a.dst = aDst.ptr;
		a.dst_type = aDst.type;
		a.dx = aDst.stride[0];
		a.dy = aDst.stride[1];
		a.dz = aDst.stride[2];
		a.border = aBorder;
		if (planar) {
			// one plane: give the absent z axis a stride past the plane so plane
			// arithmetic stays in range (only plane 0 is ever addressed)
			a.sz = (static_cast<long long>(a.nx) - 1) * a.sx + (static_cast<long long>(a.ny) - 1) * a.sy + 4;
			a.sz = (a.sz + 15) / 16 * 16;
			a.dz = (static_cast<long long>(a.nx) - 1) * a.dx + (static_cast<long long>(a.ny) - 1) * a.dy + 4;
		}
		// Engines (all FP32 products with FP32 sums, in different orders):
		//   2 (default where it applies): barrier-free, x and y filters on FP32
		//     MFMA with operands loaded per wave, long z chunks;
		//   0: the same filters on operands staged in LDS (any border, stride);
		//   1: vector ALUs only.
		// The matrix engines address a plane with unsigned 32-bit lane offsets.
		const long long src_span = (static_cast<long long>(a.nx) - 1) * a.sx + (static_cast<long long>(a.ny) - 1) * a.sy + 4;
		const long long dst_span = (static_cast<long long>(a.nx) - 1) * a.dx + (static_cast<long long>(a.ny) - 1) * a.dy + 4;
		const int engine = tune_get("sep3.engine", 2);
		const bool matrix = engine != 1 && a.sx >= 0 && a.sy >= 0 && a.dx >= 0 && a.dy >= 0 &&
			src_span < (1LL << 31) && dst_span < (1LL << 31);
		// engine 2 reads 16-byte groups of 4 samples: packed rows, widths and
		// row pitches that are multiples of 4, a 16-byte aligned source, zero border
		const bool grouped = a.sx == 4 && a.nx % 4 == 0 && a.sy % 16 == 0 && a.sz % 16 == 0 &&
			reinterpret_cast<uintptr_t>(a.src) % 16 == 0 && a.border == IANE_BORDER_ZERO;
		// engine 3 (split FP16 operands, every filter on the matrix cores) also
		// addresses a chunk of output planes from one base: planes must not overlap
		if (aPrecision == IANE_PREC_SPLIT_FP16 && !planar && matrix && grouped && a.dz >= dst_span) {
			a.src_span = static_cast<int>(src_span);
			a.dst_span = static_cast<int>(dst_span);
			if (separable::split_chunk_fits(a, 16)) {
				separable::launch_split(a, k, aStream);
				check_hip(hipGetLastError(), "Separable convolution launch");
				return;
			}
		}
		if (matrix && engine == 2 && grouped) {
			a.src_span = static_cast<int>(src_span);
			a.dst_span = static_cast<int>(dst_span);
			switch (k) {
			case 1: launch_direct_rank<1>(a, planar, aStream); break;
			case 3: launch_direct_rank<3>(a, planar, aStream); break;
			case 5: launch_direct_rank<5>(a, planar, aStream); break;
			case 7: launch_direct_rank<7>(a, planar, aStream); break;
			case 9: launch_direct_rank<9>(a, planar, aStream); break;
			case 11: launch_direct_rank<11>(a, planar, aStream); break;
			case 13: launch_direct_rank<13>(a, planar, aStream); break;
			default: launch_direct_rank<15>(a, planar, aStream); break;
			}
			check_hip(hipGetLastError(), "Separable convolution launch");
			return;
		}
		if (matrix) {
			a.src_span = static_cast<int>(src_span);
			a.dst_span = static_cast<int>(dst_span);
			switch (k) {
			case 1: launch_mfma<1>(a, aStream); break;
			case 3: launch_mfma<3>(a, aStream); break;
			case 5: launch_mfma<5>(a, aStream); break;
			case 7: launch_mfma<7>(a, aStream); break;
			case 9: launch_mfma<9>(a, aStream); break;
			case 11: launch_mfma<11>(a, aStream); break;
			case 13: launch_mfma<13>(a, aStream); break;
			default: launch_mfma<15>(a, aStream); break;
			}
			check_hip(hipGetLastError(), "Separable convolution launch");
			return;
		}
		const int threads = tune_get("sep3.threads", 512);
		switch (k) {
		case 1: launch_fused<1>(a, aStream, threads); break;
		case 3: launch_fused<3>(a, aStream, threads); break;
		case 5: launch_fused<5>(a, aStream, threads); break;
		case 7: launch_fused<7>(a, aStream, threads); break;
		case 9: launch_fused<9>(a, aStream, threads); break;
		case 11: launch_fused<11>(a, aStream, threads); break;
		case 13: launch_fused<13>(a, aStream, threads); break;
		default: launch_fused<15>(a, aStream, threads); break;
		}
		check_hip(hipGetLastError(), "Separable convolution launch");
		return;
	}
	// general path: one pass per axis
	const bool wide = aPrecision == IANE_PREC_REFERENCE;
	const int temp_type = wide ? IANE_F64 : IANE_F32;
	const long long count = element_count(aDst);
	const size_t temp_bytes = (static_cast<size_t>(count) * dtype_size(temp_type) + 255) / 256 * 256;
	void* temps[2] = {aWorkspace, static_cast<char*>(aWorkspace) + temp_bytes};
	const int passes = rank;
	view4 in = aSrc;
	for (int axis = 0; axis < passes; ++axis) {
		const bool last = axis == passes - 1;
		const view4 out = last ? aDst : packed_temp(aDst, temp_type, temps[axis % 2]);
		if (wide) {
			axis_pass_kernel<double><<<ceil_div(count, 256), 256, 0, aStream>>>(in, out, aTaps[axis].ptr,
				aTaps[axis].type, aTaps[axis].stride[0], aTaps[axis].extent[0], axis, aBorder, count);
		} else {
			axis_pass_kernel<float><<<ceil_div(count, 256), 256, 0, aStream>>>(in, out, aTaps[axis].ptr,
				aTaps[axis].type, aTaps[axis].stride[0], aTaps[axis].extent[0], axis, aBorder, count);
		}
		check_hip(hipGetLastError(), "Separable pass launch");
		in = out;
	}
}

} // namespace iane_hip

using namespace iane_hip;

namespace {

struct separable_call {
	view4 src, dst;
	view4 taps[IANE_MAX_RANK];
	int n;
};

separable_call check_separable(const iane_view_desc* aSrc, const iane_view_desc* aTaps, int aN,
	const iane_view_desc* aDst, int aBorder, int aPrecision)
{
	separable_call c{};
	c.src = make_view4(aSrc, "source");
	c.dst = make_view4(aDst, "destination");
	if (!aTaps || aN != c.src.rank) {
		invalid("separable convolution needs one tap vector per axis");
	}
	c.n = aN;
	view4 shape = c.src;
	for (int i = 0; i < aN; ++i) {
		c.taps[i] = make_view4(&aTaps[i], "taps");
		if (c.taps[i].rank != 1 || (c.taps[i].type != IANE_F32 && c.taps[i].type != IANE_F64)) {
			invalid("separable taps must be rank-1 f32 or f64 views");
		}
		shape.extent[i] = c.taps[i].extent[0];
	}
	for (int i = aN; i < 4; ++i) {
		shape.extent[i] = 1;
	}
	shape.type = c.taps[0].type;
	validate_convolution(c.src, shape, c.dst);
	if (aBorder != IANE_BORDER_CLAMP && aBorder != IANE_BORDER_ZERO) {
		invalid("unknown border mode");
	}
	if (aPrecision != IANE_PREC_REFERENCE && aPrecision != IANE_PREC_FP32 && aPrecision != IANE_PREC_SPLIT_FP16) {
		fail(IANE_ERR_UNSUPPORTED, "separable convolution supports REFERENCE, FP32 and SPLIT_FP16 precision");
	}
	return c;
}

int run_separable(const iane_view_desc* aSrc, const iane_view_desc* aTaps, int aN, const iane_view_desc* aDst,
	int aBorder, int aPrecision, void* aWorkspace, size_t aWorkspaceBytes, void* aStream, bool aAllowFused)
{
	return guarded([&] {
		const separable_call c = check_separable(aSrc, aTaps, aN, aDst, aBorder, aPrecision);
		if (element_count(c.dst) == 0) {
			return;
		}
		check_workspace(aWorkspace, aWorkspaceBytes,
			separable_workspace(c.src, c.taps, c.n, c.dst, aPrecision, aAllowFused));
		convolve_separable(c.src, c.taps, c.n, c.dst, aBorder, aPrecision, aWorkspace,
			static_cast<hipStream_t>(aStream), aAllowFused);
	});
}

int separable_bytes(const iane_view_desc* aSrc, const iane_view_desc* aTaps, int aN, const iane_view_desc* aDst,
	int aBorder, int aPrecision, size_t* aBytes, bool aAllowFused)
{
	return guarded([&] {
		if (!aBytes) {
			invalid("workspace size output is null");
		}
		const separable_call c = check_separable(aSrc, aTaps, aN, aDst, aBorder, aPrecision);
		*aBytes = separable_workspace(c.src, c.taps, c.n, c.dst, aPrecision, aAllowFused);
	});
}

} // namespace

extern "C" int iane_hip_convolve_separable_workspace_bytes(const iane_view_desc* aSrc, const iane_view_desc* aTaps,
	int aN, const iane_view_desc* aDst, int aBorder, int aPrecision, size_t* aBytes)
{
	return separable_bytes(aSrc, aTaps, aN, aDst, aBorder, aPrecision, aBytes, true);
}

extern "C" int iane_hip_convolve_separable(const iane_view_desc* aSrc, const iane_view_desc* aTaps, int aN,
	const iane_view_desc* aDst, int aBorder, int aPrecision, void* aWorkspace, size_t aWorkspaceBytes, void* aStream)
{
	return run_separable(aSrc, aTaps, aN, aDst, aBorder, aPrecision, aWorkspace, aWorkspaceBytes, aStream, true);
}

// Test hooks: the per-axis pass path forced, with its own workspace query.
extern "C" int iane_hip_convolve_separable_passes_workspace_bytes(const iane_view_desc* aSrc,
	const iane_view_desc* aTaps, int aN, const iane_view_desc* aDst, int aBorder, int aPrecision, size_t* aBytes)
{
	return separable_bytes(aSrc, aTaps, aN, aDst, aBorder, aPrecision, aBytes, false);
}

extern "C" int iane_hip_convolve_separable_passes(const iane_view_desc* aSrc, const iane_view_desc* aTaps, int aN,
	const iane_view_desc* aDst, int aBorder, int aPrecision, void* aWorkspace, size_t aWorkspaceBytes, void* aStream)
{
	return run_separable(aSrc, aTaps, aN, aDst, aBorder, aPrecision, aWorkspace, aWorkspaceBytes, aStream, false);
}
