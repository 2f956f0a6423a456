// Direct N-D convolution (rank 1..4) behind iane_hip_convolve.
//
// Semantics follow the reference convolve (include/iane/algorithms/convolution.hpp:23-117):
//   dst[p] = conv( sum over kernel positions q in storage order (axis 0 fastest)
//                  of src_b[p + c - q] * k[q] ),   c = extent / 2 per axis,
// where src_b applies the border rule per axis (include/iane/locator/image_locator.hpp:40-56):
// an out-of-range coordinate makes the whole sample zero (zero) or is clamped
// to the view's extent (clamp). In REFERENCE precision the sum is an FP64 fma
// chain in exactly that order, so results are bit-identical to the reference.
//
// Tiled kernel: one workgroup (256 threads) computes a 64 x 64 output tile of
// one (z, w) plane (64 x 32 when that tile's LDS would not fit). For every
// kernel plane (qz, qw) it stages the input plane region with its x/y halo
// (border applied at staging) into LDS as the accumulator type, then each
// thread accumulates a 4 x 4 (4 x 2) register block, reading each input row's
// window along x once into registers and adding it to every output row it
// reaches (kernel width is a template parameter).
//
// Streaming kernel (rank <= 3, small kernels with kz >= 2): a workgroup owns a
// 64 x 16 (x, y) tile and walks a chunk of output planes; every input plane is
// staged once into an LDS ring of input planes (next plane prefetched into
// registers, one barrier per plane), so the input is read from memory once per
// chunk instead of once per kernel plane. The ring has kz + 2 slots: the plane
// stored at step z + 1 never overwrites a plane that step z still reads, so
// the barrier after each store is the only one. Each output keeps the same FP64
// fma chain in the reference's tap order (FP32 chains: kernel planes from the
// back, see plane_tap), so both kernels produce identical bits.
//
// Column kernel (3 x 3 x 3 kernels on x-contiguous images): no LDS and no
// barriers; input planes pass through registers and each plane adds its taps
// to the three output planes it reaches (described at conv_column_kernel).
#include "common.hpp"

#include <atomic>

namespace iane_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kThreadsX = 16;
constexpr int kRx = 4;                 // outputs per thread along x
constexpr int kTileX = kThreadsX * kRx; // 64
constexpr size_t kMaxLds = 64 * 1024;

struct conv_args {
	view4 src;
	view4 ker;
	view4 dst;
	int border;
	int center[4];
	unsigned tiles_x;
	unsigned tiles_y;
	int src_span; // bytes of one (x, y) plane (streaming kernels)
	int dst_span;
	int column_dot2; // column kernel: u16 images with small nonnegative taps on v_dot2_u32_u16
	int column_alternate; // column kernel, exact integer sums: odd z chunks march down
};

// Sample the source at an integer coordinate with the border rule; returns
// false when the zero border makes the sample zero.
__device__ __forceinline__ bool border_coord(int& aCoord, int aExtent, int aBorder)
{
	if (aCoord >= 0 && aCoord < aExtent) {
		return true;
	}
	if (aBorder == IANE_BORDER_ZERO || aExtent <= 0) {
		return false;
	}
	aCoord = aCoord < 0 ? 0 : aExtent - 1;
	return true;
}

// Kernel plane of the step-th term of an output's fma chain. FP64 chains take
// kernel planes in the reference's storage order (qz ascending: input planes
// from the highest z down). FP32 chains take them from the back, i.e. input
// planes in ascending z, the order in which a kernel marching up z receives
// them; rows and x taps keep the storage order in both.
template<class Acc>
__device__ __forceinline__ int plane_tap(int aStep, int aKz)
{
	return std::is_same_v<Acc, float> ? aKz - 1 - aStep : aStep;
}

template<class Acc, int KX, int RY>
__global__ __launch_bounds__(kThreads) void conv_tiled_kernel(conv_args a)
{
	constexpr int kTy = (kThreads / kThreadsX) * RY;  // output rows per tile
	extern __shared__ __align__(16) char smem[];
	const int kx = KX > 0 ? KX : a.ker.extent[0];
	const int ky = a.ker.extent[1];
	const int kz = a.ker.extent[2];
	const int kw = a.ker.extent[3];
	const int cols = kTileX + kx - 1;
	const int rows = kTy + ky - 1;
	Acc* tile = reinterpret_cast<Acc*>(smem);
	Acc* coef = tile + rows * cols;

	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const unsigned tile_x = block % a.tiles_x;
	const unsigned rest = block / a.tiles_x;
	const unsigned tile_y = rest % a.tiles_y;
	const unsigned plane = rest / a.tiles_y;
	const int pz = static_cast<int>(plane % static_cast<unsigned>(a.dst.extent[2]));
	const int pw = static_cast<int>(plane / static_cast<unsigned>(a.dst.extent[2]));
	const int x0 = static_cast<int>(tile_x) * kTileX;
	const int y0 = static_cast<int>(tile_y) * kTy;
	const int tid = threadIdx.x;
	const int tx = tid % kThreadsX;
	const int ty = tid / kThreadsX;

	Acc acc[RY][kRx];
#pragma unroll
	for (int ry = 0; ry < RY; ++ry) {
#pragma unroll
		for (int rx = 0; rx < kRx; ++rx) {
			acc[ry][rx] = Acc(0);
		}
	}

	for (int qw = 0; qw < kw; ++qw) {
		for (int step = 0; step < kz; ++step) {
			const int qz = plane_tap<Acc>(step, kz);
			__syncthreads();
			// Coefficients of kernel plane (qz, qw), converted like the reference does.
			for (int i = tid; i < kx * ky; i += kThreads) {
				const int qx = i % kx;
				const int qy = i / kx;
				coef[i] = load_as<Acc>(a.ker.ptr + qx * a.ker.stride[0] + qy * a.ker.stride[1] +
					qz * a.ker.stride[2] + qw * a.ker.stride[3], a.ker.type);
			}
			int zi = pz + a.center[2] - qz;
			int wi = pw + a.center[3] - qw;
			const bool plane_in = border_coord(zi, a.src.extent[2], a.border) &&
				border_coord(wi, a.src.extent[3], a.border);
			const char* base = a.src.ptr + (plane_in ? zi * a.src.stride[2] + wi * a.src.stride[3] : 0);
			// Stage the tile kStage elements per thread at a time: every load of a
			// batch is issued (a zero-border sample reads the plane's first sample)
			// before any is used, with the element type fixed outside the loop.
			constexpr int kStage = 8;
			const int total = rows * cols;
			with_element_type(a.src.type, [&](auto aType) {
				using S = decltype(aType);
				for (int i0 = tid; i0 < total; i0 += kStage * kThreads) {
					S raw[kStage];
					bool inside[kStage];
#pragma unroll
					for (int u = 0; u < kStage; ++u) {
						const int i = min(i0 + u * kThreads, total - 1);
						const int ly = i / cols;
						const int lx = i - ly * cols;
						int xi = x0 - a.center[0] + lx;
						int yi = y0 - a.center[1] + ly;
						inside[u] = plane_in && border_coord(xi, a.src.extent[0], a.border) &&
							border_coord(yi, a.src.extent[1], a.border);
						const long long offset = inside[u] ? xi * a.src.stride[0] + yi * a.src.stride[1] : 0;
						raw[u] = *reinterpret_cast<const S*>(base + offset);
					}
#pragma unroll
					for (int u = 0; u < kStage; ++u) {
						const int i = i0 + u * kThreads;
						if (i < total) {
							tile[i] = inside[u] ? static_cast<Acc>(raw[u]) : Acc(0);
						}
					}
				}
			});
			__syncthreads();
			if constexpr (KX > 0) {
				// Tile row ty * RY + rho meets output row ry through kernel row
				// qy = ry + ky - 1 - rho. Rows are taken from the last down, so
				// each output's chain still runs qy ascending, x taps ascending
				// (the reference's order), and every window is read from LDS once.
				constexpr int span = kRx + KX - 1;
				for (int rho = RY + ky - 2; rho >= 0; --rho) {
					const Acc* row = tile + (ty * RY + rho) * cols + tx * kRx;
					Acc window[span];
#pragma unroll
					for (int j = 0; j < span; ++j) {
						window[j] = row[j];
					}
#pragma unroll
					for (int ry = 0; ry < RY; ++ry) {
						const int qy = ry + ky - 1 - rho;
						if (qy >= 0 && qy < ky) {
							const Acc* krow = coef + qy * kx;
#pragma unroll
							for (int qx = 0; qx < KX; ++qx) {
								const Acc k = krow[qx];
#pragma unroll
								for (int rx = 0; rx < kRx; ++rx) {
									acc[ry][rx] = fma(window[rx + KX - 1 - qx], k, acc[ry][rx]);
								}
							}
						}
					}
				}
			} else {
				for (int qy = 0; qy < ky; ++qy) {
					const Acc* krow = coef + qy * kx;
					for (int qx = 0; qx < kx; ++qx) {
						const Acc k = krow[qx];
#pragma unroll
						for (int ry = 0; ry < RY; ++ry) {
							const Acc* row = tile + (ty * RY + ry + (ky - 1 - qy)) * cols + tx * kRx + (kx - 1 - qx);
#pragma unroll
							for (int rx = 0; rx < kRx; ++rx) {
								acc[ry][rx] = fma(row[rx], k, acc[ry][rx]);
							}
						}
					}
				}
			}
		}
	}

	char* out = a.dst.ptr + pz * a.dst.stride[2] + pw * a.dst.stride[3];
#pragma unroll
	for (int ry = 0; ry < RY; ++ry) {
		const int py = y0 + ty * RY + ry;
		if (py >= a.dst.extent[1]) {
			continue;
		}
#pragma unroll
		for (int rx = 0; rx < kRx; ++rx) {
			const int px = x0 + tx * kRx + rx;
			if (px < a.dst.extent[0]) {
				store_converted(out + px * a.dst.stride[0] + py * a.dst.stride[1], a.dst.type, acc[ry][rx]);
			}
		}
	}
}

// Fallback for kernels whose staged tile does not fit in LDS: one thread per
// output, every tap read from global memory with the border rule, same order.
template<class Acc>
__global__ __launch_bounds__(kThreads) void conv_simple_kernel(conv_args a, long long aCount)
{
	const long long index = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (index >= aCount) {
		return;
	}
	int p[4];
	long long rest = index;
	for (int i = 0; i < 4; ++i) {
		p[i] = static_cast<int>(rest % a.dst.extent[i]);
		rest /= a.dst.extent[i];
	}
	Acc sum = Acc(0);
	for (int qw = 0; qw < a.ker.extent[3]; ++qw) {
		for (int step = 0; step < a.ker.extent[2]; ++step) {
			const int qz = plane_tap<Acc>(step, a.ker.extent[2]);
			for (int qy = 0; qy < a.ker.extent[1]; ++qy) {
				for (int qx = 0; qx < a.ker.extent[0]; ++qx) {
					const int q[4] = {qx, qy, qz, qw};
					bool inside = true;
					long long offset = 0;
					for (int i = 0; i < 4; ++i) {
						int c = p[i] + a.center[i] - q[i];
						inside = inside && border_coord(c, a.src.extent[i], a.border);
						offset += static_cast<long long>(c) * a.src.stride[i];
					}
					const Acc sample = inside ? load_as<Acc>(a.src.ptr + offset, a.src.type) : Acc(0);
					const Acc k = load_as<Acc>(a.ker.ptr + qx * a.ker.stride[0] + qy * a.ker.stride[1] +
						qz * a.ker.stride[2] + qw * a.ker.stride[3], a.ker.type);
					sum = fma(sample, k, sum);
				}
			}
		}
	}
	char* out = a.dst.ptr;
	for (int i = 0; i < 4; ++i) {
		out += static_cast<long long>(p[i]) * a.dst.stride[i];
	}
	store_converted(out, a.dst.type, sum);
}

constexpr int kStreamTileY = 16;    // output rows per streaming workgroup (16 x 16 threads, 4 x-outputs each)
constexpr int kStreamStage = 6;     // staged elements per thread and plane (at most)
constexpr int kStreamDepth = 2;     // input planes of loads in flight per thread

// LDS row length of a streamed plane, a multiple of 4 samples: every thread's
// window (4 outputs + KX - 1 halo samples) starts 16-byte aligned.
__host__ __device__ constexpr int stream_cols(int aKx)
{
	return (kTileX + aKx - 1 + 3) & ~3;
}

// Read SPAN consecutive 4-byte samples from a 16-byte aligned LDS row with
// 16- and 8-byte reads.
template<int SPAN, class T>
__device__ __forceinline__ void read_window(const T* aRow, T (&aOut)[SPAN])
{
	static_assert(sizeof(T) == 4 && SPAN % 2 == 0);
	using V4 = T __attribute__((ext_vector_type(4)));
	using V2 = T __attribute__((ext_vector_type(2)));
#pragma unroll
	for (int j = 0; j + 4 <= SPAN; j += 4) {
		const V4 v = *reinterpret_cast<const V4*>(aRow + j);
		aOut[j] = v.x;
		aOut[j + 1] = v.y;
		aOut[j + 2] = v.z;
		aOut[j + 3] = v.w;
	}
	if constexpr (SPAN % 4 == 2) {
		const V2 v = *reinterpret_cast<const V2*>(aRow + SPAN - 2);
		aOut[SPAN - 2] = v.x;
		aOut[SPAN - 1] = v.y;
	}
}

constexpr int kOob = 0x7fffff00; // byte offset past every plane a streaming kernel addresses

// Raw buffer resource of one plane (GFX9 word 3: 32-bit data format): loads at
// offsets at or past `aBytes` read 0 and stores there are dropped, so border
// samples, tile overhang and a zero-border plane need no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t stream_resource(const char* aBase, int aBytes)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(aBase), static_cast<short>(0), aBytes, 0x00020000);
}

template<class T>
__device__ __forceinline__ T buffer_load_as(__amdgpu_buffer_rsrc_t aResource, int aOffset)
{
	if constexpr (sizeof(T) == 1) {
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b8(aResource, aOffset, 0, 0));
	} else if constexpr (sizeof(T) == 2) {
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(aResource, aOffset, 0, 0));
	} else if constexpr (sizeof(T) == 4) {
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(aResource, aOffset, 0, 0));
	} else {
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(aResource, aOffset, 0, 0));
	}
}

// The b8 .. b64 builtins move raw words: bit casts, not conversions.
template<class T>
__device__ __forceinline__ void buffer_store_as(__amdgpu_buffer_rsrc_t aResource, int aOffset, T aValue)
{
	if constexpr (sizeof(T) == 1) {
		__builtin_amdgcn_raw_buffer_store_b8(__builtin_bit_cast(unsigned char, aValue), aResource, aOffset, 0, 0);
	} else if constexpr (sizeof(T) == 2) {
		__builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, aValue), aResource, aOffset, 0, 0);
	} else if constexpr (sizeof(T) == 4) {
		__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, aValue), aResource, aOffset, 0, 0);
	} else {
		using v2u = unsigned __attribute__((ext_vector_type(2)));
		__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, aValue), aResource, aOffset, 0, 0);
	}
}

// The output conversion of store_converted for a static destination type.
template<class D, class Acc>
__device__ __forceinline__ D convert_out(Acc aSum)
{
	if constexpr (std::is_same_v<D, float> && std::is_same_v<Acc, float>) {
		return aSum;
	} else {
		const double sum = static_cast<double>(aSum);
		if constexpr (std::is_same_v<D, uint8_t>) {
			return round_byte(sum);
		} else if constexpr (std::is_same_v<D, uint16_t> || std::is_same_v<D, int32_t>) {
			return saturate_truncate<D>(sum);
		} else {
			return static_cast<D>(sum);
		}
	}
}

// convert_out<D>(double(aSum)) for an exact integer |aSum| < 2^31: the double
// is exact, so rounding and truncation do nothing and only the clamps remain
// (u8 rounds half up after clamping, u16 and i32 saturate; float rounds the
// integer once either way).
template<class D>
__device__ __forceinline__ D convert_exact_int(int aSum)
{
	if constexpr (std::is_same_v<D, uint8_t>) {
		return static_cast<uint8_t>(min(max(aSum, 0), 255));
	} else if constexpr (std::is_same_v<D, uint16_t>) {
		return static_cast<uint16_t>(min(max(aSum, 0), 65535));
	} else {
		return static_cast<D>(aSum);
	}
}

// Where a streaming workgroup works and which samples each of its threads
// stages: a 64 x 16 output tile of planes [z0, z1), fed by input planes
// [z0 - cz, z1 + cz) that pass through an LDS ring of kz + 2 planes. A thread
// stages elements tid + e * 256 (e < kStreamStage) of every plane, with the x and
// y border applied (kOob where the zero border or the tile overhang makes the
// sample 0), and writes outputs x0 + 4 tx + r of row y0 + ty.
struct stream_tile {
	int x0, y0, z0, z1, cz;
	int rows, cols, slots, plane_elems;
	int tx, ty, tid;
	int offset[kStreamStage];
	int out_offset[kRx];

	__device__ stream_tile(const conv_args& a, int aCols, int aChunk)
	{
		const int ky = a.ker.extent[1];
		const int kz = a.ker.extent[2];
		const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
		const unsigned tile_x = block % a.tiles_x;
		const unsigned rest = block / a.tiles_x;
		const unsigned tile_y = rest % a.tiles_y;
		const unsigned chunk = rest / a.tiles_y;
		x0 = static_cast<int>(tile_x) * kTileX;
		y0 = static_cast<int>(tile_y) * kStreamTileY;
		z0 = static_cast<int>(chunk) * aChunk;
		z1 = min(z0 + aChunk, a.dst.extent[2]);
		cz = a.center[2];
		rows = kStreamTileY + ky - 1;
		cols = aCols;
		slots = kz + 2;
		plane_elems = rows * cols;
		tid = threadIdx.x;
		tx = tid % kThreadsX;
		ty = tid / kThreadsX;
#pragma unroll
		for (int e = 0; e < kStreamStage; ++e) {
			const int i = tid + e * kThreads;
			const int ly = i / cols;
			const int lx = i - ly * cols;
			int xi = x0 - a.center[0] + lx;
			int yi = y0 - a.center[1] + ly;
			const bool inside = i < plane_elems && border_coord(xi, a.src.extent[0], a.border) &&
				border_coord(yi, a.src.extent[1], a.border);
			offset[e] = inside ? static_cast<int>(xi * a.src.stride[0] + yi * a.src.stride[1]) : kOob;
		}
		const int py = y0 + ty;
#pragma unroll
		for (int r = 0; r < kRx; ++r) {
			const int px = x0 + tx * kRx + r;
			out_offset[r] = px < a.dst.extent[0] && py < a.dst.extent[1] ?
				static_cast<int>(px * a.dst.stride[0] + py * a.dst.stride[1]) : kOob;
		}
	}

	// Ring slot of the plane qz steps before the one in slot aTop.
	__device__ int slot_before(int aTop, int aSteps) const
	{
		const int slot = aTop - aSteps;
		return slot < 0 ? slot + slots : slot;
	}

	// Issue the loads of one input plane (z border applied) into registers;
	// nothing uses them before the plane is stored to the ring P steps later,
	// so they stay in flight across the taps of the planes in between.
	// u8 / u16 samples are held one per 32-bit register: packing two into one
	// register would use them (and wait for them) right after the load.
	template<class S>
	using staged_t = std::conditional_t<(sizeof(S) < 4), uint32_t, S>;

	// A load with aWanted false is still issued (so every path through the
	// pipeline has the same loads in flight) but reads 0 without touching memory.
	template<class S>
	__device__ void load(const conv_args& a, int aPlane, bool aWanted, staged_t<S> (&aRaw)[kStreamStage]) const
	{
		int zi = aPlane;
		const bool plane_in = aWanted && border_coord(zi, a.src.extent[2], a.border);
		const auto plane = stream_resource(a.src.ptr + static_cast<long long>(plane_in ? zi : 0) * a.src.stride[2],
			plane_in ? a.src_span : 0);
#pragma unroll
		for (int e = 0; e < kStreamStage; ++e) {
			aRaw[e] = static_cast<staged_t<S>>(buffer_load_as<S>(plane, offset[e]));
		}
	}

	// Store a staged plane to its ring slot. Elements past the plane go to a
	// per-thread spare word after the ring (aSpare[tid]) instead of behind a
	// branch.
	template<class T, class S>
	__device__ void store(T* aRing, T* aSpare, int aSlot, const S (&aRaw)[kStreamStage]) const
	{
		T* slot = aRing + aSlot * plane_elems;
#pragma unroll
		for (int e = 0; e < kStreamStage; ++e) {
			const int i = tid + e * kThreads;
			T* at = i < plane_elems ? slot + i : aSpare + tid;
			*at = static_cast<T>(aRaw[e]);
		}
	}

	template<class D, class Acc>
	__device__ void write(const conv_args& a, int aZ, const Acc (&aSums)[kRx]) const
	{
		const auto plane = stream_resource(a.dst.ptr + static_cast<long long>(aZ) * a.dst.stride[2], a.dst_span);
#pragma unroll
		for (int r = 0; r < kRx; ++r) {
			if constexpr (std::is_same_v<Acc, int>) {
				buffer_store_as(plane, out_offset[r], convert_exact_int<D>(aSums[r]));
			} else {
				buffer_store_as(plane, out_offset[r], convert_out<D>(aSums[r]));
			}
		}
	}

	// Stream the chunk with P planes of loads in flight: each input plane is
	// loaded P steps before it is stored to the ring; after the store of plane
	// z + cz (to slot s) and a barrier, aCompute(z, s) reads planes
	// z + cz - kz + 1 .. z + cz from slots s - kz + 1 .. s (mod kz + 2).
	// The ring has kz + 2 slots, so the store of the next plane never lands on a
	// slot that a thread still computing the previous output plane reads.
	// The main loop issues a load after every store (past the chunk: a load
	// that reads nothing), so at each store the P - 1 younger planes' loads are
	// known to be in flight and the wait covers only the plane stored; the
	// last count mod P planes run in a tail with no loads.
	template<int P, class S, class T, class Compute>
	__device__ void run(const conv_args& a, T* aRing, T* aSpare, Compute aCompute) const
	{
		const int first = z0 - cz;
		const int count = (z1 - z0) + 2 * cz;
		const int lead = 2 * cz; // planes stored before the first output plane
		staged_t<S> staged[P][kStreamStage];
#pragma unroll
		for (int q = 0; q < P; ++q) {
			load<S>(a, first + q, q < count, staged[q]);
		}
		int slot = 0; // ring slot of plane first + i (the ring starts empty at every chunk)
		int base = 0;
		for (; base + P <= count; base += P) {
#pragma unroll
			for (int q = 0; q < P; ++q) {
				const int i = base + q;
				store(aRing, aSpare, slot, staged[q]);
				load<S>(a, first + i + P, i + P < count, staged[q]);
				if (i >= lead) {
					__syncthreads();
					aCompute(first + i - cz, slot);
				}
				slot = slot + 1 == slots ? 0 : slot + 1;
			}
		}
#pragma unroll
		for (int q = 0; q + 1 < P; ++q) {
			const int i = base + q;
			if (i < count) {
				store(aRing, aSpare, slot, staged[q]);
				if (i >= lead) {
					__syncthreads();
					aCompute(first + i - cz, slot);
				}
				slot = slot + 1 == slots ? 0 : slot + 1;
			}
		}
	}
};

__device__ __forceinline__ int load_as_int(const char* aPtr, int aType)
{
	switch (aType) {
	case IANE_U8: return *reinterpret_cast<const uint8_t*>(aPtr);
	case IANE_U16: return *reinterpret_cast<const uint16_t*>(aPtr);
	default: return *reinterpret_cast<const int32_t*>(aPtr);
	}
}

// Kernel coefficients of a streaming workgroup: the KX x ky x kz taps in LDS
// (x fastest); with KYZ > 0 (ky = kz = KYZ) also in registers, where the fully
// unrolled tap loops read them.
template<class T, int KX, int KYZ>
struct stream_coefficients {
	T* lds;
	T reg[KYZ > 0 ? KX * KYZ * KYZ : 1];

	// Fills the LDS copy; every thread must pass a barrier before load_registers.
	__device__ stream_coefficients(const conv_args& a, T* aLds, int aTid) : lds(aLds)
	{
		const int ky = a.ker.extent[1];
		const int taps = KX * ky * a.ker.extent[2];
		for (int i = aTid; i < taps; i += kThreads) {
			const int qx = i % KX;
			const int qy = (i / KX) % ky;
			const int qz = i / (KX * ky);
			const char* at = a.ker.ptr + qx * a.ker.stride[0] + qy * a.ker.stride[1] + qz * a.ker.stride[2];
			if constexpr (std::is_same_v<T, float>) {
				lds[i] = load_as<float>(at, a.ker.type);
			} else {
				lds[i] = load_as_int(at, a.ker.type);
			}
		}
	}

	// The coefficients are the same for every lane: read into scalar registers.
	__device__ void load_registers()
	{
		if constexpr (KYZ > 0) {
#pragma unroll
			for (int i = 0; i < KX * KYZ * KYZ; ++i) {
				reg[i] = __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lds[i])));
			}
		}
	}
};

// Accumulate the taps of the output plane whose newest input plane is in ring
// slot aTop: aMac(acc, sample, coefficient) over kernel planes in plane_tap
// order, rows qy (the tiled kernel's order: rows from the back) and x taps from
// the back. UNROLL
// false keeps runtime loops over the LDS coefficients even when KYZ > 0.
template<int KX, int KYZ, bool UNROLL = true, class T, class Acc, class Mac>
__device__ __forceinline__ void stream_taps(const stream_tile& t, const T* aRing,
	const stream_coefficients<T, KX, KYZ>& aCoef, int aKy, int aKz, int aTop, Acc (&aAcc)[kRx], Mac aMac)
{
	constexpr int span = kRx + KX - 1;
	const auto row_taps = [&](const T* aSlot, int aRow, const T* aK) {
		T window[span];
		read_window(aSlot + (t.ty + aRow) * t.cols + t.tx * kRx, window);
#pragma unroll
		for (int qx = 0; qx < KX; ++qx) {
			const T k = aK[qx];
#pragma unroll
			for (int rx = 0; rx < kRx; ++rx) {
				aMac(aAcc[rx], window[rx + KX - 1 - qx], k);
			}
		}
	};
	if constexpr (KYZ > 0 && UNROLL) {
#pragma unroll
		for (int step = 0; step < KYZ; ++step) {
			const int qz = plane_tap<Acc>(step, KYZ);
			const T* slot = aRing + t.slot_before(aTop, qz) * t.plane_elems;
#pragma unroll
			for (int qy = 0; qy < KYZ; ++qy) {
				row_taps(slot, KYZ - 1 - qy, aCoef.reg + (qz * KYZ + qy) * KX);
			}
		}
	} else {
		for (int step = 0; step < aKz; ++step) {
			const int qz = plane_tap<Acc>(step, aKz);
			const T* slot = aRing + t.slot_before(aTop, qz) * t.plane_elems;
			for (int qy = 0; qy < aKy; ++qy) {
				row_taps(slot, aKy - 1 - qy, aCoef.lds + (qz * aKy + qy) * KX);
			}
		}
	}
}

// Streaming kernel for FP32 accumulation: the same fma chain per output as the
// tiled kernel (kernel rows from the back, x taps within a row from the back),
// hence the same bits.
template<class S, class D, int KX, int KYZ>
__global__ __launch_bounds__(kThreads) void conv_stream_kernel(conv_args a, int aChunk)
{
	extern __shared__ __align__(16) char smem[];
	const int ky = a.ker.extent[1];
	const int kz = a.ker.extent[2];
	const stream_tile t(a, stream_cols(KX), aChunk);
	float* ring = reinterpret_cast<float*>(smem);
	stream_coefficients<float, KX, KYZ> coef(a, ring + t.slots * t.plane_elems, t.tid);
	float* spare = coef.lds + KX * ky * kz;
	__syncthreads();
	coef.load_registers();
	const auto compute = [&](int z, int aTop) {
		float acc[kRx] = {0.0f, 0.0f, 0.0f, 0.0f};
		stream_taps(t, ring, coef, ky, kz, aTop, acc,
			[](float& aSum, float aSample, float aK) { aSum = fma(aSample, aK, aSum); });
		t.write<D>(a, z, acc);
	};
	t.run<kStreamDepth, S>(a, ring, spare, compute);
}

// Streaming kernel for integer images (u8, u16) with integer kernels: the ring
// holds int32 samples. When every partial sum provably fits in int32 (largest
// source value x sum |k| < 2^31, and |k| < 2^23 so the 24-bit multiply is
// exact), taps accumulate with 24-bit integer multiply-adds; the sum is the
// exact integer the reference's FP64 chain represents (it is exact below 2^53),
// converted the same way. Otherwise the workgroup runs the FP64 chain in the
// tiled kernel's tap order on the same ring.
template<class S, class D, int KX, int KYZ>
__global__ __launch_bounds__(kThreads) void conv_stream_int_kernel(conv_args a, int aChunk)
{
	extern __shared__ __align__(16) char smem[];
	const int ky = a.ker.extent[1];
	const int kz = a.ker.extent[2];
	const stream_tile t(a, stream_cols(KX), aChunk);
	int* ring = reinterpret_cast<int*>(smem);
	stream_coefficients<int, KX, KYZ> coef(a, ring + t.slots * t.plane_elems, t.tid);
	int* spare = coef.lds + KX * ky * kz;
	__shared__ int exact_int;

	__syncthreads();
	coef.load_registers();
	if (t.tid == 0) {
		constexpr double source_max = sizeof(S) == 1 ? 255.0 : 65535.0;
		const int taps = KX * ky * kz;
		double total = 0.0;
		bool small = true;
		for (int i = 0; i < taps; ++i) {
			const double k = coef.lds[i] < 0 ? -static_cast<double>(coef.lds[i]) : static_cast<double>(coef.lds[i]);
			total += k;
			small = small && k < 8388608.0;
		}
		exact_int = small && total * source_max < 2147483648.0 ? 1 : 0;
	}
	// exact_int is read after the barrier that precedes the first output plane
	const auto compute = [&](int z, int aTop) {
		if (exact_int != 0) { // uniform over the workgroup
			int acc[kRx] = {0, 0, 0, 0};
			stream_taps(t, ring, coef, ky, kz, aTop, acc,
				[](int& aSum, int aSample, int aK) { aSum += __mul24(aSample, aK); });
			t.write<D>(a, z, acc);
		} else {
			double sums[kRx] = {0.0, 0.0, 0.0, 0.0};
			stream_taps<KX, KYZ, false>(t, ring, coef, ky, kz, aTop, sums, [](double& aSum, int aSample, int aK) {
				aSum = fma(static_cast<double>(aSample), static_cast<double>(aK), aSum);
			});
			t.write<D>(a, z, sums);
		}
	};
	t.run<kStreamDepth, S>(a, ring, spare, compute);
}

// ---------------------------------------------------------------------------
// Column kernel: 3 x 3 x 3 kernels on images whose rows are x-contiguous and
// aligned for 4-sample vector accesses (x extent a multiple of 4). A workgroup
// of four waves owns a 64 x 16R (x, y) tile of one z chunk; lane (tx, ty) of
// wave w owns outputs x0 + 4 tx .. + 3 of the R rows starting at
// y0 + R (4 w + ty). Input planes pass through registers, loaded P planes
// ahead: R + 2 rows of one 4-sample vector per lane. The x neighbours (x - 1,
// x + 4) come from the adjacent lanes of the lane's 16-lane DPP row; the lanes
// at the tile's x ends load the samples beyond it. Each plane adds its taps to
// the three output planes it reaches (a ring of three register accumulator
// sets) and an output plane is stored after its last input plane, so there is
// no LDS and no barrier. FP32 sums keep the FP32 chain of the other direct
// kernels (plane_tap order); integer sums are exact in int32 when the kernel's
// bound allows (otherwise each output is the FP64 chain in the reference's
// order, read straight from memory).
constexpr int kColumnThreads = 256;

using column_v4 = unsigned __attribute__((ext_vector_type(4)));
using column_v2 = unsigned __attribute__((ext_vector_type(2)));

// One 4-sample vector of S as a single buffer access, kept as raw words until
// it is used (unpacking at the load would wait for it there).
template<class S>
struct column_io;

template<>
struct column_io<float> {
	using raw = column_v4;
	__device__ static raw load(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b128(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static void unpack(const raw& aRaw, T (&aOut)[4])
	{
		// Bit-cast the whole vector: this compiler reads a bit cast of one
		// vector component (v.y) from the vector's first element.
		using f4 = float __attribute__((ext_vector_type(4)));
		const f4 f = __builtin_bit_cast(f4, aRaw);
		aOut[0] = static_cast<T>(f.x);
		aOut[1] = static_cast<T>(f.y);
		aOut[2] = static_cast<T>(f.z);
		aOut[3] = static_cast<T>(f.w);
	}
	__device__ static unsigned load_single(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b32(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static T single(unsigned aRaw) { return static_cast<T>(__builtin_bit_cast(float, aRaw)); }
};

template<>
struct column_io<uint16_t> {
	using raw = column_v2;
	__device__ static raw load(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b64(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static void unpack(const raw& aRaw, T (&aOut)[4])
	{
		aOut[0] = static_cast<T>(static_cast<int>(aRaw.x & 0xffffu));
		aOut[1] = static_cast<T>(static_cast<int>(aRaw.x >> 16));
		aOut[2] = static_cast<T>(static_cast<int>(aRaw.y & 0xffffu));
		aOut[3] = static_cast<T>(static_cast<int>(aRaw.y >> 16));
	}
	__device__ static unsigned load_single(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b16(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static T single(unsigned aRaw) { return static_cast<T>(static_cast<int>(aRaw)); }
};

template<>
struct column_io<uint8_t> {
	using raw = unsigned;
	__device__ static raw load(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b32(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static void unpack(const raw& aRaw, T (&aOut)[4])
	{
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			aOut[e] = static_cast<T>(static_cast<int>((aRaw >> (8 * e)) & 0xffu));
		}
	}
	__device__ static unsigned load_single(__amdgpu_buffer_rsrc_t aRes, int aOffset)
	{
		return __builtin_amdgcn_raw_buffer_load_b8(aRes, aOffset, 0, 0);
	}
	template<class T>
	__device__ static T single(unsigned aRaw) { return static_cast<T>(static_cast<int>(aRaw)); }
};

// Store four converted outputs of one row with one buffer access (two for f64).
template<class D>
__device__ __forceinline__ void column_store(__amdgpu_buffer_rsrc_t aRes, int aOffset, const D (&aOut)[4])
{
	if constexpr (sizeof(D) == 4) {
		const column_v4 w = {__builtin_bit_cast(unsigned, aOut[0]), __builtin_bit_cast(unsigned, aOut[1]),
			__builtin_bit_cast(unsigned, aOut[2]), __builtin_bit_cast(unsigned, aOut[3])};
		__builtin_amdgcn_raw_buffer_store_b128(w, aRes, aOffset, 0, 0);
	} else if constexpr (sizeof(D) == 2) {
		const column_v2 w = {
			static_cast<unsigned>(__builtin_bit_cast(uint16_t, aOut[0])) |
				(static_cast<unsigned>(__builtin_bit_cast(uint16_t, aOut[1])) << 16),
			static_cast<unsigned>(__builtin_bit_cast(uint16_t, aOut[2])) |
				(static_cast<unsigned>(__builtin_bit_cast(uint16_t, aOut[3])) << 16)};
		__builtin_amdgcn_raw_buffer_store_b64(w, aRes, aOffset, 0, 0);
	} else if constexpr (sizeof(D) == 1) {
		const unsigned w = static_cast<unsigned>(aOut[0]) | (static_cast<unsigned>(aOut[1]) << 8) |
			(static_cast<unsigned>(aOut[2]) << 16) | (static_cast<unsigned>(aOut[3]) << 24);
		__builtin_amdgcn_raw_buffer_store_b32(w, aRes, aOffset, 0, 0);
	} else {
		using v2d = double __attribute__((ext_vector_type(2)));
		__builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(column_v4, v2d{aOut[0], aOut[1]}), aRes, aOffset,
			0, 0);
		__builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(column_v4, v2d{aOut[2], aOut[3]}), aRes,
			aOffset + 16, 0, 0);
	}
}

// v of the lane one to the left (row_shr:1) or right (row_shl:1) in the lane's
// 16-lane row; the row's first (last) lane gets aEdge.
template<class T>
__device__ __forceinline__ T column_from_left(T aEdge, T aValue)
{
	return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, aEdge),
		__builtin_bit_cast(int, aValue), 0x111, 0xf, 0xf, false));
}

template<class T>
__device__ __forceinline__ T column_from_right(T aEdge, T aValue)
{
	return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, aEdge),
		__builtin_bit_cast(int, aValue), 0x101, 0xf, 0xf, false));
}

// Where a column lane works: its tile, its x, its first output row and its z chunk.
struct column_geom {
	int x0, x, tx, yb, z0, z1;
	int ext_x, ext_y, ext_z;
	int chunk;
	int step;  // +1: the chunk is marched up in z, -1: down
	bool clamp;

	template<int R>
	__device__ static column_geom make(const conv_args& a, int aChunk)
	{
		const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
		const unsigned tile_x = block % a.tiles_x;
		const unsigned rest = block / a.tiles_x;
		const unsigned tile_y = rest % a.tiles_y;
		const int chunk = static_cast<int>(rest / a.tiles_y);
		const int lane = static_cast<int>(threadIdx.x) % 64;
		const int group = static_cast<int>(threadIdx.x) / 64 * 4 + lane / 16;
		column_geom g;
		g.ext_x = a.src.extent[0];
		g.ext_y = a.src.extent[1];
		g.ext_z = a.src.extent[2];
		g.tx = lane % 16;
		g.x0 = static_cast<int>(tile_x) * 64;
		g.x = g.x0 + 4 * g.tx;
		g.yb = static_cast<int>(tile_y) * (16 * R) + group * R;
		g.z0 = chunk * aChunk;
		g.z1 = min(g.z0 + aChunk, g.ext_z);
		g.chunk = chunk;
		g.step = 1;
		g.clamp = a.border == IANE_BORDER_CLAMP;
		return g;
	}
};

__device__ __forceinline__ bool column_in_chunk(const column_geom& g, int aZ)
{
	return aZ >= g.z0 && aZ < g.z1;
}

// Coefficient qx + 3 qy + 9 qz of a 3^3 kernel, the same for every lane: read
// into scalar registers, converted like load_as_int / load_as_float, with one
// element-type dispatch for all 27 (a switch per coefficient would put each
// load in its own block and wait for it there).
template<class T>
__device__ __forceinline__ void column_coefficients(const conv_args& a, T (&aK)[27])
{
	with_element_type(a.ker.type, [&](auto aType) {
		using E = decltype(aType);
		E raw[27];
#pragma unroll
		for (int i = 0; i < 27; ++i) {
			raw[i] = *reinterpret_cast<const E*>(a.ker.ptr + (i % 3) * a.ker.stride[0] +
				(i / 3 % 3) * a.ker.stride[1] + (i / 9) * a.ker.stride[2]);
		}
#pragma unroll
		for (int i = 0; i < 27; ++i) {
			if constexpr (std::is_same_v<T, int>) {
				aK[i] = __builtin_amdgcn_readfirstlane(static_cast<int>(raw[i]));
			} else {
				aK[i] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int,
					static_cast<float>(raw[i]))));
			}
		}
	});
}

// March the lane's column through the z chunk with T arithmetic: float fma
// chains (FP32 sums, or exact integer sums below 2^24) or int32 24-bit
// multiply-adds. EXACT: the sums are exact integers (converted like the
// integer streaming kernel); otherwise FP32 sums converted like the others.
template<class S, class D, class T, bool EXACT, int R, int P>
__device__ __forceinline__ void column_march(const conv_args& a, const column_geom& g, const T (&k)[27])
{
	using io = column_io<S>;
	using raw_t = typename io::raw;
	constexpr int rows = R + 2;
	constexpr int ss = sizeof(S);
	constexpr int ds = sizeof(D);

	// Byte offsets of the lane's vector and of its x-edge sample in each input
	// row (kOob where the zero border makes the row zero or the lane has none).
	int row_off[rows];
	int edge_off[rows];
	int edge_x = -1;
	if (g.tx == 0 && g.x0 > 0) {
		edge_x = g.x0 - 1;
	} else if (g.tx == 15 && g.x0 + 64 < g.ext_x) {
		edge_x = g.x0 + 64;
	}
#pragma unroll
	for (int j = 0; j < rows; ++j) {
		int yi = g.yb - 1 + j;
		const bool in = border_coord(yi, g.ext_y, a.border);
		const long long base = static_cast<long long>(yi) * a.src.stride[1];
		row_off[j] = in && g.x < g.ext_x ? static_cast<int>(base + g.x * ss) : kOob;
		edge_off[j] = in && edge_x >= 0 ? static_cast<int>(base + edge_x * ss) : kOob;
	}

	const auto load_plane = [&](int aPlane, bool aWanted, raw_t (&aVec)[rows], unsigned (&aEdge)[rows]) {
		int zi = aPlane;
		const bool in = aWanted && border_coord(zi, g.ext_z, a.border);
		const auto plane = stream_resource(a.src.ptr + static_cast<long long>(in ? zi : 0) * a.src.stride[2],
			in ? a.src_span : 0);
#pragma unroll
		for (int j = 0; j < rows; ++j) {
			aVec[j] = io::load(plane, row_off[j]);
			aEdge[j] = io::load_single(plane, edge_off[j]);
		}
	};

	// Samples x - 1 .. x + 4 of one input row.
	const auto window = [&](const raw_t& aVec, unsigned aEdge, T (&aOut)[6]) {
		T s[4];
		io::unpack(aVec, s);
		const T edge = io::template single<T>(aEdge);
		const T beyond_left = g.clamp ? s[0] : T(0);
		const T beyond_right = g.clamp ? s[3] : T(0);
		aOut[0] = column_from_left(g.x0 == 0 ? beyond_left : edge, s[3]);
		const T right = column_from_right(g.x + 4 < g.ext_x ? edge : beyond_right, s[0]);
		aOut[5] = g.x + 4 < g.ext_x ? right : beyond_right;
		aOut[1] = s[0];
		aOut[2] = s[1];
		aOut[3] = s[2];
		aOut[4] = s[3];
	};

	T acc[3][R][4];
#pragma unroll
	for (int set = 0; set < 3; ++set) {
#pragma unroll
		for (int r = 0; r < R; ++r) {
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				acc[set][r][e] = T(0);
			}
		}
	}

	// Add one input plane's taps of kernel plane qz to an accumulator set
	// (rows and x taps from the back, as in every direct kernel).
	const auto add_plane = [&](const T (&aWin)[rows][6], T (&aAcc)[R][4], int aQz) {
#pragma unroll
		for (int r = 0; r < R; ++r) {
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				T sum = aAcc[r][e];
#pragma unroll
				for (int qy = 0; qy < 3; ++qy) {
#pragma unroll
					for (int qx = 0; qx < 3; ++qx) {
						const T sample = aWin[r + 2 - qy][e + 2 - qx];
						const T kk = k[qx + 3 * qy + 9 * aQz];
						if constexpr (std::is_same_v<T, int>) {
							sum += __mul24(sample, kk);
						} else {
							sum = fma(sample, kk, sum);
						}
					}
				}
				aAcc[r][e] = sum;
			}
		}
	};

	const auto store_plane = [&](int aZ, T (&aAcc)[R][4]) {
		const auto plane = stream_resource(a.dst.ptr + static_cast<long long>(aZ) * a.dst.stride[2], a.dst_span);
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int y = g.yb + r;
			if (g.x < g.ext_x && y < g.ext_y) {
				D out[4];
#pragma unroll
				for (int e = 0; e < 4; ++e) {
					if constexpr (EXACT) {
						out[e] = convert_exact_int<D>(static_cast<int>(aAcc[r][e]));
					} else {
						out[e] = convert_out<D>(aAcc[r][e]);
					}
				}
				column_store(plane, static_cast<int>(y * a.dst.stride[1] + g.x * ds), out);
			}
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				aAcc[r][e] = T(0);
			}
		}
	};

	// Input planes first, first + step, .. (n of them: z0 - 1 .. z1 marching up,
	// z1 .. z0 - 1 marching down). Plane p = first + step i adds kernel plane 2
	// to output p + step (set i % 3), kernel plane 1 to output p (set
	// (i + 2) % 3) and kernel plane 0 to output p - step (set (i + 1) % 3),
	// which is then complete; a downward march takes the kernel with its z
	// planes swapped (only exact integer sums, whose order is free). The loop is
	// unrolled by 6 (a multiple of the ring and of P) so sets and staging slots
	// are registers; the skipped terms are those of outputs outside the chunk.
	const int step = g.step;
	const int first = step > 0 ? g.z0 - 1 : g.z1;
	const int n = g.z1 - g.z0 + 2;
	raw_t vec[P][rows];
	unsigned edge[P][rows];
#pragma unroll
	for (int q = 0; q < P; ++q) {
		load_plane(first + step * q, q < n, vec[q], edge[q]);
	}
	for (int i0 = 0; i0 < n; i0 += 6) {
#pragma unroll
		for (int ph = 0; ph < 6; ++ph) {
			const int i = i0 + ph;
			if (i < n) {
				T win[rows][6];
#pragma unroll
				for (int j = 0; j < rows; ++j) {
					window(vec[ph % P][j], edge[ph % P][j], win[j]);
				}
				const int p = first + step * i;
				load_plane(p + step * P, i + P < n, vec[ph % P], edge[ph % P]);
				if (column_in_chunk(g, p + step)) {
					add_plane(win, acc[ph % 3], 2);
				}
				if (column_in_chunk(g, p)) {
					add_plane(win, acc[(ph + 2) % 3], 1);
				}
				if (column_in_chunk(g, p - step)) {
					add_plane(win, acc[(ph + 1) % 3], 0);
					store_plane(p - step, acc[(ph + 1) % 3]);
				}
			}
		}
	}
}

// u16 images with coefficients in [0, 2^16) whose sums stay below 2^31: the
// samples stay packed two per dword and every tap pair is one v_dot2_u32_u16
// into an exact u32 sum. With kernel row (qy, qz) packed as B10 = {k1, k0},
// B21 = {k2, k1}, B2h = {0, k2}, B0l = {k0, 0} (low, high half; k_q the tap
// qx = q, which meets sample x + e + 1 - q), the lane's pairs d0 = {s0, s1} and
// d1 = {s2, s3}, the left lane's d1 (high half s-1) and the right lane's d0
// (low half s4) give the four outputs of a row in eight instructions:
//   e = 0: d0.B10 + L.B2h    e = 1: d0.B21 + d1.B0l
//   e = 2: d1.B10 + d0.B2h   e = 3: d1.B21 + R.B0l
// Loads, borders, lane shifts and the three-plane accumulator ring are those
// of column_march.
typedef unsigned short column_u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned column_dot2(unsigned aPair, unsigned aCoeff, unsigned aSum)
{
	return __builtin_amdgcn_udot2(__builtin_bit_cast(column_u16x2, aPair), __builtin_bit_cast(column_u16x2, aCoeff),
		aSum, false);
}

template<class D, int R, int P>
__device__ __forceinline__ void column_march_dot2(const conv_args& a, const column_geom& g, const int (&k)[27])
{
	using io = column_io<uint16_t>;
	constexpr int rows = R + 2;
	constexpr int ss = 2;
	constexpr int ds = sizeof(D);

	unsigned b10[9], b21[9], b2h[9], b0l[9];  // kernel row qy + 3 qz
#pragma unroll
	for (int row = 0; row < 9; ++row) {
		const unsigned k0 = static_cast<unsigned>(k[3 * row]);
		const unsigned k1 = static_cast<unsigned>(k[3 * row + 1]);
		const unsigned k2 = static_cast<unsigned>(k[3 * row + 2]);
		b10[row] = k1 | (k0 << 16);
		b21[row] = k2 | (k1 << 16);
		b2h[row] = k2 << 16;
		b0l[row] = k0;
	}

	int row_off[rows];
	int edge_off[rows];
	int edge_x = -1;
	if (g.tx == 0 && g.x0 > 0) {
		edge_x = g.x0 - 1;
	} else if (g.tx == 15 && g.x0 + 64 < g.ext_x) {
		edge_x = g.x0 + 64;
	}
#pragma unroll
	for (int j = 0; j < rows; ++j) {
		int yi = g.yb - 1 + j;
		const bool in = border_coord(yi, g.ext_y, a.border);
		const long long base = static_cast<long long>(yi) * a.src.stride[1];
		row_off[j] = in && g.x < g.ext_x ? static_cast<int>(base + g.x * ss) : kOob;
		edge_off[j] = in && edge_x >= 0 ? static_cast<int>(base + edge_x * ss) : kOob;
	}

	const auto load_plane = [&](int aPlane, bool aWanted, column_v2 (&aVec)[rows], unsigned (&aEdge)[rows]) {
		int zi = aPlane;
		const bool in = aWanted && border_coord(zi, g.ext_z, a.border);
		const auto plane = stream_resource(a.src.ptr + static_cast<long long>(in ? zi : 0) * a.src.stride[2],
			in ? a.src_span : 0);
#pragma unroll
		for (int j = 0; j < rows; ++j) {
			aVec[j] = io::load(plane, row_off[j]);
			aEdge[j] = io::load_single(plane, edge_off[j]);
		}
	};

	// The lane's pairs of one input row and its neighbours' (left: high half
	// s-1; right: low half s4), the zero or clamp rule applied at the image ends.
	struct pairs {
		unsigned d0, d1, left, right;
	};
	const auto window = [&](const column_v2& aVec, unsigned aEdge) {
		pairs w;
		w.d0 = aVec.x;
		w.d1 = aVec.y;
		const unsigned beyond_left = g.clamp ? (aVec.x & 0xffffu) : 0u;
		const unsigned beyond_right = g.clamp ? (aVec.y >> 16) : 0u;
		w.left = column_from_left((g.x0 == 0 ? beyond_left : aEdge) << 16, aVec.y);
		const unsigned right = column_from_right(g.x + 4 < g.ext_x ? aEdge : beyond_right, aVec.x);
		w.right = g.x + 4 < g.ext_x ? right : beyond_right;
		return w;
	};

	unsigned acc[3][R][4];
#pragma unroll
	for (int set = 0; set < 3; ++set) {
#pragma unroll
		for (int r = 0; r < R; ++r) {
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				acc[set][r][e] = 0u;
			}
		}
	}

	const auto add_plane = [&](const pairs (&aWin)[rows], unsigned (&aAcc)[R][4], int aQz) {
#pragma unroll
		for (int r = 0; r < R; ++r) {
#pragma unroll
			for (int qy = 0; qy < 3; ++qy) {
				const pairs& w = aWin[r + 2 - qy];
				const int row = qy + 3 * aQz;
				aAcc[r][0] = column_dot2(w.left, b2h[row], column_dot2(w.d0, b10[row], aAcc[r][0]));
				aAcc[r][1] = column_dot2(w.d1, b0l[row], column_dot2(w.d0, b21[row], aAcc[r][1]));
				aAcc[r][2] = column_dot2(w.d0, b2h[row], column_dot2(w.d1, b10[row], aAcc[r][2]));
				aAcc[r][3] = column_dot2(w.right, b0l[row], column_dot2(w.d1, b21[row], aAcc[r][3]));
			}
		}
	};

	const auto store_plane = [&](int aZ, unsigned (&aAcc)[R][4]) {
		const auto plane = stream_resource(a.dst.ptr + static_cast<long long>(aZ) * a.dst.stride[2], a.dst_span);
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int y = g.yb + r;
			if (g.x < g.ext_x && y < g.ext_y) {
				D out[4];
#pragma unroll
				for (int e = 0; e < 4; ++e) {
					out[e] = convert_exact_int<D>(static_cast<int>(aAcc[r][e]));
				}
				column_store(plane, static_cast<int>(y * a.dst.stride[1] + g.x * ds), out);
			}
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				aAcc[r][e] = 0u;
			}
		}
	};

	// the march of column_march: plane p = first + step i adds kernel plane 2
	// to output p + step, plane 1 to output p, plane 0 to output p - step
	const int step = g.step;
	const int first = step > 0 ? g.z0 - 1 : g.z1;
	const int n = g.z1 - g.z0 + 2;
	column_v2 vec[P][rows];
	unsigned edge[P][rows];
#pragma unroll
	for (int q = 0; q < P; ++q) {
		load_plane(first + step * q, q < n, vec[q], edge[q]);
	}
	for (int i0 = 0; i0 < n; i0 += 6) {
#pragma unroll
		for (int ph = 0; ph < 6; ++ph) {
			const int i = i0 + ph;
			if (i < n) {
				pairs win[rows];
#pragma unroll
				for (int j = 0; j < rows; ++j) {
					win[j] = window(vec[ph % P][j], edge[ph % P][j]);
				}
				const int p = first + step * i;
				load_plane(p + step * P, i + P < n, vec[ph % P], edge[ph % P]);
				if (column_in_chunk(g, p + step)) {
					add_plane(win, acc[ph % 3], 2);
				}
				if (column_in_chunk(g, p)) {
					add_plane(win, acc[(ph + 2) % 3], 1);
				}
				if (column_in_chunk(g, p - step)) {
					add_plane(win, acc[(ph + 1) % 3], 0);
					store_plane(p - step, acc[(ph + 1) % 3]);
				}
			}
		}
	}
}

// Integer images whose sums fit neither FP32 nor int32 exactly (coefficients
// of 2^23 and more, or huge sums): every output of the lane is the FP64 chain
// in the reference's order, read straight from memory (compact loops).
template<class S, class D, int R>
__device__ void column_fp64_chain(const conv_args& a, const column_geom& g)
{
	constexpr int ds = sizeof(D);
#pragma unroll 1
	for (int z = g.z0; z < g.z1; ++z) {
#pragma unroll 1
		for (int r = 0; r < R; ++r) {
			const int y = g.yb + r;
			if (g.x >= g.ext_x || y >= g.ext_y) {
				continue;
			}
			D out[4];
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				double sum = 0.0;
#pragma unroll 1
				for (int qz = 0; qz < 3; ++qz) {
#pragma unroll 1
					for (int qy = 0; qy < 3; ++qy) {
#pragma unroll 1
						for (int qx = 0; qx < 3; ++qx) {
							int xi = g.x + e + 1 - qx;
							int yi = y + 1 - qy;
							int zi = z + 1 - qz;
							const bool inside = border_coord(xi, g.ext_x, a.border) &&
								border_coord(yi, g.ext_y, a.border) && border_coord(zi, g.ext_z, a.border);
							const double sample = inside ? static_cast<double>(*reinterpret_cast<const S*>(
								a.src.ptr + xi * a.src.stride[0] + yi * a.src.stride[1] + zi * a.src.stride[2])) :
								0.0;
							const int kk = load_as_int(a.ker.ptr + qx * a.ker.stride[0] + qy * a.ker.stride[1] +
								qz * a.ker.stride[2], a.ker.type);
							sum = fma(sample, static_cast<double>(kk), sum);
						}
					}
				}
				out[e] = convert_out<D>(sum);
			}
			const auto plane = stream_resource(a.dst.ptr + static_cast<long long>(z) * a.dst.stride[2], a.dst_span);
			column_store(plane, static_cast<int>(y * a.dst.stride[1] + g.x * ds), out);
		}
	}
}

template<class S, class D, bool INTEGER, int R, int P>
__global__ __launch_bounds__(kColumnThreads) void conv_column_kernel(conv_args a, int aChunk)
{
	const column_geom g = column_geom::make<R>(a, aChunk);
	if constexpr (!INTEGER) {
		float k[27];
		column_coefficients(a, k);
		column_march<S, D, float, false, R, P>(a, g, k);
	} else {
		// Every partial sum is bounded by max source x sum |k|. Below 2^24 the
		// FP32 fma chain is exact (integer products and sums), below 2^31 with
		// |k| < 2^23 the 24-bit integer multiply-adds are; both give the
		// integer the reference's FP64 chain represents.
		int k[27];
		column_coefficients(a, k);
		// Exact sums do not depend on the order of their terms, so odd chunks
		// march down (the kernel's z planes swapped): two neighbouring chunks
		// then read their two shared planes at the same time, at both ends,
		// where the second read finds them in L2 (marching the same way, one
		// reads them a whole chunk after the other).
		column_geom gz = g;
		if (a.column_alternate && (g.chunk & 1)) {
			gz.step = -1;
			int flipped[27];
#pragma unroll
			for (int i = 0; i < 27; ++i) {
				flipped[i] = k[i % 9 + 9 * (2 - i / 9)];
			}
#pragma unroll
			for (int i = 0; i < 27; ++i) {
				k[i] = flipped[i];
			}
		}
		long long total = 0;
		long long largest = 0;
#pragma unroll
		for (int i = 0; i < 27; ++i) {
			const long long m = k[i] < 0 ? -static_cast<long long>(k[i]) : static_cast<long long>(k[i]);
			total += m;
			largest = m > largest ? m : largest;
		}
		constexpr long long source_max = sizeof(S) == 1 ? 255 : 65535;
		bool small_nonnegative = true;
#pragma unroll
		for (int i = 0; i < 27; ++i) {
			small_nonnegative = small_nonnegative && k[i] >= 0 && k[i] < (1 << 16);
		}
		if (std::is_same_v<S, uint16_t> && small_nonnegative && total * source_max < (1LL << 31) &&
			a.column_dot2) {
			if constexpr (std::is_same_v<S, uint16_t>) {
				column_march_dot2<D, R, P>(a, gz, k);
			}
		} else if (total * source_max < (1LL << 24)) {
			float kf[27];
#pragma unroll
			for (int i = 0; i < 27; ++i) {
				kf[i] = static_cast<float>(k[i]);
			}
			column_march<S, D, float, true, R, P>(a, gz, kf);
		} else if (largest < (1LL << 23) && total * source_max < (1LL << 31)) {
			column_march<S, D, int, true, R, P>(a, gz, k);
		} else {
			column_fp64_chain<S, D, R>(a, g);
		}
	}
}

template<bool INTEGER, class S, class D, int KX, int KYZ>
void launch_stream_shape(const conv_args& a, int aChunk, unsigned aBlocks, size_t aLds, hipStream_t aStream)
{
	if constexpr (INTEGER) {
		conv_stream_int_kernel<S, D, KX, KYZ><<<aBlocks, kThreads, aLds, aStream>>>(a, aChunk);
	} else {
		conv_stream_kernel<S, D, KX, KYZ><<<aBlocks, kThreads, aLds, aStream>>>(a, aChunk);
	}
}

// 3^3 kernels get fully unrolled tap loops with scalar-register coefficients.
template<bool INTEGER, class S, class D, int KX>
void launch_stream_kx(const conv_args& a, int aChunk, unsigned aBlocks, size_t aLds, hipStream_t aStream)
{
	const bool cubic = a.ker.extent[1] == KX && a.ker.extent[2] == KX && tune_get("conv.stream_unroll", 1) != 0;
	if constexpr (KX == 3) {
		if (cubic) {
			launch_stream_shape<INTEGER, S, D, KX, KX>(a, aChunk, aBlocks, aLds, aStream);
			return;
		}
	}
	launch_stream_shape<INTEGER, S, D, KX, 0>(a, aChunk, aBlocks, aLds, aStream);
}

template<bool INTEGER, class S, class D>
void launch_stream_types(const conv_args& a, int aChunk, unsigned aBlocks, size_t aLds, hipStream_t aStream)
{
	switch (a.ker.extent[0]) {
	case 1: launch_stream_kx<INTEGER, S, D, 1>(a, aChunk, aBlocks, aLds, aStream); break;
	case 3: launch_stream_kx<INTEGER, S, D, 3>(a, aChunk, aBlocks, aLds, aStream); break;
	case 5: launch_stream_kx<INTEGER, S, D, 5>(a, aChunk, aBlocks, aLds, aStream); break;
	default: launch_stream_kx<INTEGER, S, D, 7>(a, aChunk, aBlocks, aLds, aStream); break;
	}
}

template<class S, class D, bool INTEGER, int R>
void launch_column_rows(const conv_args& a, int aChunk, unsigned aBlocks, hipStream_t aStream)
{
	conv_column_kernel<S, D, INTEGER, R, 2><<<aBlocks, kColumnThreads, 0, aStream>>>(a, aChunk);
}

// ---------------------------------------------------------------------------
// Plane kernel: square K x K kernels with one plane (3 <= K <= 9) on images
// whose rows are x-contiguous and aligned for 4-sample vectors. No LDS and no
// barriers: a lane holds 4 x-outputs of R rows; it loads its 4 samples of the
// R + K - 1 input rows with one vector access each and takes the K / 2 samples
// on either side from the neighbouring lanes of its 16-lane DPP row (the lanes
// at the ends of a row load the vector beyond the tile). Input rows are used
// from the last down, so every output's fma chain runs kernel rows ascending
// and x taps ascending (input rows and samples from the back): the order of
// the tiled kernel, whose bits it reproduces: the reference's FP64 chain in
// REFERENCE precision, FP32 sums in FP32, exact sums for integer images.
enum plane_mode { PLANE_FP32, PLANE_REFERENCE, PLANE_INTEGER };

// The K x K coefficients as T (converted like load_as_float / load_as_double:
// one rounding from the stored type), with the element type dispatched once.
template<class T, int K>
__device__ __forceinline__ void plane_coefficients(const conv_args& a, T (&aK)[K][K])
{
	with_element_type(a.ker.type, [&](auto aType) {
		using E = decltype(aType);
#pragma unroll
		for (int qy = 0; qy < K; ++qy) {
#pragma unroll
			for (int qx = 0; qx < K; ++qx) {
				const E v = *reinterpret_cast<const E*>(a.ker.ptr + qx * a.ker.stride[0] + qy * a.ker.stride[1]);
				if constexpr (std::is_same_v<T, int>) {
					aK[qy][qx] = static_cast<int>(v);
				} else {
					aK[qy][qx] = static_cast<T>(v);
				}
			}
		}
	});
}

template<class S, class D, class T, bool EXACT, int K, int R>
__device__ __forceinline__ void plane_rows(const conv_args& a, int aPlane, int aTileX, int aYb)
{
	using io = column_io<S>;
	using raw_t = typename io::raw;
	using U = std::conditional_t<std::is_same_v<S, float>, float, int>;  // a sample in 32 bits
	constexpr int C = K / 2;
	constexpr int rows = R + K - 1;
	constexpr int ss = sizeof(S);
	constexpr int ds = sizeof(D);
	const int lane = static_cast<int>(threadIdx.x) % 64;
	const int tx = lane % 16;
	const int ext_x = a.src.extent[0];
	const int ext_y = a.src.extent[1];
	const int x0 = aTileX * 64;
	const int x = x0 + 4 * tx;
	const bool clamp = a.border == IANE_BORDER_CLAMP;

	T k[K][K];
	plane_coefficients<T, K>(a, k);
	// the vector beyond the tile for the lanes at the ends of a row
	int edge_x = -1;
	if (tx == 0 && x0 > 0) {
		edge_x = x0 - 4;
	} else if (tx == 15 && x0 + 64 < ext_x) {
		edge_x = x0 + 64;
	}
	const auto src = stream_resource(a.src.ptr + static_cast<long long>(aPlane) * a.src.stride[2], a.src_span);
	raw_t vec[rows];
	raw_t edge[rows];
#pragma unroll
	for (int j = 0; j < rows; ++j) {
		int yi = aYb - C + j;
		const bool in = border_coord(yi, ext_y, a.border);
		const long long base = static_cast<long long>(yi) * a.src.stride[1];
		vec[j] = io::load(src, in && x < ext_x ? static_cast<int>(base + x * ss) : kOob);
		edge[j] = io::load(src, in && edge_x >= 0 ? static_cast<int>(base + edge_x * ss) : kOob);
	}

	T acc[R][4];
#pragma unroll
	for (int r = 0; r < R; ++r) {
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			acc[r][e] = T(0);
		}
	}
#pragma unroll
	for (int j = rows - 1; j >= 0; --j) {
		// samples x - C .. x + 3 + C of input row j (the border rule at the image ends)
		U mine[4];
		U beyond[4];
		io::unpack(vec[j], mine);
		io::unpack(edge[j], beyond);
		const U before = clamp ? mine[0] : U(0);
		const U after = clamp ? mine[3] : U(0);
		T w[4 + 2 * C];
#pragma unroll
		for (int c = 0; c < C; ++c) {
			const int from = 4 - C + c;  // component of the left lane's vector
			const U left = column_from_left(x0 == 0 ? before : beyond[from], mine[from]);
			w[c] = static_cast<T>(left);
			const U right = column_from_right(x + 4 < ext_x ? beyond[c] : after, mine[c]);
			w[C + 4 + c] = static_cast<T>(x + 4 < ext_x ? right : after);
		}
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			w[C + e] = static_cast<T>(mine[e]);
		}
		// outputs r meet row j through kernel row qy = r + K - 1 - j
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int qy = r + K - 1 - j;
			if (qy >= 0 && qy < K) {
#pragma unroll
				for (int e = 0; e < 4; ++e) {
#pragma unroll
					for (int qx = 0; qx < K; ++qx) {
						if constexpr (std::is_same_v<T, int>) {
							acc[r][e] += __mul24(w[e + K - 1 - qx], k[qy][qx]);
						} else {
							acc[r][e] = fma(w[e + K - 1 - qx], k[qy][qx], acc[r][e]);
						}
					}
				}
			}
		}
	}

	const auto dst = stream_resource(a.dst.ptr + static_cast<long long>(aPlane) * a.dst.stride[2], a.dst_span);
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const int y = aYb + r;
		if (x < ext_x && y < ext_y) {
			D out[4];
#pragma unroll
			for (int e = 0; e < 4; ++e) {
				if constexpr (EXACT) {
					out[e] = convert_out<D>(static_cast<double>(acc[r][e]));
				} else {
					out[e] = convert_out<D>(acc[r][e]);
				}
			}
			column_store(dst, static_cast<int>(y * a.dst.stride[1] + x * ds), out);
		}
	}
}

// Integer images whose sums fit neither FP32 nor int32 exactly: each of the
// lane's outputs is the FP64 chain in the reference's order, read straight
// from memory (compact loops).
template<class S, class D, int K, int R>
__device__ void plane_fp64_chain(const conv_args& a, int aPlane, int aTileX, int aYb)
{
	constexpr int C = K / 2;
	constexpr int ds = sizeof(D);
	const int x = aTileX * 64 + 4 * (static_cast<int>(threadIdx.x) % 16);
	const char* src = a.src.ptr + static_cast<long long>(aPlane) * a.src.stride[2];
	const auto dst = stream_resource(a.dst.ptr + static_cast<long long>(aPlane) * a.dst.stride[2], a.dst_span);
#pragma unroll 1
	for (int r = 0; r < R; ++r) {
		const int y = aYb + r;
		if (x >= a.src.extent[0] || y >= a.src.extent[1]) {
			continue;
		}
		D out[4];
#pragma unroll
		for (int e = 0; e < 4; ++e) {
			double sum = 0.0;
#pragma unroll 1
			for (int qy = 0; qy < K; ++qy) {
#pragma unroll 1
				for (int qx = 0; qx < K; ++qx) {
					int xi = x + e + C - qx;
					int yi = y + C - qy;
					const bool inside = border_coord(xi, a.src.extent[0], a.border) &&
						border_coord(yi, a.src.extent[1], a.border);
					const double sample = inside ? static_cast<double>(*reinterpret_cast<const S*>(
						src + xi * a.src.stride[0] + yi * a.src.stride[1])) : 0.0;
					sum = fma(sample, static_cast<double>(load_as_int(a.ker.ptr + qx * a.ker.stride[0] +
						qy * a.ker.stride[1], a.ker.type)), sum);
				}
			}
			out[e] = convert_out<D>(sum);
		}
		column_store(dst, static_cast<int>(y * a.dst.stride[1] + x * ds), out);
	}
}

template<class S, class D, int MODE, int K, int R>
__global__ __launch_bounds__(kColumnThreads) void conv_plane_kernel(conv_args a)
{
	const unsigned block = xcd_remap(blockIdx.x, gridDim.x);
	const int tile_x = static_cast<int>(block % a.tiles_x);
	const unsigned rest = block / a.tiles_x;
	const int tile_y = static_cast<int>(rest % a.tiles_y);
	const int plane = static_cast<int>(rest / a.tiles_y);
	const int group = static_cast<int>(threadIdx.x) / 64 * 4 + static_cast<int>(threadIdx.x) % 64 / 16;
	const int yb = tile_y * (16 * R) + group * R;
	if constexpr (MODE == PLANE_FP32) {
		plane_rows<S, D, float, false, K, R>(a, plane, tile_x, yb);
	} else if constexpr (MODE == PLANE_REFERENCE) {
		plane_rows<S, D, double, false, K, R>(a, plane, tile_x, yb);
	} else {
		// Integer samples and coefficients: every partial sum is bounded by
		// max source x sum |k|. Below 2^24 the FP32 chain is exact, below 2^31
		// with |k| < 2^23 the 24-bit integer multiply-adds are; both give the
		// integer of the reference's FP64 chain. Beyond, each output is that
		// FP64 chain read straight from memory.
		int k[K][K];
		plane_coefficients<int, K>(a, k);
		long long total = 0;
		long long largest = 0;
#pragma unroll
		for (int qy = 0; qy < K; ++qy) {
#pragma unroll
			for (int qx = 0; qx < K; ++qx) {
				const int kk = k[qy][qx];
				const long long m = kk < 0 ? -static_cast<long long>(kk) : static_cast<long long>(kk);
				total += m;
				largest = m > largest ? m : largest;
			}
		}
		constexpr long long source_max = sizeof(S) == 1 ? 255 : 65535;
		if (total * source_max < (1LL << 24)) {
			plane_rows<S, D, float, true, K, R>(a, plane, tile_x, yb);
		} else if (largest < (1LL << 23) && total * source_max < (1LL << 31)) {
			plane_rows<S, D, int, true, K, R>(a, plane, tile_x, yb);
		} else {
			plane_fp64_chain<S, D, K, R>(a, plane, tile_x, yb);
		}
	}
}

// Rows per lane the kernel instantiates: FP64 or integer sums of 7 x 7 kernels
// keep at most 2 (4 would not fit the register file).
template<int MODE, int K>
constexpr int plane_max_rows() { return MODE != PLANE_FP32 && K >= 7 ? 2 : 4; }

template<class S, class D, int MODE, int K>
void launch_plane_rows(const conv_args& a, int aRows, unsigned aBlocks, hipStream_t aStream)
{
	if (aRows == 4) {
		if constexpr (plane_max_rows<MODE, K>() >= 4) {
			conv_plane_kernel<S, D, MODE, K, 4><<<aBlocks, kColumnThreads, 0, aStream>>>(a);
		}
	} else if (aRows == 2) {
		conv_plane_kernel<S, D, MODE, K, 2><<<aBlocks, kColumnThreads, 0, aStream>>>(a);
	} else {
		conv_plane_kernel<S, D, MODE, K, 1><<<aBlocks, kColumnThreads, 0, aStream>>>(a);
	}
}

template<class S, class D, int MODE>
void launch_plane_k(const conv_args& a, int aRows, unsigned aBlocks, hipStream_t aStream)
{
	switch (a.ker.extent[0]) {
	case 3: launch_plane_rows<S, D, MODE, 3>(a, aRows, aBlocks, aStream); break;
	case 5: launch_plane_rows<S, D, MODE, 5>(a, aRows, aBlocks, aStream); break;
	case 7: launch_plane_rows<S, D, MODE, 7>(a, aRows, aBlocks, aStream); break;
	default:
		if constexpr (MODE == PLANE_FP32) {
			launch_plane_rows<S, D, MODE, 9>(a, aRows, aBlocks, aStream);
		}
		break;
	}
}

// Compute units of the current device (cached per device).
int compute_units()
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	static std::atomic<int> cache[64];
	if (device >= 0 && device < 64) {
		const int cached = cache[device].load(std::memory_order_relaxed);
		if (cached > 0) {
			return cached;
		}
	}
	int units = 0;
	check_hip(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
	units = std::max(units, 1);
	if (device >= 0 && device < 64) {
		cache[device].store(units, std::memory_order_relaxed);
	}
	return units;
}

// Byte span of a view's (x, y) plane when both strides are non-negative and the
// span fits a buffer resource; -1 otherwise.
int plane_span(const view4& aView)
{
	if (aView.stride[0] < 0 || aView.stride[1] < 0) {
		return -1;
	}
	const long long span = (aView.extent[0] - 1) * aView.stride[0] + (aView.extent[1] - 1) * aView.stride[1] +
		dtype_size(aView.type);
	return span < kOob - 64 ? static_cast<int>(span) : -1;
}

// Streaming launch for rank <= 3 with 2 <= kz <= 7, kx in {1, 3, 5, 7}, ky <= 7 and
// a ring that fits, for FP32 accumulation of f32 -> f32, u8 -> u8, u16 -> u16,
// u8 / u16 -> f32, or integer images (u8, u16) with integer kernels at
// reference precision; returns false when the tiled kernel should run instead.
template<class Acc>
bool launch_stream(conv_args& a, hipStream_t aStream)
{
	const int kx = a.ker.extent[0];
	const int ky = a.ker.extent[1];
	const int kz = a.ker.extent[2];
	if (a.src.extent[3] != 1 || a.ker.extent[3] != 1 || kz < 2 || kz > 7 || ky > 7 ||
		!(kx == 1 || kx == 3 || kx == 5 || kx == 7)) {
		return false;
	}
	const int rows = kStreamTileY + ky - 1;
	const int cols = stream_cols(kx);
	if (rows * cols > kStreamStage * kThreads) {
		return false;
	}
	const int st = a.src.type;
	const int dt = a.dst.type;
	const bool int_source = st == IANE_U8 || st == IANE_U16;
	const bool integer = int_source && (a.ker.type == IANE_U8 || a.ker.type == IANE_U16 || a.ker.type == IANE_I32) &&
		sizeof(Acc) == 8;
	const bool fp32 = sizeof(Acc) == sizeof(float) &&
		((st == IANE_F32 && dt == IANE_F32) || (int_source && (dt == st || dt == IANE_F32)));
	if (!integer && !fp32) {
		return false; // FP64 window reads from LDS: the tiled kernel is faster
	}
	a.src_span = plane_span(a.src);
	a.dst_span = plane_span(a.dst);
	if (a.src_span < 0 || a.dst_span < 0) {
		return false;
	}
	const size_t lds = 4 * (static_cast<size_t>(kz + 2) * rows * cols + static_cast<size_t>(kx) * ky * kz + kThreads);
	if (lds > 96 * 1024) {
		return false;
	}
	a.tiles_x = ceil_div(a.dst.extent[0], kTileX);
	a.tiles_y = ceil_div(a.dst.extent[1], kStreamTileY);
	// One round of workgroups: as many as fit on the chip at once (LDS bound),
	// each streaming an equal z chunk; halo planes re-read per chunk: kz - 1.
	const long long xy = static_cast<long long>(a.tiles_x) * a.tiles_y;
	const long long resident = static_cast<long long>(compute_units()) *
		std::max<long long>(1, std::min<long long>(8, (160 * 1024) / static_cast<long long>(lds)));
	const long long columns = std::max<long long>(1, resident / xy);
	int chunk = static_cast<int>(std::max<long long>(4, ceil_div(a.dst.extent[2], columns)));
	const int forced = tune_get("conv.stream_chunk", 0);
	if (forced > 0) {
		chunk = std::min(forced, 64);
	}
	const long long blocks = xy * ceil_div(a.dst.extent[2], chunk);
	if (blocks > 0x7fffffffLL || blocks == 0) {
		return false;
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	if (!integer) {
		if (st == IANE_F32) {
			launch_stream_types<false, float, float>(a, chunk, nb, lds, aStream);
		} else if (st == IANE_U8) {
			if (dt == IANE_U8) {
				launch_stream_types<false, uint8_t, uint8_t>(a, chunk, nb, lds, aStream);
			} else {
				launch_stream_types<false, uint8_t, float>(a, chunk, nb, lds, aStream);
			}
		} else if (dt == IANE_U16) {
			launch_stream_types<false, uint16_t, uint16_t>(a, chunk, nb, lds, aStream);
		} else {
			launch_stream_types<false, uint16_t, float>(a, chunk, nb, lds, aStream);
		}
		return true;
	}
	const auto with_dst = [&](auto aSource) {
		using S = decltype(aSource);
		switch (dt) {
		case IANE_U8: launch_stream_types<true, S, uint8_t>(a, chunk, nb, lds, aStream); break;
		case IANE_U16: launch_stream_types<true, S, uint16_t>(a, chunk, nb, lds, aStream); break;
		case IANE_I32: launch_stream_types<true, S, int32_t>(a, chunk, nb, lds, aStream); break;
		case IANE_F32: launch_stream_types<true, S, float>(a, chunk, nb, lds, aStream); break;
		default: launch_stream_types<true, S, double>(a, chunk, nb, lds, aStream); break;
		}
	};
	if (st == IANE_U8) {
		with_dst(uint8_t{});
	} else {
		with_dst(uint16_t{});
	}
	return true;
}

// Column launch for 3 x 3 x 3 kernels on rank-3 images with x-contiguous rows
// aligned for 4-sample vectors (source and destination), an x extent that is a
// multiple of 4, and the type pairs f32 -> f32 (FP32), or u8 / u16 with an
// integer kernel into i32, f32 or their own type (exact); returns false when
// the streaming or tiled kernel should run instead.
template<class Acc>
bool launch_column(conv_args& a, hipStream_t aStream)
{
	if (tune_get("conv.column", 1) == 0) {
		return false;
	}
	const view4& s = a.src;
	const view4& d = a.dst;
	if (a.ker.extent[0] != 3 || a.ker.extent[1] != 3 || a.ker.extent[2] != 3 || a.ker.extent[3] != 1 ||
		s.extent[3] != 1 || s.extent[0] % 4 != 0) {
		return false;
	}
	const int st = s.type;
	const int dt = d.type;
	const bool int_source = st == IANE_U8 || st == IANE_U16;
	const bool integer = int_source && sizeof(Acc) == 8 &&
		(a.ker.type == IANE_U8 || a.ker.type == IANE_U16 || a.ker.type == IANE_I32) &&
		(dt == st || dt == IANE_I32 || dt == IANE_F32);
	const bool fp32 = sizeof(Acc) == 4 && st == IANE_F32 && dt == IANE_F32;
	if (!integer && !fp32) {
		return false;
	}
	const long long ss = dtype_size(st);
	const long long ds = dtype_size(dt);
	const auto aligned = [](const view4& v, long long aVector) {
		const auto address = reinterpret_cast<uintptr_t>(v.ptr);
		return address % aVector == 0 && v.stride[1] % aVector == 0 && v.stride[2] % aVector == 0;
	};
	if (s.stride[0] != ss || d.stride[0] != ds || !aligned(s, 4 * ss) || !aligned(d, 4 * ds)) {
		return false;
	}
	a.src_span = plane_span(s);
	a.dst_span = plane_span(d);
	if (a.src_span < 0 || a.dst_span < 0) {
		return false;
	}
	a.column_dot2 = tune_get("conv.column_dot2", 1) != 0;
	a.column_alternate = tune_get("conv.column_alternate", 1) != 0;
	const int r = tune_get("conv.column_rows", 4) == 2 ? 2 : 4;
	a.tiles_x = ceil_div(s.extent[0], 64);
	a.tiles_y = ceil_div(s.extent[1], 16 * r);
	// Up to four workgroups per compute unit in one round, each marching an
	// equal z chunk of at least 8 planes: the chunk's two halo planes are read
	// again by its neighbour (from HBM: 256^3 f32 measured 1.25 x the input at
	// 8 planes), while longer chunks leave compute units idle (8 to 16 planes
	// measured fastest on 256^3).
	const long long xy = static_cast<long long>(a.tiles_x) * a.tiles_y;
	const long long want = 4LL * compute_units();
	const long long chunks = std::max<long long>(1, ceil_div(want, xy));
	int chunk = static_cast<int>(std::max<long long>(8, ceil_div(s.extent[2], chunks)));
	const int forced = tune_get("conv.column_chunk", 0);
	if (forced > 0) {
		chunk = forced;
	}
	const long long blocks = xy * ceil_div(s.extent[2], chunk);
	if (blocks > 0x7fffffffLL || blocks == 0) {
		return false;
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	const auto go = [&](auto aSource, auto aDest, auto aInteger) {
		using S = decltype(aSource);
		using D = decltype(aDest);
		constexpr bool I = decltype(aInteger)::value;
		if (r == 4) {
			launch_column_rows<S, D, I, 4>(a, chunk, nb, aStream);
		} else {
			launch_column_rows<S, D, I, 2>(a, chunk, nb, aStream);
		}
	};
	if (fp32) {
		go(float{}, float{}, std::false_type{});
		return true;
	}
	const auto with_dst = [&](auto aSource) {
		using S = decltype(aSource);
		if (dt == IANE_I32) {
			go(S{}, int32_t{}, std::true_type{});
		} else if (dt == IANE_F32) {
			go(S{}, float{}, std::true_type{});
		} else {
			go(S{}, S{}, std::true_type{});
		}
	};
	if (st == IANE_U8) {
		with_dst(uint8_t{});
	} else {
		with_dst(uint16_t{});
	}
	return true;
}

// Plane launch for square single-plane kernels (K in 3..7, 9 with FP32 sums) on
// images of one or more planes (rank <= 3) with x-contiguous rows aligned for
// 4-sample vectors and an x extent that is a multiple of 4, for f32 -> f32
// (FP32 or REFERENCE accumulation) and u8 / u16 with an integer kernel into
// i32 or their own type; returns false when another kernel should run.
template<class Acc>
bool launch_plane(conv_args& a, hipStream_t aStream)
{
	if (tune_get("conv.plane", 1) == 0) {
		return false;
	}
	const view4& s = a.src;
	const view4& d = a.dst;
	const int kx = a.ker.extent[0];
	const bool fp32 = sizeof(Acc) == 4;
	if (kx != a.ker.extent[1] || a.ker.extent[2] != 1 || a.ker.extent[3] != 1 || s.extent[3] != 1 ||
		s.extent[0] % 4 != 0 || kx < 3 || kx > (fp32 ? 9 : 7)) {
		return false;
	}
	const int st = s.type;
	const int dt = d.type;
	const bool integer = (st == IANE_U8 || st == IANE_U16) && sizeof(Acc) == 8 &&
		(a.ker.type == IANE_U8 || a.ker.type == IANE_U16 || a.ker.type == IANE_I32) && (dt == st || dt == IANE_I32);
	const bool real = st == IANE_F32 && dt == IANE_F32;
	if (!integer && !real) {
		return false;
	}
	const long long ss = dtype_size(st);
	const long long ds = dtype_size(dt);
	const auto aligned = [](const view4& v, long long aVector) {
		const auto address = reinterpret_cast<uintptr_t>(v.ptr);
		return address % aVector == 0 && v.stride[1] % aVector == 0 && v.stride[2] % aVector == 0;
	};
	if (s.stride[0] != ss || d.stride[0] != ds || !aligned(s, 4 * ss) || !aligned(d, 4 * ds)) {
		return false;
	}
	a.src_span = plane_span(s);
	a.dst_span = plane_span(d);
	if (a.src_span < 0 || a.dst_span < 0) {
		return false;
	}
	// Rows per lane: the most (fewer input rows loaded per output row) that
	// still gives a workgroup per compute unit (1024^2 f32 7 x 7 in a graph:
	// 6.0 us at 1 row, 5.2 us at 2 and 4).
	a.tiles_x = ceil_div(s.extent[0], 64);
	const long long planes = s.extent[2];
	const auto blocks_for = [&](int aRows) {
		return static_cast<long long>(a.tiles_x) * ceil_div(s.extent[1], 16 * aRows) * planes;
	};
	const int max_rows = fp32 || kx < 7 ? 4 : 2;
	int rows = tune_get("conv.plane_rows", 0);
	if (rows != 1 && rows != 2 && rows != 4) {
		const long long want = compute_units();
		rows = blocks_for(4) >= want ? 4 : (blocks_for(2) >= want ? 2 : 1);
	}
	rows = std::min(rows, max_rows);
	a.tiles_y = ceil_div(s.extent[1], 16 * rows);
	const long long blocks = blocks_for(rows);
	if (blocks > 0x7fffffffLL || blocks == 0) {
		return false;
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	if (real) {
		if (fp32) {
			launch_plane_k<float, float, PLANE_FP32>(a, rows, nb, aStream);
		} else {
			launch_plane_k<float, float, PLANE_REFERENCE>(a, rows, nb, aStream);
		}
	} else if (st == IANE_U8) {
		if (dt == IANE_I32) {
			launch_plane_k<uint8_t, int32_t, PLANE_INTEGER>(a, rows, nb, aStream);
		} else {
			launch_plane_k<uint8_t, uint8_t, PLANE_INTEGER>(a, rows, nb, aStream);
		}
	} else {
		if (dt == IANE_I32) {
			launch_plane_k<uint16_t, int32_t, PLANE_INTEGER>(a, rows, nb, aStream);
		} else {
			launch_plane_k<uint16_t, uint16_t, PLANE_INTEGER>(a, rows, nb, aStream);
		}
	}
	return true;
}

template<class Acc, int KX>
void launch_tiled(const conv_args& a, int aRows, size_t aLds, unsigned aBlocks, hipStream_t aStream)
{
	if (aRows == 4) {
		conv_tiled_kernel<Acc, KX, 4><<<aBlocks, kThreads, aLds, aStream>>>(a);
	} else {
		conv_tiled_kernel<Acc, KX, 2><<<aBlocks, kThreads, aLds, aStream>>>(a);
	}
}

template<class Acc>
void launch_conv(conv_args a, hipStream_t aStream, bool aForceSimple)
{
	const int kx = a.ker.extent[0];
	const int ky = a.ker.extent[1];
	// 4 output rows per thread (64-row tiles: each LDS window row feeds more
	// outputs) when the tile fits, else 2 (32-row tiles)
	const auto tile_lds = [&](int aRows) {
		return sizeof(Acc) * (static_cast<size_t>((kThreads / kThreadsX) * aRows + ky - 1) * (kTileX + kx - 1) +
			static_cast<size_t>(kx) * ky);
	};
	const int rows = tile_lds(4) <= kMaxLds ? 4 : 2;
	const size_t lds = tile_lds(rows);
	const long long count = element_count(a.dst);
	if (!aForceSimple && tune_get("conv.stream", 1) != 0 &&
		(launch_column<Acc>(a, aStream) || launch_plane<Acc>(a, aStream) || launch_stream<Acc>(a, aStream))) {
		return;
	}
	if (aForceSimple || lds > kMaxLds) {
		conv_simple_kernel<Acc><<<ceil_div(count, kThreads), kThreads, 0, aStream>>>(a, count);
		return;
	}
	a.tiles_x = ceil_div(a.dst.extent[0], kTileX);
	a.tiles_y = ceil_div(a.dst.extent[1], (kThreads / kThreadsX) * rows);
	const long long blocks = static_cast<long long>(a.tiles_x) * a.tiles_y * a.dst.extent[2] * a.dst.extent[3];
	if (blocks > 0x7fffffffLL) {
		fail(IANE_ERR_UNSUPPORTED, "convolution exceeds the grid size limit");
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	switch (kx) {
	case 1: launch_tiled<Acc, 1>(a, rows, lds, nb, aStream); break;
	case 3: launch_tiled<Acc, 3>(a, rows, lds, nb, aStream); break;
	case 5: launch_tiled<Acc, 5>(a, rows, lds, nb, aStream); break;
	case 7: launch_tiled<Acc, 7>(a, rows, lds, nb, aStream); break;
	case 9: launch_tiled<Acc, 9>(a, rows, lds, nb, aStream); break;
	case 11: launch_tiled<Acc, 11>(a, rows, lds, nb, aStream); break;
	case 13: launch_tiled<Acc, 13>(a, rows, lds, nb, aStream); break;
	case 15: launch_tiled<Acc, 15>(a, rows, lds, nb, aStream); break;
	default: launch_tiled<Acc, 0>(a, rows, lds, nb, aStream); break;
	}
}

} // namespace

// Validation shared by every convolution entry point (convolution.hpp:97-117, transform.hpp:51-63).
void validate_convolution(const view4& aSrc, const view4& aKer, const view4& aDst)
{
	if (aSrc.rank != aKer.rank || aSrc.rank != aDst.rank) {
		invalid("Convolution requires a kernel and images of matching rank");
	}
	for (int i = 0; i < aKer.rank; ++i) {
		if (aKer.extent[i] <= 0 || aKer.extent[i] % 2 == 0) {
			invalid("Convolution requires positive odd kernel dimensions");
		}
	}
	for (int i = 0; i < 4; ++i) {
		if (aSrc.extent[i] != aDst.extent[i]) {
			invalid("Source and destination require matching nonnegative dimensions");
		}
	}
	if (aSrc.type == IANE_C64 || aKer.type == IANE_C64 || aDst.type == IANE_C64) {
		fail(IANE_ERR_UNSUPPORTED, "convolution supports u8, u16, i32, f32 and f64 elements");
	}
}

void convolve_direct(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, int aPrecision,
	hipStream_t aStream, bool aForceSimple)
{
	conv_args a{};
	a.src = aSrc;
	a.ker = aKer;
	a.dst = aDst;
	a.border = aBorder;
	for (int i = 0; i < 4; ++i) {
		a.center[i] = aKer.extent[i] / 2;
	}
	// Integer samples with integer coefficients always accumulate in FP64:
	// that is exact below 2^53, like the reference, whatever the requested precision.
	const bool exact = dtype_integral(aSrc.type) && dtype_integral(aKer.type);
	if (aPrecision == IANE_PREC_FP32 && !exact) {
		launch_conv<float>(a, aStream, aForceSimple);
	} else {
		launch_conv<double>(a, aStream, aForceSimple);
	}
	check_hip(hipGetLastError(), "Convolution launch");
}

} // namespace iane_hip

using namespace iane_hip;

namespace {

struct convolve_call {
	view4 src, ker, dst;
	int precision;
	bool split;  // runs on the split-precision MFMA path
};

convolve_call check_convolve(const iane_view_desc* aSrc, const iane_view_desc* aKernel, const iane_view_desc* aDst,
	int aBorder, int aPrecision)
{
	convolve_call c{make_view4(aSrc, "source"), make_view4(aKernel, "kernel"), make_view4(aDst, "destination"),
		aPrecision, false};
	validate_convolution(c.src, c.ker, c.dst);
	if (aBorder != IANE_BORDER_CLAMP && aBorder != IANE_BORDER_ZERO) {
		invalid("unknown border mode");
	}
	if (aPrecision < IANE_PREC_REFERENCE || aPrecision > IANE_PREC_SPLIT_FP16) {
		invalid("unknown precision mode");
	}
	if (aPrecision == IANE_PREC_SPLIT_BF16 || aPrecision == IANE_PREC_SPLIT_FP16) {
		c.split = split_applicable(c.src, c.ker, c.dst);
		if (!c.split) {
			c.precision = IANE_PREC_FP32;  // shapes the MFMA path does not tile: FP32 accumulation
		}
	}
	return c;
}

size_t convolve_workspace(const convolve_call& c)
{
	if (!c.split || element_count(c.dst) == 0) {
		return 0;
	}
	return split_kernel_bytes(c.ker, c.precision) + split_source_bytes(c.src, c.ker, c.precision);
}

} // namespace

extern "C" int iane_hip_convolve_workspace_bytes(const iane_view_desc* aSrc, const iane_view_desc* aKernel,
	const iane_view_desc* aDst, int aBorder, int aPrecision, size_t* aBytes)
{
	return guarded([&] {
		if (!aBytes) {
			invalid("workspace size output is null");
		}
		*aBytes = convolve_workspace(check_convolve(aSrc, aKernel, aDst, aBorder, aPrecision));
	});
}

extern "C" int iane_hip_convolve(const iane_view_desc* aSrc, const iane_view_desc* aKernel,
	const iane_view_desc* aDst, int aBorder, int aPrecision, void* aWorkspace, size_t aWorkspaceBytes, void* aStream)
{
	return guarded([&] {
		const convolve_call c = check_convolve(aSrc, aKernel, aDst, aBorder, aPrecision);
		if (element_count(c.dst) == 0) {
			return;
		}
		auto stream = static_cast<hipStream_t>(aStream);
		if (c.split) {
			check_workspace(aWorkspace, aWorkspaceBytes, convolve_workspace(c));
			char* kernel_part = static_cast<char*>(aWorkspace);
			split_prepare_kernel(c.ker, c.precision, kernel_part, stream);
			split_convolve_prepared(c.src, c.ker, c.dst, aBorder, c.precision, kernel_part,
				kernel_part + split_kernel_bytes(c.ker, c.precision), stream);
			return;
		}
		convolve_direct(c.src, c.ker, c.dst, aBorder, c.precision, stream, false);
	});
}

// Test hook: the untiled fallback kernel, so tests can compare both paths.
extern "C" int iane_hip_convolve_untiled(const iane_view_desc* aSrc, const iane_view_desc* aKernel,
	const iane_view_desc* aDst, int aBorder, int aPrecision, void* aStream)
{
	return guarded([&] {
		const auto src = make_view4(aSrc, "source");
		const auto ker = make_view4(aKernel, "kernel");
		const auto dst = make_view4(aDst, "destination");
		validate_convolution(src, ker, dst);
		if (aBorder != IANE_BORDER_CLAMP && aBorder != IANE_BORDER_ZERO) {
			invalid("unknown border mode");
		}
		if (element_count(dst) == 0) {
			return;
		}
		convolve_direct(src, ker, dst, aBorder, aPrecision, static_cast<hipStream_t>(aStream), true);
	});
}
