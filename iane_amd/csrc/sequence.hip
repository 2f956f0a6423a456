// Scan and reduce behind iane_hip_sequence_*: the prepared reduce /
// reduce_along / inclusive_scan / exclusive_scan (and *_along) algorithms of
// the reference (include/iane/algorithms/{reduce,scan}.hpp over
// algorithms/detail/sequence.hpp) for device views, with a fixed operator set
// (sum, product, min, max) instead of device functors, which a C ABI cannot
// carry.
//
// The work is the same three stages as sequence.hpp:213-235 (chunk totals,
// line carries, output), laid out for the GPU:
//   row mode (the whole view, or lines along axis 0): a workgroup owns a tile
//     of up to 2048 consecutive elements of one line, staged through LDS with
//     coalesced loads; totals are block reductions, the carries of a line are
//     a block scan of its tile totals, and the output pass rescans each tile
//     from its carry (8 consecutive elements per thread, wavefront shuffles,
//     one LDS exchange across the 4 waves) and stores through LDS again;
//   column mode (lines along axis >= 1): one thread per line, consecutive
//     threads on consecutive lines (adjacent addresses when axis 0 is
//     packed), each folding its line in order; lines are cut into chunks when
//     there are too few of them to fill the GPU.
// Everything is HBM-bound: one read of the source (plus the tile/chunk totals)
// and one write of the destination. A whole-view problem is one line with a
// unit step; a whole view that is not packed is gathered into the plan's
// workspace first (and a scan's result scattered back), so the kernels only
// ever address base + index * step.
#include "common.hpp"

#include <type_traits>

struct iane_sequence_plan {
	int kind;
	int axis;
	int op;
	int acc_type;
	alignas(8) unsigned char init[8];
	int src_type;
	int src_rank;
	int src_extent[4];
	int dst_rank;
	int dst_extent[4];
	bool row;               // row mode (whole or axis 0), else column mode
	long long lines;
	long long length;
	int per;                // row mode: elements per thread of a tile (1..16, power of two); tile = 256 * per
	bool wave_reduce;       // row mode reduction of short lines: one wave per line
	long long tiles;        // row mode: tiles per line
	long long chunk;        // column mode: elements per chunk
	long long chunks;       // column mode: chunks per line
	size_t partials_bytes;  // tile / chunk totals and carries
	bool stage_src;         // whole view, source not packed: gathered into the workspace
	bool stage_dst;         // whole-view scan, destination not packed: scattered from the workspace
	size_t workspace;
};

namespace iane_hip {
namespace {

constexpr int kSeqThreads = 256;

// Element addressing of a sequence problem: element i of line l lives at
// base(l) + i * step, base(l) from the line index over the other axes.
struct seq_map {
	const char* src;
	char* dst;
	int src_type;
	int other_rank;              // axes other than the line axis, in order (0 for a whole view)
	int other_extent[3];
	long long src_other_stride[3];
	long long dst_other_stride[3];
	long long src_step, dst_step; // byte stride along the line
	long long lines, length;
	bool nested;                  // other axes nest like one axis (stride k+1 = stride k x extent k) in both views
};

// Line index -> byte offset of the line's first element (lines < 2^31).
__device__ __forceinline__ long long line_base(const long long* aStrides, const seq_map& m, long long aLine)
{
	if (m.nested) {
		return aLine * aStrides[0];
	}
	unsigned rest = static_cast<unsigned>(aLine);
	long long offset = 0;
	for (int k = 0; k < m.other_rank; ++k) {
		const unsigned extent = static_cast<unsigned>(m.other_extent[k]);
		const unsigned c = rest % extent;
		rest /= extent;
		offset += static_cast<long long>(c) * aStrides[k];
	}
	return offset;
}

__device__ __forceinline__ long long src_offset(const seq_map& m, long long aBase, long long aIndex)
{
	return aBase + aIndex * m.src_step;
}

__device__ __forceinline__ long long dst_offset(const seq_map& m, long long aBase, long long aIndex)
{
	return aBase + aIndex * m.dst_step;
}

// Whole views that are not packed: element i (storage order, axis 0 fastest)
// between the view and a packed copy.
template<int BYTES>
struct raw_element {
	unsigned char b[BYTES];
};

template<int BYTES, bool GATHER>
__global__ __launch_bounds__(kSeqThreads) void stage_kernel(view4 aView, char* aPacked, long long aCount)
{
	const long long i = static_cast<long long>(blockIdx.x) * kSeqThreads + threadIdx.x;
	if (i >= aCount) {
		return;
	}
	long long rest = i;
	long long offset = 0;
	for (int k = 0; k < 4; ++k) {
		const long long c = rest % aView.extent[k];
		rest /= aView.extent[k];
		offset += c * aView.stride[k];
	}
	auto* packed = reinterpret_cast<raw_element<BYTES>*>(aPacked) + i;
	auto* element = reinterpret_cast<raw_element<BYTES>*>(aView.ptr + offset);
	if constexpr (GATHER) {
		*packed = *element;
	} else {
		*element = *packed;
	}
}

template<class Acc>
__device__ __forceinline__ Acc load_value(const char* aPtr, int aType)
{
	switch (aType) {
	case IANE_U8: return static_cast<Acc>(*reinterpret_cast<const uint8_t*>(aPtr));
	case IANE_U16: return static_cast<Acc>(*reinterpret_cast<const uint16_t*>(aPtr));
	case IANE_I32: return static_cast<Acc>(*reinterpret_cast<const int32_t*>(aPtr));
	case IANE_F32: return static_cast<Acc>(*reinterpret_cast<const float*>(aPtr));
	default: return static_cast<Acc>(*reinterpret_cast<const double*>(aPtr));
	}
}

// op(accumulated, next) in the accumulator type. Integer sums and products
// wrap (computed in 32-bit unsigned arithmetic, then narrowed); min and max
// are std::min / std::max.
template<class Acc, int OP>
__device__ __forceinline__ Acc apply(Acc a, Acc b)
{
	if constexpr (OP == IANE_OP_SUM) {
		if constexpr (std::is_integral_v<Acc>) {
			return static_cast<Acc>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
		} else {
			return a + b;
		}
	} else if constexpr (OP == IANE_OP_PRODUCT) {
		if constexpr (std::is_integral_v<Acc>) {
			return static_cast<Acc>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b));
		} else {
			return a * b;
		}
	} else if constexpr (OP == IANE_OP_MIN) {
		return b < a ? b : a;
	} else {
		return a < b ? b : a;
	}
}

// Neutral element: padding for partial tiles and block collectives.
template<class Acc, int OP>
__device__ __forceinline__ Acc identity()
{
	if constexpr (OP == IANE_OP_SUM) {
		return Acc(0);
	} else if constexpr (OP == IANE_OP_PRODUCT) {
		return Acc(1);
	} else if constexpr (OP == IANE_OP_MIN) {
		if constexpr (std::is_floating_point_v<Acc>) {
			return std::numeric_limits<Acc>::infinity();
		} else {
			return std::numeric_limits<Acc>::max();
		}
	} else {
		if constexpr (std::is_floating_point_v<Acc>) {
			return -std::numeric_limits<Acc>::infinity();
		} else {
			return std::numeric_limits<Acc>::lowest();
		}
	}
}

template<class Acc>
__device__ __forceinline__ Acc shfl_up(Acc aValue, int aDelta)
{
	if constexpr (sizeof(Acc) < 4) {
		return static_cast<Acc>(__shfl_up(static_cast<unsigned>(aValue), aDelta, 64));
	} else {
		return __shfl_up(aValue, aDelta, 64);
	}
}

// Exclusive scan over the 256 threads of a workgroup (identity for thread 0);
// `aTotal` receives the combination of all 256 values.
template<class Acc, int OP>
__device__ __forceinline__ Acc block_exclusive_scan(Acc aValue, Acc& aTotal, Acc* aShared)
{
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	Acc inclusive = aValue;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		const Acc other = shfl_up(inclusive, d);
		if (lane >= d) {
			inclusive = apply<Acc, OP>(other, inclusive);
		}
	}
	if (lane == 63) {
		aShared[wave] = inclusive;
	}
	__syncthreads();
	Acc before = identity<Acc, OP>();
	for (int w = 0; w < wave; ++w) {
		before = apply<Acc, OP>(before, aShared[w]);
	}
	aTotal = apply<Acc, OP>(apply<Acc, OP>(apply<Acc, OP>(aShared[0], aShared[1]), aShared[2]), aShared[3]);
	Acc exclusive = shfl_up(inclusive, 1);
	if (lane == 0) {
		exclusive = identity<Acc, OP>();
	}
	__syncthreads();
	return apply<Acc, OP>(before, exclusive);
}

template<class Acc>
__device__ __forceinline__ void store_value(char* aPtr, Acc aValue)
{
	*reinterpret_cast<Acc*>(aPtr) = aValue;
}

// ---- row mode -------------------------------------------------------------
constexpr int kSeqPer = 16;                       // elements per thread of a tile, at most
constexpr int kSeqTile = kSeqThreads * kSeqPer;   // 4096

template<class Acc>
__device__ __forceinline__ Acc shfl_xor(Acc aValue, int aMask)
{
	if constexpr (sizeof(Acc) < 4) {
		return static_cast<Acc>(__shfl_xor(static_cast<unsigned>(aValue), aMask, 64));
	} else {
		return __shfl_xor(aValue, aMask, 64);
	}
}

// aOut[e] <- source element aFirst + e * aStep of a line (e < aPer), the
// neutral element where that index is >= aCount. Every load is issued before
// any is used: indices past the end are clamped to the last element (a valid
// address) and replaced afterwards, so there is no branch between the loads.
// Byte offsets within one load or store span (a tile or a short line) are
// 24-bit multiplies when the step is below 2^22 bytes and every offset of the
// span stays below 2^31 (every packed line and the axis-0 lines of ordinary
// views), instead of 64-bit arithmetic per element.
constexpr long long kSeqSmallStep = 1LL << 22;

__host__ __device__ __forceinline__ bool small_span(long long aStep, long long aCount)
{
	const long long step = aStep < 0 ? -aStep : aStep;
	return step < kSeqSmallStep && aCount < (1LL << 23) && step * aCount < (1LL << 31);
}

template<class Acc, int OP, int PER>
__device__ __forceinline__ void load_line(const seq_map& m, long long aBase, long long aBegin, long long aCount,
	long long aFirst, long long aStep, int aPer, Acc (&aOut)[PER])
{
	const char* span = m.src + src_offset(m, aBase, aBegin);
	const bool small = small_span(m.src_step, aCount);
	const int step = static_cast<int>(m.src_step);
	with_element_type(m.src_type, [&](auto aType) {
		using T = decltype(aType);
		T raw[PER];
		if (small) {
			const int first = static_cast<int>(aFirst);
			const int stride = static_cast<int>(aStep);
			const int last = static_cast<int>(aCount - 1);
#pragma unroll
			for (int e = 0; e < PER; ++e) {
				const int j = min(first + e * stride, last);
				raw[e] = *reinterpret_cast<const T*>(span + __mul24(j, step));
			}
		} else {
#pragma unroll
			for (int e = 0; e < PER; ++e) {
				const long long j = min(aFirst + e * aStep, aCount - 1);
				raw[e] = *reinterpret_cast<const T*>(span + j * m.src_step);
			}
		}
#pragma unroll
		for (int e = 0; e < PER; ++e) {
			const bool valid = e < aPer && aFirst + e * aStep < aCount;
			aOut[e] = valid ? static_cast<Acc>(raw[e]) : identity<Acc, OP>();
		}
	});
}

// One wave per line (lines of at most a few thousand elements): lanes stride
// through the line in batches of PER elements each (PER sized to the line, so
// no lane issues loads past it), then a butterfly over the wave; the result
// init (+) total goes to the destination.
template<class Acc, int OP, int PER>
__global__ __launch_bounds__(kSeqThreads) void row_reduce_wave_kernel(seq_map m, Acc aInit)
{
	const int lane = threadIdx.x & 63;
	const long long line = static_cast<long long>(blockIdx.x) * (kSeqThreads / 64) + (threadIdx.x >> 6);
	if (line >= m.lines) {
		return; // uniform over the wave
	}
	const long long base = line_base(m.src_other_stride, m, line);
	Acc acc = identity<Acc, OP>();
	for (long long start = 0; start < m.length; start += 64 * PER) {
		Acc v[PER];
		load_line<Acc, OP, PER>(m, base, start, m.length - start, lane, 64, PER, v);
#pragma unroll
		for (int e = 0; e < PER; ++e) {
			acc = apply<Acc, OP>(acc, v[e]);
		}
	}
#pragma unroll
	for (int mask = 32; mask > 0; mask >>= 1) {
		acc = apply<Acc, OP>(acc, shfl_xor(acc, mask));
	}
	if (lane == 0) {
		store_value(m.dst + line_base(m.dst_other_stride, m, line), apply<Acc, OP>(aInit, acc));
	}
}

// Short contiguous lines (a step of one element, every line start 16-byte
// aligned, the length a whole number of 16-byte vectors): a wave takes
// 8 / VPL lines; each lane loads vectors lane + 64 i (i < VPL) of every one of
// them before any is used (8 x 16 bytes in flight per lane instead of one
// element per load), folds its vectors in order, then a butterfly per line;
// lane 0 stores init (+) total.
template<class Acc, int OP, int VPL>
__global__ __launch_bounds__(kSeqThreads) void row_reduce_vec_kernel(seq_map m, Acc aInit)
{
	constexpr int kLines = 8 / VPL;
	using vec_t = unsigned __attribute__((ext_vector_type(4)));
	const int lane = threadIdx.x & 63;
	const long long first = (static_cast<long long>(blockIdx.x) * (kSeqThreads / 64) + (threadIdx.x >> 6)) * kLines;
	if (first >= m.lines) {
		return; // uniform over the wave
	}
	with_element_type(m.src_type, [&](auto aType) {
		using T = decltype(aType);
		constexpr int V = 16 / static_cast<int>(sizeof(T));
		struct parts {
			T e[V];
		};
		const int vectors = static_cast<int>(m.length / V);
		vec_t raw[kLines][VPL];
#pragma unroll
		for (int l = 0; l < kLines; ++l) {
			const long long line = min(first + l, m.lines - 1);
			const char* base = m.src + line_base(m.src_other_stride, m, line);
#pragma unroll
			for (int i = 0; i < VPL; ++i) {
				const int v = min(lane + 64 * i, vectors - 1);
				raw[l][i] = *reinterpret_cast<const vec_t*>(base + 16LL * v);
			}
		}
#pragma unroll
		for (int l = 0; l < kLines; ++l) {
			Acc acc = identity<Acc, OP>();
#pragma unroll
			for (int i = 0; i < VPL; ++i) {
				const parts p = __builtin_bit_cast(parts, raw[l][i]);
				const bool valid = lane + 64 * i < vectors;
#pragma unroll
				for (int j = 0; j < V; ++j) {
					acc = apply<Acc, OP>(acc, valid ? static_cast<Acc>(p.e[j]) : identity<Acc, OP>());
				}
			}
#pragma unroll
			for (int mask = 32; mask > 0; mask >>= 1) {
				acc = apply<Acc, OP>(acc, shfl_xor(acc, mask));
			}
			if (lane == 0 && first + l < m.lines) {
				store_value(m.dst + line_base(m.dst_other_stride, m, first + l), apply<Acc, OP>(aInit, acc));
			}
		}
	});
}

// Totals of tiles of 256 * aPer elements (all tiles for a reduction; all but the last for a scan);
// PER >= aPer is the number of loads a thread issues.
template<class Acc, int OP, int PER>
__global__ __launch_bounds__(kSeqThreads) void row_totals_kernel(seq_map m, Acc* aPartials, long long aTiles, int aPer,
	int aSkipLast)
{
	__shared__ Acc shared[4];
	const long long task = blockIdx.x;
	const long long line = task / aTiles;
	const long long t = task - line * aTiles;
	if (aSkipLast && t == aTiles - 1) {
		return; // uniform over the workgroup
	}
	const long long tile = static_cast<long long>(kSeqThreads) * aPer;
	const long long begin = t * tile;
	const long long n = min(tile, m.length - begin);
	const long long base = line_base(m.src_other_stride, m, line);
	Acc v[PER];
	load_line<Acc, OP, PER>(m, base, begin, n, threadIdx.x, kSeqThreads, aPer, v);
	Acc acc = identity<Acc, OP>();
#pragma unroll
	for (int e = 0; e < PER; ++e) {
		acc = apply<Acc, OP>(acc, v[e]);
	}
	Acc total;
	(void)block_exclusive_scan<Acc, OP>(acc, total, shared);
	if (threadIdx.x == 0) {
		aPartials[task] = total;
	}
}

// row_totals_kernel for full-size tiles (kSeqPer elements per thread) of
// contiguous lines of whole 16-byte vectors: thread t folds vectors t, t + 256,
// .. of the tile (each wave load one coalesced 1 KiB run), all loads in flight
// before the first is folded.
template<class Acc, int OP>
__global__ __launch_bounds__(kSeqThreads) void row_totals_vec_kernel(seq_map m, Acc* aPartials, long long aTiles,
	int aSkipLast)
{
	using vec_t = unsigned __attribute__((ext_vector_type(4)));
	__shared__ Acc shared[4];
	const long long task = blockIdx.x;
	const long long line = task / aTiles;
	const long long t = task - line * aTiles;
	if (aSkipLast && t == aTiles - 1) {
		return; // uniform over the workgroup
	}
	const long long begin = t * kSeqTile;
	const int n = static_cast<int>(min(static_cast<long long>(kSeqTile), m.length - begin));
	Acc acc = identity<Acc, OP>();
	with_element_type(m.src_type, [&](auto aType) {
		using T = decltype(aType);
		constexpr int SV = 16 / static_cast<int>(sizeof(T));
		constexpr int NS = kSeqPer / SV;
		struct parts {
			T e[SV];
		};
		const char* src = m.src + line_base(m.src_other_stride, m, line) + begin * static_cast<long long>(sizeof(T));
		const int vectors = n / SV;
		vec_t raw[NS];
#pragma unroll
		for (int i = 0; i < NS; ++i) {
			raw[i] = *reinterpret_cast<const vec_t*>(src + 16 * min(static_cast<int>(threadIdx.x) + kSeqThreads * i,
				vectors - 1));
		}
#pragma unroll
		for (int i = 0; i < NS; ++i) {
			const parts p = __builtin_bit_cast(parts, raw[i]);
			const bool valid = static_cast<int>(threadIdx.x) + kSeqThreads * i < vectors;
#pragma unroll
			for (int j = 0; j < SV; ++j) {
				acc = apply<Acc, OP>(acc, valid ? static_cast<Acc>(p.e[j]) : identity<Acc, OP>());
			}
		}
	});
	Acc total;
	(void)block_exclusive_scan<Acc, OP>(acc, total, shared);
	if (threadIdx.x == 0) {
		aPartials[task] = total;
	}
}

// Carries of one line: slot k <- init (+) totals of tiles 0..k-1 (scans), or the
// line's result (reduction, written to the destination). Each thread takes
// kSeqPer totals: a reduction every 256th (each wave load one coalesced run); a
// scan a consecutive segment, moved between coalesced loads / stores and the
// segments through LDS (as the output pass does).
template<class Acc, int OP, int KIND>
__global__ __launch_bounds__(kSeqThreads) void row_carries_kernel(seq_map m, Acc* aPartials, long long aTiles, Acc aInit)
{
	constexpr bool scan = KIND != IANE_SEQ_REDUCE;
	__shared__ Acc stage[scan ? kSeqTile + kSeqTile / 16 : 1];
	const auto slot = [](int aIndex) { return aIndex + (aIndex >> 4); };
	__shared__ Acc shared[4];
	const int tid = static_cast<int>(threadIdx.x);
	const long long line = blockIdx.x;
	Acc* p = aPartials + line * aTiles;
	const long long read = scan ? aTiles - 1 : aTiles;
	Acc carry = aInit;
	for (long long start = 0; start < aTiles; start += kSeqTile) {
		// clamped loads, masked afterwards: no branch between them
		Acc v[kSeqPer];
#pragma unroll
		for (int e = 0; e < kSeqPer; ++e) {
			v[e] = p[min(start + tid + static_cast<long long>(e) * kSeqThreads, read - 1)];
		}
#pragma unroll
		for (int e = 0; e < kSeqPer; ++e) {
			v[e] = start + tid + static_cast<long long>(e) * kSeqThreads < read ? v[e] : identity<Acc, OP>();
		}
		if constexpr (scan) {
#pragma unroll
			for (int e = 0; e < kSeqPer; ++e) {
				stage[slot(tid + e * kSeqThreads)] = v[e];
			}
			__syncthreads();
#pragma unroll
			for (int e = 0; e < kSeqPer; ++e) {
				v[e] = stage[slot(kSeqPer * tid + e)];
			}
		}
		Acc run = identity<Acc, OP>();
		Acc before_local[kSeqPer];
#pragma unroll
		for (int e = 0; e < kSeqPer; ++e) {
			before_local[e] = run;
			run = apply<Acc, OP>(run, v[e]);
		}
		Acc total;
		const Acc before = block_exclusive_scan<Acc, OP>(run, total, shared);
		if constexpr (scan) {
			const Acc head = apply<Acc, OP>(carry, before);
#pragma unroll
			for (int e = 0; e < kSeqPer; ++e) {
				stage[slot(kSeqPer * tid + e)] = apply<Acc, OP>(head, before_local[e]);
			}
			__syncthreads();
#pragma unroll
			for (int e = 0; e < kSeqPer; ++e) {
				const long long k = start + tid + static_cast<long long>(e) * kSeqThreads;
				if (k < aTiles) {
					p[k] = stage[slot(tid + e * kSeqThreads)];
				}
			}
			__syncthreads();
		}
		carry = apply<Acc, OP>(carry, total);
	}
	if (!scan && threadIdx.x == 0) {
		store_value(m.dst + line_base(m.dst_other_stride, m, line), carry);
	}
}

// Output of one tile: rescan from the tile's carry (init when the line is one tile);
// PER >= aPer elements per thread. The tile passes through LDS between the
// coalesced order (element j to thread j mod 256) and per-thread segments
// (aPer consecutive elements); one padding slot per 16 elements puts the
// segment reads and writes of a wave on different banks.
template<class Acc, int OP, int KIND, int PER>
__global__ __launch_bounds__(kSeqThreads) void row_output_kernel(seq_map m, const Acc* aCarries, long long aTiles,
	int aPer, Acc aInit)
{
	__shared__ Acc stage[kSeqTile + kSeqTile / 16];
	const auto slot = [](int aIndex) { return aIndex + (aIndex >> 4); };
	__shared__ Acc shared[4];
	const long long task = blockIdx.x;
	const long long line = task / aTiles;
	const long long t = task - line * aTiles;
	const long long tile = static_cast<long long>(kSeqThreads) * aPer;
	const long long begin = t * tile;
	const int n = static_cast<int>(min(tile, m.length - begin));
	const long long src_base = line_base(m.src_other_stride, m, line);
	const long long dst_base = line_base(m.dst_other_stride, m, line);
	{
		Acc v[PER];
		load_line<Acc, OP, PER>(m, src_base, begin, n, threadIdx.x, kSeqThreads, aPer, v);
#pragma unroll
		for (int e = 0; e < PER; ++e) {
			const int j = threadIdx.x + e * kSeqThreads;
			if (e < aPer && j < n) {
				stage[slot(j)] = v[e];
			}
		}
	}
	__syncthreads();
	Acc local[PER];
	Acc run = identity<Acc, OP>();
#pragma unroll
	for (int e = 0; e < PER; ++e) {
		const int index = aPer * threadIdx.x + e;
		const Acc x = (e < aPer && index < n) ? stage[slot(index)] : identity<Acc, OP>();
		local[e] = run;             // exclusive within the thread
		run = apply<Acc, OP>(run, x);
		if constexpr (KIND == IANE_SEQ_INCLUSIVE) {
			local[e] = run;
		}
	}
	Acc total;
	const Acc prefix = block_exclusive_scan<Acc, OP>(run, total, shared);
	const Acc carry = aTiles > 1 ? aCarries[task] : aInit;
	const Acc before = apply<Acc, OP>(carry, prefix);
#pragma unroll
	for (int e = 0; e < PER; ++e) {
		const int index = aPer * threadIdx.x + e;
		if (e < aPer && index < n) {
			stage[slot(index)] = apply<Acc, OP>(before, local[e]);
		}
	}
	__syncthreads();
	char* span = m.dst + dst_offset(m, dst_base, begin);
	if (small_span(m.dst_step, n)) {
		const int step = static_cast<int>(m.dst_step);
#pragma unroll
		for (int e = 0; e < PER; ++e) {
			const int j = threadIdx.x + e * kSeqThreads;
			if (e < aPer && j < n) {
				store_value(span + __mul24(j, step), stage[slot(j)]);
			}
		}
	} else {
#pragma unroll
		for (int e = 0; e < PER; ++e) {
			const int j = threadIdx.x + e * kSeqThreads;
			if (e < aPer && j < n) {
				store_value(span + j * m.dst_step, stage[slot(j)]);
			}
		}
	}
}

// row_output_kernel for full-size tiles (kSeqPer elements per thread) of
// contiguous lines whose source and destination rows are whole 16-byte vectors:
// the global loads and stores move 16-byte vectors (thread t takes vectors t,
// t + 256, .. of the tile, each wave access one coalesced run) to and from the
// same LDS stage, and the segments, scans and carries between them are those
// of row_output_kernel, so the results are the same bits.
template<class Acc, int OP, int KIND>
__global__ __launch_bounds__(kSeqThreads) void row_output_vec_kernel(seq_map m, const Acc* aCarries, long long aTiles,
	Acc aInit)
{
	using vec_t = unsigned __attribute__((ext_vector_type(4)));
	__shared__ Acc stage[kSeqTile + kSeqTile / 16];
	const auto slot = [](int aIndex) { return aIndex + (aIndex >> 4); };
	__shared__ Acc shared[4];
	const long long task = blockIdx.x;
	const long long line = task / aTiles;
	const long long t = task - line * aTiles;
	const long long begin = t * kSeqTile;
	const int n = static_cast<int>(min(static_cast<long long>(kSeqTile), m.length - begin));
	const int tid = static_cast<int>(threadIdx.x);
	with_element_type(m.src_type, [&](auto aType) {
		using T = decltype(aType);
		constexpr int SV = 16 / static_cast<int>(sizeof(T));
		constexpr int NS = kSeqPer / SV;
		struct parts {
			T e[SV];
		};
		const char* src = m.src + line_base(m.src_other_stride, m, line) + begin * static_cast<long long>(sizeof(T));
		const int vectors = n / SV;
		vec_t raw[NS];
#pragma unroll
		for (int i = 0; i < NS; ++i) {
			raw[i] = *reinterpret_cast<const vec_t*>(src + 16 * min(tid + kSeqThreads * i, vectors - 1));
		}
#pragma unroll
		for (int i = 0; i < NS; ++i) {
			const int v = tid + kSeqThreads * i;
			if (v < vectors) {
				const parts p = __builtin_bit_cast(parts, raw[i]);
#pragma unroll
				for (int j = 0; j < SV; ++j) {
					stage[slot(v * SV + j)] = static_cast<Acc>(p.e[j]);
				}
			}
		}
	});
	__syncthreads();
	Acc local[kSeqPer];
	Acc run = identity<Acc, OP>();
#pragma unroll
	for (int e = 0; e < kSeqPer; ++e) {
		const int index = kSeqPer * tid + e;
		const Acc x = index < n ? stage[slot(index)] : identity<Acc, OP>();
		local[e] = run;
		run = apply<Acc, OP>(run, x);
		if constexpr (KIND == IANE_SEQ_INCLUSIVE) {
			local[e] = run;
		}
	}
	Acc total;
	const Acc prefix = block_exclusive_scan<Acc, OP>(run, total, shared);
	const Acc carry = aTiles > 1 ? aCarries[task] : aInit;
	const Acc before = apply<Acc, OP>(carry, prefix);
#pragma unroll
	for (int e = 0; e < kSeqPer; ++e) {
		const int index = kSeqPer * tid + e;
		if (index < n) {
			stage[slot(index)] = apply<Acc, OP>(before, local[e]);
		}
	}
	__syncthreads();
	constexpr int DV = 16 / static_cast<int>(sizeof(Acc));
	constexpr int ND = kSeqPer / DV;
	struct out_parts {
		Acc e[DV];
	};
	char* dst = m.dst + line_base(m.dst_other_stride, m, line) + begin * static_cast<long long>(sizeof(Acc));
	const int out_vectors = n / DV;
#pragma unroll
	for (int i = 0; i < ND; ++i) {
		const int v = tid + kSeqThreads * i;
		if (v < out_vectors) {
			out_parts o;
#pragma unroll
			for (int j = 0; j < DV; ++j) {
				o.e[j] = stage[slot(v * DV + j)];
			}
			*reinterpret_cast<vec_t*>(dst + 16LL * v) = __builtin_bit_cast(vec_t, o);
		}
	}
}

// ---- column mode ----------------------------------------------------------
template<class Acc, int OP>
__global__ __launch_bounds__(kSeqThreads) void col_totals_kernel(seq_map m, Acc* aPartials, long long aChunks,
	long long aChunk, int aSkipLast)
{
	const long long index = static_cast<long long>(blockIdx.x) * kSeqThreads + threadIdx.x;
	if (index >= m.lines * aChunks) {
		return;
	}
	const long long line = index % m.lines;
	const long long c = index / m.lines;
	if (aSkipLast && c == aChunks - 1) {
		return;
	}
	const long long begin = c * aChunk;
	const long long end = min(begin + aChunk, m.length);
	const long long base = line_base(m.src_other_stride, m, line);
	Acc acc = load_value<Acc>(m.src + base + begin * m.src_step, m.src_type);
	for (long long i = begin + 1; i < end; ++i) {
		acc = apply<Acc, OP>(acc, load_value<Acc>(m.src + base + i * m.src_step, m.src_type));
	}
	aPartials[index] = acc;
}

template<class Acc, int OP, int KIND>
__global__ __launch_bounds__(kSeqThreads) void col_carries_kernel(seq_map m, Acc* aPartials, long long aChunks, Acc aInit)
{
	const long long line = static_cast<long long>(blockIdx.x) * kSeqThreads + threadIdx.x;
	if (line >= m.lines) {
		return;
	}
	const long long parts = KIND == IANE_SEQ_REDUCE ? aChunks : aChunks - 1;
	Acc total = aInit;
	for (long long c = 0; c < parts; ++c) {
		const Acc x = aPartials[c * m.lines + line];
		if (KIND != IANE_SEQ_REDUCE) {
			aPartials[c * m.lines + line] = total;
		}
		total = apply<Acc, OP>(total, x);
	}
	if (KIND == IANE_SEQ_REDUCE) {
		store_value(m.dst + line_base(m.dst_other_stride, m, line), total);
	} else {
		aPartials[(aChunks - 1) * m.lines + line] = total;
	}
}

// Output of one (chunk, line) from its carry; a reduction of unchunked lines writes the result.
template<class Acc, int OP, int KIND>
__global__ __launch_bounds__(kSeqThreads) void col_output_kernel(seq_map m, const Acc* aCarries, long long aChunks,
	long long aChunk, Acc aInit)
{
	const long long index = static_cast<long long>(blockIdx.x) * kSeqThreads + threadIdx.x;
	if (index >= m.lines * aChunks) {
		return;
	}
	const long long line = index % m.lines;
	const long long c = index / m.lines;
	const long long begin = c * aChunk;
	const long long end = min(begin + aChunk, m.length);
	const long long src_base = line_base(m.src_other_stride, m, line);
	const long long dst_base = line_base(m.dst_other_stride, m, line);
	Acc total = aChunks > 1 ? aCarries[index] : aInit;
	for (long long i = begin; i < end; ++i) {
		const Acc x = load_value<Acc>(m.src + src_base + i * m.src_step, m.src_type);
		if constexpr (KIND == IANE_SEQ_EXCLUSIVE) {
			store_value(m.dst + dst_base + i * m.dst_step, total);
		}
		total = apply<Acc, OP>(total, x);
		if constexpr (KIND == IANE_SEQ_INCLUSIVE) {
			store_value(m.dst + dst_base + i * m.dst_step, total);
		}
	}
	if constexpr (KIND == IANE_SEQ_REDUCE) {
		store_value(m.dst + dst_base, total);
	}
}

// Every destination position <- init (reductions of empty lines).
template<class Acc>
__global__ __launch_bounds__(kSeqThreads) void fill_init_kernel(seq_map m, Acc aInit)
{
	const long long line = static_cast<long long>(blockIdx.x) * kSeqThreads + threadIdx.x;
	if (line < m.lines) {
		store_value(m.dst + line_base(m.dst_other_stride, m, line), aInit);
	}
}

unsigned blocks_of(long long aCount)
{
	const long long blocks = (aCount + kSeqThreads - 1) / kSeqThreads;
	if (blocks > 0x7fffffffLL) {
		fail(IANE_ERR_UNSUPPORTED, "sequence exceeds the grid size limit");
	}
	return static_cast<unsigned>(blocks);
}

unsigned grid_of(long long aBlocks)
{
	if (aBlocks > 0x7fffffffLL) {
		fail(IANE_ERR_UNSUPPORTED, "sequence exceeds the grid size limit");
	}
	return static_cast<unsigned>(aBlocks);
}

// Vectors per lane of row_reduce_vec_kernel for this run's lines (1, 2 or 4),
// or 0 when the lines are not contiguous, 16-byte aligned whole vectors of at
// most 256 vectors.
int vector_lines(const seq_map& m)
{
	const long long size = dtype_size(m.src_type);
	if (m.src_step != size || (m.length * size) % 16 != 0 || reinterpret_cast<uintptr_t>(m.src) % 16 != 0) {
		return 0;
	}
	for (int k = 0; k < m.other_rank; ++k) {
		if (m.src_other_stride[k] % 16 != 0) {
			return 0;
		}
	}
	const long long vectors = m.length * size / 16;
	return vectors <= 64 ? 1 : (vectors <= 128 ? 2 : (vectors <= 256 ? 4 : 0));
}

// Lines of whole 16-byte source vectors (unit step, 16-byte aligned line
// starts): full-size tiles take row_totals_vec_kernel.
bool vector_source_rows(const seq_map& m)
{
	const long long ss = dtype_size(m.src_type);
	if (m.src_step != ss || (m.length * ss) % 16 != 0 || reinterpret_cast<uintptr_t>(m.src) % 16 != 0) {
		return false;
	}
	for (int k = 0; k < m.other_rank; ++k) {
		if (m.src_other_stride[k] % 16 != 0) {
			return false;
		}
	}
	return true;
}

// Scan destinations of whole 16-byte vectors of Acc (unit step, 16-byte
// aligned line starts): full-size tiles take row_output_vec_kernel.
bool vector_dest_rows(const seq_map& m, size_t aAccBytes)
{
	const long long ds = static_cast<long long>(aAccBytes);
	if (m.dst_step != ds || (m.length * ds) % 16 != 0 || reinterpret_cast<uintptr_t>(m.dst) % 16 != 0) {
		return false;
	}
	for (int k = 0; k < m.other_rank; ++k) {
		if (m.dst_other_stride[k] % 16 != 0) {
			return false;
		}
	}
	return true;
}

// Tile totals with as many loads per thread as the tile needs (4 or 16).
template<class Acc, int OP>
void launch_row_totals(const iane_sequence_plan& p, const seq_map& m, Acc* aPartials, long long aTasks, int aSkipLast,
	hipStream_t aStream)
{
	if (p.per == kSeqPer && vector_source_rows(m)) {
		row_totals_vec_kernel<Acc, OP><<<grid_of(aTasks), kSeqThreads, 0, aStream>>>(m, aPartials, p.tiles, aSkipLast);
	} else if (p.per <= 4) {
		row_totals_kernel<Acc, OP, 4><<<grid_of(aTasks), kSeqThreads, 0, aStream>>>(m, aPartials, p.tiles, p.per,
			aSkipLast);
	} else {
		row_totals_kernel<Acc, OP, kSeqPer><<<grid_of(aTasks), kSeqThreads, 0, aStream>>>(m, aPartials, p.tiles, p.per,
			aSkipLast);
	}
}

template<class Acc, int OP, int KIND>
void run_kind(const iane_sequence_plan& p, const seq_map& m, void* aPartials, hipStream_t s)
{
	Acc init;
	std::memcpy(&init, p.init, sizeof(Acc));
	auto* partials = static_cast<Acc*>(aPartials);
	if (m.lines == 0) {
		return;
	}
	if (m.length == 0) {
		if constexpr (KIND == IANE_SEQ_REDUCE) {
			fill_init_kernel<Acc><<<blocks_of(m.lines), kSeqThreads, 0, s>>>(m, init);
		}
		return;
	}
	if (p.row) {
		const long long tasks = m.lines * p.tiles;
		if constexpr (KIND == IANE_SEQ_REDUCE) {
			const int vpl = vector_lines(m);
			if (p.wave_reduce && vpl > 0) {
				const long long lines_per_block = 4LL * (8 / vpl);
				const unsigned grid = grid_of((m.lines + lines_per_block - 1) / lines_per_block);
				if (vpl == 1) {
					row_reduce_vec_kernel<Acc, OP, 1><<<grid, kSeqThreads, 0, s>>>(m, init);
				} else if (vpl == 2) {
					row_reduce_vec_kernel<Acc, OP, 2><<<grid, kSeqThreads, 0, s>>>(m, init);
				} else {
					row_reduce_vec_kernel<Acc, OP, 4><<<grid, kSeqThreads, 0, s>>>(m, init);
				}
				return;
			}
			if (p.wave_reduce) {
				const unsigned grid = grid_of((m.lines + 3) / 4);
				if (m.length <= 64 * 4) {
					row_reduce_wave_kernel<Acc, OP, 4><<<grid, kSeqThreads, 0, s>>>(m, init);
				} else if (m.length <= 64 * 8) {
					row_reduce_wave_kernel<Acc, OP, 8><<<grid, kSeqThreads, 0, s>>>(m, init);
				} else {
					row_reduce_wave_kernel<Acc, OP, kSeqPer><<<grid, kSeqThreads, 0, s>>>(m, init);
				}
				return;
			}
			launch_row_totals<Acc, OP>(p, m, partials, tasks, 0, s);
			row_carries_kernel<Acc, OP, KIND><<<grid_of(m.lines), kSeqThreads, 0, s>>>(m, partials, p.tiles, init);
		} else {
			if (p.tiles > 1) {
				launch_row_totals<Acc, OP>(p, m, partials, tasks, 1, s);
				row_carries_kernel<Acc, OP, KIND><<<grid_of(m.lines), kSeqThreads, 0, s>>>(m, partials, p.tiles, init);
			}
			if (p.per == kSeqPer && vector_source_rows(m) && vector_dest_rows(m, sizeof(Acc))) {
				row_output_vec_kernel<Acc, OP, KIND><<<grid_of(tasks), kSeqThreads, 0, s>>>(m, partials, p.tiles, init);
			} else if (p.per <= 4) {
				row_output_kernel<Acc, OP, KIND, 4><<<grid_of(tasks), kSeqThreads, 0, s>>>(m, partials, p.tiles, p.per, init);
			} else {
				row_output_kernel<Acc, OP, KIND, kSeqPer><<<grid_of(tasks), kSeqThreads, 0, s>>>(m, partials, p.tiles, p.per,
					init);
			}
		}
		return;
	}
	const long long tasks = m.lines * p.chunks;
	if (p.chunks > 1) {
		col_totals_kernel<Acc, OP><<<blocks_of(tasks), kSeqThreads, 0, s>>>(m, partials, p.chunks, p.chunk,
			KIND == IANE_SEQ_REDUCE ? 0 : 1);
		col_carries_kernel<Acc, OP, KIND><<<blocks_of(m.lines), kSeqThreads, 0, s>>>(m, partials, p.chunks, init);
		if constexpr (KIND == IANE_SEQ_REDUCE) {
			return;
		}
	}
	col_output_kernel<Acc, OP, KIND><<<blocks_of(tasks), kSeqThreads, 0, s>>>(m, partials, p.chunks, p.chunk, init);
}

template<class Acc, int OP>
void run_op(const iane_sequence_plan& p, const seq_map& m, void* w, hipStream_t s)
{
	switch (p.kind) {
	case IANE_SEQ_INCLUSIVE: run_kind<Acc, OP, IANE_SEQ_INCLUSIVE>(p, m, w, s); break;
	case IANE_SEQ_EXCLUSIVE: run_kind<Acc, OP, IANE_SEQ_EXCLUSIVE>(p, m, w, s); break;
	default: run_kind<Acc, OP, IANE_SEQ_REDUCE>(p, m, w, s); break;
	}
}

template<class Acc>
void run_acc(const iane_sequence_plan& p, const seq_map& m, void* w, hipStream_t s)
{
	switch (p.op) {
	case IANE_OP_SUM: run_op<Acc, IANE_OP_SUM>(p, m, w, s); break;
	case IANE_OP_PRODUCT: run_op<Acc, IANE_OP_PRODUCT>(p, m, w, s); break;
	case IANE_OP_MIN: run_op<Acc, IANE_OP_MIN>(p, m, w, s); break;
	default: run_op<Acc, IANE_OP_MAX>(p, m, w, s); break;
	}
}

long long pow2_at_least(long long aValue)
{
	long long p = 1;
	while (p < aValue) {
		p <<= 1;
	}
	return p;
}

} // namespace
} // namespace iane_hip

using namespace iane_hip;

extern "C" int iane_hip_sequence_plan_create(iane_sequence_plan** aPlan, int aKind, int aAxis, int aOp, int aInitType,
	const void* aInit, const iane_view_desc* aSrc, const iane_view_desc* aDst)
{
	return guarded([&] {
		if (aPlan == nullptr || aInit == nullptr) {
			invalid("sequence_plan_create: null argument");
		}
		*aPlan = nullptr;
		const auto src = make_view4(aSrc, "source");
		const auto dst = make_view4(aDst, "destination");
		if (aKind < IANE_SEQ_INCLUSIVE || aKind > IANE_SEQ_REDUCE) {
			invalid("unknown sequence kind");
		}
		if (aOp < IANE_OP_SUM || aOp > IANE_OP_MAX) {
			invalid("unknown sequence operator");
		}
		if (aInitType == IANE_C64 || src.type == IANE_C64) {
			fail(IANE_ERR_UNSUPPORTED, "scan and reduce support u8, u16, i32, f32 and f64 elements");
		}
		dtype_size(aInitType);
		if (dst.type != aInitType) {
			invalid("the destination element type must be the type of the initial value");
		}
		if (aAxis < -1 || aAxis >= src.rank) {
			invalid("the sequence axis must be below the source rank");
		}
		// expected destination shape (sequence.hpp:276-290)
		int expected_rank = src.rank;
		int expected[4] = {1, 1, 1, 1};
		if (aKind != IANE_SEQ_REDUCE) {
			for (int i = 0; i < src.rank; ++i) {
				expected[i] = src.extent[i];
			}
		} else if (aAxis < 0 || src.rank == 1) {
			expected_rank = 1;
		} else {
			expected_rank = src.rank - 1;
			for (int i = 0; i < expected_rank; ++i) {
				expected[i] = src.extent[i + (i >= aAxis)];
			}
		}
		if (dst.rank != expected_rank) {
			invalid("the destination rank does not fit the sequence");
		}
		for (int i = 0; i < expected_rank; ++i) {
			if (dst.extent[i] != expected[i]) {
				invalid("Sequence views must match the prepared shape");
			}
		}
		auto plan = new iane_sequence_plan{};
		plan->kind = aKind;
		plan->axis = aAxis;
		plan->op = aOp;
		plan->acc_type = aInitType;
		std::memcpy(plan->init, aInit, static_cast<size_t>(dtype_size(aInitType)));
		plan->src_type = src.type;
		plan->src_rank = src.rank;
		plan->dst_rank = dst.rank;
		for (int i = 0; i < 4; ++i) {
			plan->src_extent[i] = src.extent[i];
			plan->dst_extent[i] = dst.extent[i];
		}
		const long long count = element_count(src);
		if (aAxis < 0) {
			plan->lines = (count == 0 && aKind != IANE_SEQ_REDUCE) ? 0 : 1;
			plan->length = count;
		} else {
			plan->length = src.extent[aAxis];
			long long others = 1;
			for (int i = 0; i < src.rank; ++i) {
				others *= i == aAxis ? 1 : src.extent[i];
			}
			plan->lines = (count != 0 || aKind == IANE_SEQ_REDUCE) ? others : 0;
		}
		if (plan->lines > 0x7fffffffLL || plan->length > (1LL << 40)) {
			delete plan;
			fail(IANE_ERR_UNSUPPORTED, "sequence exceeds the supported line count");
		}
		plan->row = aAxis <= 0;
		const size_t acc = static_cast<size_t>(dtype_size(aInitType));
		if (plan->row) {
			const long long per_thread = (plan->length + kSeqThreads - 1) / kSeqThreads;
			plan->per = static_cast<int>(std::min<long long>(kSeqPer, std::max<long long>(1, pow2_at_least(per_thread))));
			const long long tile = static_cast<long long>(kSeqThreads) * plan->per;
			plan->tiles = plan->length == 0 ? 1 : (plan->length + tile - 1) / tile;
			plan->wave_reduce = aKind == IANE_SEQ_REDUCE && plan->length <= kSeqTile;
			const bool totals = (aKind == IANE_SEQ_REDUCE && !plan->wave_reduce) || plan->tiles > 1;
			plan->partials_bytes = totals ? static_cast<size_t>(plan->lines * plan->tiles) * acc : 0;
		} else {
			// enough (line, chunk) tasks to fill the GPU, chunks of at least 32 elements
			long long chunks = 1;
			if (plan->lines > 0 && plan->lines < 131072 && plan->length > 32) {
				chunks = std::min((131072 + plan->lines - 1) / plan->lines, (plan->length + 31) / 32);
			}
			plan->chunk = std::max<long long>(1, (plan->length + chunks - 1) / chunks);
			plan->chunks = plan->length == 0 ? 1 : (plan->length + plan->chunk - 1) / plan->chunk;
			plan->partials_bytes = plan->chunks > 1 ? static_cast<size_t>(plan->lines * plan->chunks) * acc : 0;
		}
		// whole views that are not packed pass through a packed copy in the workspace
		plan->stage_src = aAxis < 0 && count > 0 && !is_packed(src);
		plan->stage_dst = aAxis < 0 && aKind != IANE_SEQ_REDUCE && count > 0 && !is_packed(dst);
		const auto round = [](size_t aBytes) { return (aBytes + 255) / 256 * 256; };
		plan->workspace = round(plan->partials_bytes) +
			(plan->stage_src ? round(static_cast<size_t>(count) * dtype_size(src.type)) : 0) +
			(plan->stage_dst ? round(static_cast<size_t>(count) * acc) : 0);
		*aPlan = plan;
	});
}

extern "C" int iane_hip_sequence_plan_workspace_bytes(const iane_sequence_plan* aPlan, size_t* aBytes)
{
	return guarded([&] {
		if (aPlan == nullptr || aBytes == nullptr) {
			invalid("sequence_plan_workspace_bytes: null argument");
		}
		*aBytes = aPlan->workspace;
	});
}

extern "C" int iane_hip_sequence_plan_destroy(iane_sequence_plan* aPlan)
{
	delete aPlan;
	return IANE_OK;
}

extern "C" int iane_hip_sequence_run(const iane_sequence_plan* aPlan, const iane_view_desc* aSrc,
	const iane_view_desc* aDst, void* aWorkspace, void* aStream)
{
	return guarded([&] {
		if (aPlan == nullptr) {
			invalid("sequence_run: null plan");
		}
		const auto src = make_view4(aSrc, "source");
		const auto dst = make_view4(aDst, "destination");
		if (src.rank != aPlan->src_rank || dst.rank != aPlan->dst_rank || src.type != aPlan->src_type ||
			dst.type != aPlan->acc_type) {
			invalid("Sequence views must match the prepared shape");
		}
		for (int i = 0; i < 4; ++i) {
			if (src.extent[i] != aPlan->src_extent[i] || dst.extent[i] != aPlan->dst_extent[i]) {
				invalid("Sequence views must match the prepared shape");
			}
		}
		if (aPlan->workspace && aWorkspace == nullptr) {
			invalid("sequence_run: the plan needs a workspace");
		}
		const bool whole = aPlan->axis < 0;
		if (whole && ((!aPlan->stage_src && element_count(src) > 0 && !is_packed(src)) ||
				(!aPlan->stage_dst && aPlan->kind != IANE_SEQ_REDUCE && element_count(dst) > 0 && !is_packed(dst)))) {
			invalid("a whole-view sequence prepared with packed views needs packed views");
		}
		auto stream = static_cast<hipStream_t>(aStream);
		const auto round = [](size_t aBytes) { return (aBytes + 255) / 256 * 256; };
		char* ws = static_cast<char*>(aWorkspace);
		char* staged_src = ws + round(aPlan->partials_bytes);
		char* staged_dst = staged_src + (aPlan->stage_src ? round(static_cast<size_t>(element_count(src)) *
			dtype_size(src.type)) : 0);
		const int src_bytes = dtype_size(src.type);
		const int dst_bytes = dtype_size(dst.type);
		const long long count = element_count(src);
		const auto stage = [&](bool aGather, const view4& aView, char* aPacked, int aBytes) {
			const unsigned blocks = blocks_of(count);
			switch (aBytes) {
			case 1: aGather ? stage_kernel<1, true><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count)
			                : stage_kernel<1, false><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count); break;
			case 2: aGather ? stage_kernel<2, true><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count)
			                : stage_kernel<2, false><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count); break;
			case 4: aGather ? stage_kernel<4, true><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count)
			                : stage_kernel<4, false><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count); break;
			default: aGather ? stage_kernel<8, true><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count)
			                 : stage_kernel<8, false><<<blocks, kSeqThreads, 0, stream>>>(aView, aPacked, count); break;
			}
		};
		seq_map m{};
		m.src = src.ptr;
		m.dst = dst.ptr;
		m.src_type = src.type;
		m.lines = aPlan->lines;
		m.length = aPlan->length;
		if (whole) {
			// one line, unit step over a packed view (staged when the view is not packed)
			if (aPlan->stage_src && !is_packed(src)) {
				stage(true, src, staged_src, src_bytes);
				m.src = staged_src;
			}
			if (aPlan->stage_dst && !is_packed(dst)) {
				m.dst = staged_dst;
			}
			m.src_step = src_bytes;
			m.dst_step = aPlan->kind == IANE_SEQ_REDUCE ? 0 : dst_bytes;
			m.other_rank = 0;
		} else {
			const int axis = aPlan->axis;
			m.src_step = src.stride[axis];
			m.dst_step = aPlan->kind == IANE_SEQ_REDUCE ? 0 : dst.stride[axis];
			int k = 0;
			for (int i = 0; i < src.rank; ++i) {
				if (i == axis) {
					continue;
				}
				m.other_extent[k] = src.extent[i];
				m.src_other_stride[k] = src.stride[i];
				// scans keep the axes; a reduction along `axis` drops it from the destination
				m.dst_other_stride[k] = aPlan->kind == IANE_SEQ_REDUCE ? (src.rank == 1 ? 0 : dst.stride[k])
				                                                        : dst.stride[i];
				++k;
			}
			m.other_rank = k;
			// lines at line x stride: the kernels skip the per-line index division
			const auto nests = [&](const long long* aStride) {
				for (int j = 1; j < m.other_rank; ++j) {
					if (aStride[j] != aStride[j - 1] * m.other_extent[j - 1]) {
						return false;
					}
				}
				return true;
			};
			m.nested = nests(m.src_other_stride) && nests(m.dst_other_stride);
		}
		void* partials = aPlan->partials_bytes ? ws : nullptr;
		switch (aPlan->acc_type) {
		case IANE_U8: run_acc<uint8_t>(*aPlan, m, partials, stream); break;
		case IANE_U16: run_acc<uint16_t>(*aPlan, m, partials, stream); break;
		case IANE_I32: run_acc<int32_t>(*aPlan, m, partials, stream); break;
		case IANE_F32: run_acc<float>(*aPlan, m, partials, stream); break;
		default: run_acc<double>(*aPlan, m, partials, stream); break;
		}
		if (whole && m.dst == staged_dst && aPlan->stage_dst) {
			stage(false, dst, staged_dst, dst_bytes);
		}
		check_hip(hipGetLastError(), "Sequence launch");
	});
}
