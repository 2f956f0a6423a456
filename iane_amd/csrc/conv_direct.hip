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
// Tiled kernel: one workgroup (256 threads) computes a 64 x 32 output tile of
// one (z, w) plane. For every kernel plane (qz, qw) it stages the input plane
// region with its x/y halo (border applied at staging) into LDS as the
// accumulator type, then each thread accumulates a 4 x 2 register block with a
// sliding register window along x (kernel width is a template parameter).
#include "common.hpp"

namespace iane_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kThreadsX = 16;
constexpr int kRx = 4;                 // outputs per thread along x
constexpr int kRy = 2;                 // outputs per thread along y
constexpr int kTileX = kThreadsX * kRx; // 64
constexpr int kTileY = (kThreads / kThreadsX) * kRy; // 32
constexpr size_t kMaxLds = 64 * 1024;

struct conv_args {
	view4 src;
	view4 ker;
	view4 dst;
	int border;
	int center[4];
	unsigned tiles_x;
	unsigned tiles_y;
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

template<class Acc, int KX>
__global__ __launch_bounds__(kThreads) void conv_tiled_kernel(conv_args a)
{
	extern __shared__ __align__(16) char smem[];
	const int kx = KX > 0 ? KX : a.ker.extent[0];
	const int ky = a.ker.extent[1];
	const int kz = a.ker.extent[2];
	const int kw = a.ker.extent[3];
	const int cols = kTileX + kx - 1;
	const int rows = kTileY + ky - 1;
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
	const int y0 = static_cast<int>(tile_y) * kTileY;
	const int tid = threadIdx.x;
	const int tx = tid % kThreadsX;
	const int ty = tid / kThreadsX;

	Acc acc[kRy][kRx];
#pragma unroll
	for (int ry = 0; ry < kRy; ++ry) {
#pragma unroll
		for (int rx = 0; rx < kRx; ++rx) {
			acc[ry][rx] = Acc(0);
		}
	}

	for (int qw = 0; qw < kw; ++qw) {
		for (int qz = 0; qz < kz; ++qz) {
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
			const char* base = a.src.ptr + zi * a.src.stride[2] + wi * a.src.stride[3];
			for (int i = tid; i < rows * cols; i += kThreads) {
				const int ly = i / cols;
				const int lx = i - ly * cols;
				int xi = x0 - a.center[0] + lx;
				int yi = y0 - a.center[1] + ly;
				Acc value = Acc(0);
				if (plane_in && border_coord(xi, a.src.extent[0], a.border) &&
					border_coord(yi, a.src.extent[1], a.border)) {
					value = load_as<Acc>(base + xi * a.src.stride[0] + yi * a.src.stride[1], a.src.type);
				}
				tile[i] = value;
			}
			__syncthreads();
			for (int qy = 0; qy < ky; ++qy) {
				const Acc* krow = coef + qy * kx;
				if constexpr (KX > 0) {
					constexpr int span = kRx + KX - 1;
#pragma unroll
					for (int ry = 0; ry < kRy; ++ry) {
						const Acc* row = tile + (ty * kRy + ry + (ky - 1 - qy)) * cols + tx * kRx;
						Acc window[span];
#pragma unroll
						for (int j = 0; j < span; ++j) {
							window[j] = row[j];
						}
#pragma unroll
						for (int qx = 0; qx < KX; ++qx) {
							const Acc k = krow[qx];
#pragma unroll
							for (int rx = 0; rx < kRx; ++rx) {
								acc[ry][rx] = fma(window[rx + KX - 1 - qx], k, acc[ry][rx]);
							}
						}
					}
				} else {
					for (int qx = 0; qx < kx; ++qx) {
						const Acc k = krow[qx];
#pragma unroll
						for (int ry = 0; ry < kRy; ++ry) {
							const Acc* row = tile + (ty * kRy + ry + (ky - 1 - qy)) * cols + tx * kRx + (kx - 1 - qx);
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
	for (int ry = 0; ry < kRy; ++ry) {
		const int py = y0 + ty * kRy + ry;
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
		for (int qz = 0; qz < a.ker.extent[2]; ++qz) {
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

template<class Acc, int KX>
void launch_tiled(const conv_args& a, size_t aLds, unsigned aBlocks, hipStream_t aStream)
{
	conv_tiled_kernel<Acc, KX><<<aBlocks, kThreads, aLds, aStream>>>(a);
}

template<class Acc>
void launch_conv(conv_args a, hipStream_t aStream, bool aForceSimple)
{
	const int kx = a.ker.extent[0];
	const int ky = a.ker.extent[1];
	const size_t lds = sizeof(Acc) * (static_cast<size_t>(kTileY + ky - 1) * (kTileX + kx - 1) +
		static_cast<size_t>(kx) * ky);
	const long long count = element_count(a.dst);
	if (aForceSimple || lds > kMaxLds) {
		conv_simple_kernel<Acc><<<ceil_div(count, kThreads), kThreads, 0, aStream>>>(a, count);
		return;
	}
	a.tiles_x = ceil_div(a.dst.extent[0], kTileX);
	a.tiles_y = ceil_div(a.dst.extent[1], kTileY);
	const long long blocks = static_cast<long long>(a.tiles_x) * a.tiles_y * a.dst.extent[2] * a.dst.extent[3];
	if (blocks > 0x7fffffffLL) {
		fail(IANE_ERR_UNSUPPORTED, "convolution exceeds the grid size limit");
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	switch (kx) {
	case 1: launch_tiled<Acc, 1>(a, lds, nb, aStream); break;
	case 3: launch_tiled<Acc, 3>(a, lds, nb, aStream); break;
	case 5: launch_tiled<Acc, 5>(a, lds, nb, aStream); break;
	case 7: launch_tiled<Acc, 7>(a, lds, nb, aStream); break;
	case 9: launch_tiled<Acc, 9>(a, lds, nb, aStream); break;
	case 11: launch_tiled<Acc, 11>(a, lds, nb, aStream); break;
	case 13: launch_tiled<Acc, 13>(a, lds, nb, aStream); break;
	case 15: launch_tiled<Acc, 15>(a, lds, nb, aStream); break;
	default: launch_tiled<Acc, 0>(a, lds, nb, aStream); break;
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

// Implemented in conv_mfma.hip: split-precision MFMA path for large float kernels.
int convolve_split(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, int aPrecision,
	hipStream_t aStream);

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

extern "C" int iane_hip_convolve(const iane_view_desc* aSrc, const iane_view_desc* aKernel,
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
		if (aPrecision < IANE_PREC_REFERENCE || aPrecision > IANE_PREC_SPLIT_FP16) {
			invalid("unknown precision mode");
		}
		if (element_count(dst) == 0) {
			return;
		}
		auto stream = static_cast<hipStream_t>(aStream);
		if (aPrecision == IANE_PREC_SPLIT_BF16 || aPrecision == IANE_PREC_SPLIT_FP16) {
			if (convolve_split(src, ker, dst, aBorder, aPrecision, stream)) {
				return;
			}
			// shapes the MFMA path does not tile fall back to FP32 accumulation
			aPrecision = IANE_PREC_FP32;
		}
		convolve_direct(src, ker, dst, aBorder, aPrecision, stream, false);
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
