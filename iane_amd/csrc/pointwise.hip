// Pointwise expressions behind iane_hip_pointwise, and convolution of a
// constant source behind iane_hip_convolve_constant.
//
// The reference evaluates lazy views (constant_view / map_view / zip_view and
// the view algebra, include/iane/view/procedural.hpp:93-150, view/algebra.hpp:52-151)
// by running the caller's functor per pixel inside copy / transform
// (include/iane/algorithms/copy.hpp:28-47, transform.hpp:72-92). A caller of
// this library has no device compiler, so its headers describe the per-pixel
// function as a short register program over the input views (iane_pw_program)
// and this kernel interprets it: one output element per thread iteration, the
// program uniform over the grid (its opcode switch is a scalar branch), the
// registers in LDS (indexed [register][thread], so a wave's accesses fall in
// distinct banks). Every value is an FP32 number; inputs are converted on load
// and the result on store, so +, -, *, /, sqrt, min, max, the comparisons and
// select give the same bits as the same float operations on the host.
#include "common.hpp"

#include <algorithm>

namespace iane_hip {
namespace {

constexpr int kPwThreads = 256;

struct pw_input {
	const char* ptr;
	int type;
	int packed;          // element i sits at ptr + i * size
	int size;
	int rank;
	int extent[3];       // read by point samples (any extent)
	long long stride[4];
};

struct pw_args {
	int n_inputs;
	int n_ops;
	int result;
	pw_input inputs[IANE_PW_MAX_INPUTS];
	iane_pw_op ops[IANE_PW_MAX_OPS];
	view4 dst;
	int dst_packed;
	long long count;
};

__device__ __forceinline__ float pw_load(const pw_input& aIn, long long aIndex, const int (&aCoord)[4])
{
	const char* at;
	if (aIn.packed) {
		at = aIn.ptr + aIndex * aIn.size;
	} else {
		long long offset = 0;
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			offset += static_cast<long long>(aCoord[k]) * aIn.stride[k];
		}
		at = aIn.ptr + offset;
	}
	return load_as_float(at, aIn.type);
}

// Point sample at view-local coordinates (s0, s1, s2) of a rank <= 3 input:
// clamp to the edge texel centres, read texel floor(coordinate).
__device__ __forceinline__ float pw_sample(const pw_input& aIn, float aS0, float aS1, float aS2)
{
	const float s[3] = {aS0, aS1, aS2};
	long long offset = 0;
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		if (k < aIn.rank) {
			const int extent = aIn.extent[k];
			const float c = s[k];
			if (extent <= 0 || !isfinite(c)) {
				return 0.0f;
			}
			const float last = static_cast<float>(extent) - 0.5f;
			const float clamped = c < 0.5f ? 0.5f : (c > last ? last : c);
			offset += static_cast<long long>(floorf(clamped)) * aIn.stride[k];
		}
	}
	return load_as_float(aIn.ptr + offset, aIn.type);
}

// C++ assignment of a float to the destination element; out-of-range and NaN
// values saturate (0 for NaN) instead of being undefined.
__device__ __forceinline__ void pw_store(char* aPtr, int aType, float aValue)
{
	switch (aType) {
	case IANE_U8: *reinterpret_cast<uint8_t*>(aPtr) = saturate_truncate<uint8_t>(aValue); break;
	case IANE_U16: *reinterpret_cast<uint16_t*>(aPtr) = saturate_truncate<uint16_t>(aValue); break;
	case IANE_I32: *reinterpret_cast<int32_t*>(aPtr) = saturate_truncate<int32_t>(aValue); break;
	case IANE_F32: *reinterpret_cast<float*>(aPtr) = aValue; break;
	default: *reinterpret_cast<double*>(aPtr) = static_cast<double>(aValue); break;
	}
}

__global__ __launch_bounds__(kPwThreads) void pointwise_kernel(pw_args a)
{
	__shared__ float regs[IANE_PW_MAX_OPS][kPwThreads];
	const int t = threadIdx.x;
	for (long long i = blockIdx.x * static_cast<long long>(kPwThreads) + t; i < a.count;
		 i += static_cast<long long>(gridDim.x) * kPwThreads) {
		int coord[4];
		long long rest = i;
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			coord[k] = static_cast<int>(rest % a.dst.extent[k]);
			rest /= a.dst.extent[k];
		}
		coord[3] = static_cast<int>(rest);
		for (int o = 0; o < a.n_ops; ++o) {
			const iane_pw_op op = a.ops[o];
			const float x = regs[op.a][t];
			const float y = regs[op.b][t];
			float r = 0.0f;
			switch (op.code) {
			case IANE_PW_INPUT: r = pw_load(a.inputs[op.a], i, coord); break;
			case IANE_PW_CONST: r = op.value; break;
			case IANE_PW_ADD: r = x + y; break;
			case IANE_PW_SUB: r = x - y; break;
			case IANE_PW_MUL: r = x * y; break;
			case IANE_PW_DIV: r = x / y; break;
			case IANE_PW_MIN: r = y < x ? y : x; break;
			case IANE_PW_MAX: r = x < y ? y : x; break;
			case IANE_PW_NEG: r = -x; break;
			case IANE_PW_ABS: r = fabsf(x); break;
			case IANE_PW_SQRT: r = sqrtf(x); break;
			case IANE_PW_EXP: r = expf(x); break;
			case IANE_PW_LOG: r = logf(x); break;
			case IANE_PW_SIN: r = sinf(x); break;
			case IANE_PW_COS: r = cosf(x); break;
			case IANE_PW_LOG1P: r = log1pf(x); break;
			case IANE_PW_LOG2: r = log2f(x); break;
			case IANE_PW_LOG10: r = log10f(x); break;
			case IANE_PW_EXP2: r = exp2f(x); break;
			case IANE_PW_POW: r = powf(x, y); break;
			case IANE_PW_LT: r = x < y ? 1.0f : 0.0f; break;
			case IANE_PW_LE: r = x <= y ? 1.0f : 0.0f; break;
			case IANE_PW_GT: r = x > y ? 1.0f : 0.0f; break;
			case IANE_PW_GE: r = x >= y ? 1.0f : 0.0f; break;
			case IANE_PW_EQ: r = x == y ? 1.0f : 0.0f; break;
			case IANE_PW_NE: r = x != y ? 1.0f : 0.0f; break;
			case IANE_PW_SELECT: r = x != 0.0f ? y : regs[op.c][t]; break;
			case IANE_PW_FMA: r = fmaf(x, y, regs[op.c][t]); break;
			case IANE_PW_COORD:
				r = static_cast<float>(op.a == 0 ? coord[0] : op.a == 1 ? coord[1] : op.a == 2 ? coord[2] : coord[3]);
				break;
			default: // IANE_PW_SAMPLE: input a at (r[b], r[c], r[value])
				r = pw_sample(a.inputs[op.a], y, regs[op.c][t], regs[static_cast<int>(op.value)][t]);
				break;
			}
			regs[op.dst][t] = r;
		}
		const float value = regs[a.result][t];
		char* at;
		if (a.dst_packed) {
			at = a.dst.ptr + i * dtype_size_device(a.dst.type);
		} else {
			at = a.dst.ptr;
#pragma unroll
			for (int k = 0; k < 4; ++k) {
				at += static_cast<long long>(coord[k]) * a.dst.stride[k];
			}
		}
		pw_store(at, a.dst.type, value);
	}
}

// Constant source: the sample at every in-range (or clamped) position is
// `value`, a zero-border sample outside the view is 0. Taps in kernel storage
// order with the fma chain of the direct kernels (FP64 in REFERENCE precision,
// FP32 otherwise) and the reference's output conversion.
template<class Acc>
__global__ __launch_bounds__(256) void convolve_constant_kernel(Acc value, view4 ker, view4 dst, int border, long long count)
{
	for (long long i = blockIdx.x * 256LL + threadIdx.x; i < count; i += static_cast<long long>(gridDim.x) * 256) {
		int p[4];
		long long rest = i;
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			p[k] = static_cast<int>(rest % dst.extent[k]);
			rest /= dst.extent[k];
		}
		Acc sum = Acc(0);
		for (int q3 = 0; q3 < ker.extent[3]; ++q3) {
			for (int q2 = 0; q2 < ker.extent[2]; ++q2) {
				for (int q1 = 0; q1 < ker.extent[1]; ++q1) {
					for (int q0 = 0; q0 < ker.extent[0]; ++q0) {
						const int q[4] = {q0, q1, q2, q3};
						bool zero = false;
#pragma unroll
						for (int k = 0; k < 4; ++k) {
							const int c = p[k] + ker.extent[k] / 2 - q[k];
							zero = zero || ((c < 0 || c >= dst.extent[k]) && border == IANE_BORDER_ZERO);
						}
						const char* kp = ker.ptr + q0 * ker.stride[0] + q1 * ker.stride[1] + q2 * ker.stride[2] +
							q3 * ker.stride[3];
						sum = fma(zero ? Acc(0) : value, load_as<Acc>(kp, ker.type), sum);
					}
				}
			}
		}
		char* at = dst.ptr;
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			at += static_cast<long long>(p[k]) * dst.stride[k];
		}
		store_converted(at, dst.type, sum);
	}
}

unsigned grid_for(long long aCount, int aThreads)
{
	return static_cast<unsigned>(std::min<long long>(std::max<long long>(1, (aCount + aThreads - 1) / aThreads), 4096));
}

} // namespace
} // namespace iane_hip

using namespace iane_hip;

extern "C" int iane_hip_pointwise(const iane_pw_program* aProgram, const iane_view_desc* aDst, void* aStream)
{
	return guarded([&] {
		if (!aProgram) {
			invalid("pointwise program is null");
		}
		const auto dst = make_view4(aDst, "destination");
		if (dst.type == IANE_C64) {
			fail(IANE_ERR_UNSUPPORTED, "pointwise expressions write u8, u16, i32, f32 or f64 elements");
		}
		const iane_pw_program& p = *aProgram;
		if (p.n_inputs < 0 || p.n_inputs > IANE_PW_MAX_INPUTS || p.n_ops < 1 || p.n_ops > IANE_PW_MAX_OPS) {
			invalid("pointwise program exceeds its input or operation limit");
		}
		pw_args a{};
		a.n_inputs = p.n_inputs;
		a.n_ops = p.n_ops;
		a.dst = dst;
		a.dst_packed = is_packed(dst) ? 1 : 0;
		a.count = element_count(dst);
		// inputs read at the output position must match its extents; inputs
		// that are only point-sampled may have any extent (rank 1..3)
		bool loaded[IANE_PW_MAX_INPUTS] = {};
		bool sampled[IANE_PW_MAX_INPUTS] = {};
		for (int o = 0; o < p.n_ops; ++o) {
			const iane_pw_op& op = p.ops[o];
			if ((op.code == IANE_PW_INPUT || op.code == IANE_PW_SAMPLE) && op.a >= 0 && op.a < p.n_inputs) {
				(op.code == IANE_PW_INPUT ? loaded : sampled)[op.a] = true;
			}
		}
		for (int i = 0; i < p.n_inputs; ++i) {
			const auto in = make_view4(&p.inputs[i], "pointwise input");
			if (in.type == IANE_C64) {
				fail(IANE_ERR_UNSUPPORTED, "pointwise expressions read u8, u16, i32, f32 or f64 elements");
			}
			if (loaded[i] || !sampled[i]) {
				if (in.rank != dst.rank) {
					invalid("Zip operands require matching nonnegative dimensions");
				}
				for (int k = 0; k < 4; ++k) {
					if (in.extent[k] != dst.extent[k]) {
						invalid("Source and destination require matching nonnegative dimensions");
					}
				}
			}
			if (sampled[i] && in.rank > 3) {
				fail(IANE_ERR_UNSUPPORTED, "point samples read views of rank 1..3");
			}
			a.inputs[i].ptr = in.ptr;
			a.inputs[i].type = in.type;
			a.inputs[i].size = dtype_size(in.type);
			a.inputs[i].packed = is_packed(in) && a.dst_packed && loaded[i] ? 1 : 0;
			a.inputs[i].rank = in.rank;
			for (int k = 0; k < 3; ++k) {
				a.inputs[i].extent[k] = in.extent[k];
			}
			for (int k = 0; k < 4; ++k) {
				a.inputs[i].stride[k] = in.stride[k];
			}
		}
		// every operand must be a register an earlier operation wrote
		bool defined[IANE_PW_MAX_OPS] = {};
		const auto reg = [&](int r) {
			if (r < 0 || r >= IANE_PW_MAX_OPS || !defined[r]) {
				invalid("pointwise program reads a register no earlier operation wrote");
			}
		};
		for (int o = 0; o < p.n_ops; ++o) {
			const iane_pw_op& op = p.ops[o];
			if (op.code < IANE_PW_INPUT || op.code > IANE_PW_SAMPLE) {
				invalid("unknown pointwise operation");
			}
			if (op.code == IANE_PW_INPUT) {
				if (op.a < 0 || op.a >= p.n_inputs) {
					invalid("pointwise program reads a missing input");
				}
			} else if (op.code == IANE_PW_COORD) {
				if (op.a < 0 || op.a >= 4) {
					invalid("pointwise coordinate axis out of range");
				}
			} else if (op.code == IANE_PW_SAMPLE) {
				if (op.a < 0 || op.a >= p.n_inputs) {
					invalid("pointwise program reads a missing input");
				}
				const int rank = p.inputs[op.a].rank;
				reg(op.b);
				if (rank >= 2) {
					reg(op.c);
				}
				if (rank >= 3) {
					if (op.value != static_cast<float>(static_cast<int>(op.value))) {
						invalid("pointwise program reads a register no earlier operation wrote");
					}
					reg(static_cast<int>(op.value));
				}
			} else if (op.code != IANE_PW_CONST) {
				reg(op.a);
				const bool binary = op.code < IANE_PW_NEG || op.code >= IANE_PW_POW;  // SELECT and FMA read b too
				if (binary) {
					reg(op.b);
				}
				if (op.code == IANE_PW_SELECT || op.code == IANE_PW_FMA) {
					reg(op.c);
				}
			}
			if (op.dst < 0 || op.dst >= IANE_PW_MAX_OPS) {
				invalid("pointwise program writes past its registers");
			}
			a.ops[o] = op;
			// unused operand fields still index the register file: keep them in range
			a.ops[o].a = (op.code == IANE_PW_INPUT || op.code == IANE_PW_CONST) ? op.dst : op.a;
			const bool binary = op.code >= IANE_PW_ADD && op.code <= IANE_PW_FMA &&
				(op.code < IANE_PW_NEG || op.code >= IANE_PW_POW);
			a.ops[o].b = binary ? op.b : a.ops[o].a;
			a.ops[o].c = (op.code == IANE_PW_SELECT || op.code == IANE_PW_FMA) ? op.c : a.ops[o].a;
			if (op.code == IANE_PW_INPUT || op.code == IANE_PW_COORD) {
				a.ops[o].a = op.a;  // the input index or axis, not a register
				a.ops[o].b = op.dst;
				a.ops[o].c = op.dst;
			}
			if (op.code == IANE_PW_SAMPLE) {
				const int rank = p.inputs[op.a].rank;
				a.ops[o].a = op.a;  // the input index; coordinates in b, c, value
				a.ops[o].b = op.b;
				a.ops[o].c = rank >= 2 ? op.c : op.b;
				a.ops[o].value = static_cast<float>(rank >= 3 ? static_cast<int>(op.value) : op.b);
			}
			defined[op.dst] = true;
		}
		reg(p.result);
		a.result = p.result;
		if (a.count == 0) {
			return;
		}
		pointwise_kernel<<<grid_for(a.count, kPwThreads), kPwThreads, 0, static_cast<hipStream_t>(aStream)>>>(a);
		check_hip(hipGetLastError(), "pointwise launch");
	});
}

extern "C" int iane_hip_convolve_constant(double aValue, const iane_view_desc* aKernel, const iane_view_desc* aDst,
	int aBorder, int aPrecision, void* aStream)
{
	return guarded([&] {
		const auto ker = make_view4(aKernel, "kernel");
		const auto dst = make_view4(aDst, "destination");
		validate_convolution(dst, ker, dst);
		if (aBorder != IANE_BORDER_CLAMP && aBorder != IANE_BORDER_ZERO) {
			invalid("unknown border mode");
		}
		if (aPrecision < IANE_PREC_REFERENCE || aPrecision > IANE_PREC_SPLIT_FP16) {
			invalid("unknown precision mode");
		}
		const long long count = element_count(dst);
		if (count == 0) {
			return;
		}
		auto stream = static_cast<hipStream_t>(aStream);
		if (aPrecision == IANE_PREC_REFERENCE) {
			convolve_constant_kernel<double><<<grid_for(count, 256), 256, 0, stream>>>(aValue, ker, dst, aBorder, count);
		} else {
			convolve_constant_kernel<float><<<grid_for(count, 256), 256, 0, stream>>>(static_cast<float>(aValue), ker, dst,
				aBorder, count);
		}
		check_hip(hipGetLastError(), "constant convolution launch");
	});
}
