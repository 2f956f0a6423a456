// Shared HIP helpers of libiane_hip.so: error state, element access by dtype,
// view validation and the XCD-aware block remap used by every tiled kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "../../include/iane_hip.h"

namespace iane_hip {

// ---------------------------------------------------------------------------
// Error reporting: C-ABI functions catch these and turn them into status codes
// plus a thread-local message (iane_hip_last_error).
struct status_error : std::runtime_error {
	int code;
	status_error(int aCode, const std::string& aMessage) : std::runtime_error(aMessage), code(aCode) {}
};

[[noreturn]] inline void fail(int aCode, const std::string& aMessage)
{
	throw status_error(aCode, aMessage);
}

inline void invalid(const std::string& aMessage) { fail(IANE_ERR_INVALID_ARGUMENT, aMessage); }

inline void check_hip(hipError_t aStatus, const char* aOperation)
{
	if (aStatus != hipSuccess) {
		fail(IANE_ERR_HIP, std::string(aOperation) + ": " + hipGetErrorString(aStatus));
	}
}

void set_last_error(const std::string& aMessage);

// Launch-shape knobs (tile sizes, thread counts, chunking) set by the
// autotuner through iane_hip_tune_set; `aDefault` when unset.
int tune_get(const char* aKey, int aDefault);

// Run a C-ABI body, mapping exceptions to status codes.
template<class F>
int guarded(F&& aBody)
{
	try {
		aBody();
		return IANE_OK;
	} catch (const status_error& e) {
		set_last_error(e.what());
		return e.code;
	} catch (const std::bad_alloc&) {
		set_last_error("host allocation failed");
		return IANE_ERR_HIP;
	} catch (const std::exception& e) {
		set_last_error(e.what());
		return IANE_ERR_HIP;
	}
}

// ---------------------------------------------------------------------------
// Element types.
inline int dtype_size(int aType)
{
	switch (aType) {
	case IANE_U8: return 1;
	case IANE_U16: return 2;
	case IANE_I32: return 4;
	case IANE_F32: return 4;
	case IANE_F64: return 8;
	case IANE_C64: return 8;
	default: invalid("unknown element type " + std::to_string(aType)); return 0;
	}
}

inline bool dtype_integral(int aType) { return aType == IANE_U8 || aType == IANE_U16 || aType == IANE_I32; }

// Element bytes in device code (the host-side dtype_size also validates).
__device__ __forceinline__ int dtype_size_device(int aType)
{
	return aType == IANE_U8 ? 1 : aType == IANE_U16 ? 2 : (aType == IANE_F64 || aType == IANE_C64) ? 8 : 4;
}

// Read one element of runtime type `aType` as double (exact for every type).
__device__ __forceinline__ double load_as_double(const char* aPtr, int aType)
{
	switch (aType) {
	case IANE_U8: return static_cast<double>(*reinterpret_cast<const uint8_t*>(aPtr));
	case IANE_U16: return static_cast<double>(*reinterpret_cast<const uint16_t*>(aPtr));
	case IANE_I32: return static_cast<double>(*reinterpret_cast<const int32_t*>(aPtr));
	case IANE_F32: return static_cast<double>(*reinterpret_cast<const float*>(aPtr));
	default: return *reinterpret_cast<const double*>(aPtr);
	}
}

__device__ __forceinline__ float load_as_float(const char* aPtr, int aType)
{
	switch (aType) {
	case IANE_U8: return static_cast<float>(*reinterpret_cast<const uint8_t*>(aPtr));
	case IANE_U16: return static_cast<float>(*reinterpret_cast<const uint16_t*>(aPtr));
	case IANE_I32: return static_cast<float>(*reinterpret_cast<const int32_t*>(aPtr));
	case IANE_F32: return *reinterpret_cast<const float*>(aPtr);
	default: return static_cast<float>(*reinterpret_cast<const double*>(aPtr));
	}
}

// Call f(T{}) with the element type T of runtime code `aType`. Used to hoist
// the type dispatch out of load loops: a switch around each load puts every
// load in its own basic block and the compiler then waits for each one at the
// join, serialising the memory traffic.
template<class F>
__device__ __forceinline__ void with_element_type(int aType, F&& aFunction)
{
	switch (aType) {
	case IANE_U8: aFunction(uint8_t{}); break;
	case IANE_U16: aFunction(uint16_t{}); break;
	case IANE_I32: aFunction(int32_t{}); break;
	case IANE_F32: aFunction(float{}); break;
	default: aFunction(double{}); break;
	}
}

template<class Acc>
__device__ __forceinline__ Acc load_as(const char* aPtr, int aType)
{
	if constexpr (sizeof(Acc) == 8) {
		return load_as_double(aPtr, aType);
	} else {
		return load_as_float(aPtr, aType);
	}
}

// Output conversion of the reference (include/iane/algorithms/convolution.hpp:51-75):
// u8 rounds half up after saturating at 0 (also NaN) and 255; other integral
// types map NaN to 0, saturate, then truncate toward zero; floats use static_cast.
template<class TInt>
__device__ __forceinline__ TInt saturate_truncate(double aSum)
{
	constexpr double lower = static_cast<double>(std::numeric_limits<TInt>::lowest());
	constexpr double upper = static_cast<double>(std::numeric_limits<TInt>::max());
	if (aSum != aSum) {
		return TInt{};
	}
	if (aSum <= lower) {
		return std::numeric_limits<TInt>::lowest();
	}
	if (aSum >= upper) {
		return std::numeric_limits<TInt>::max();
	}
	return static_cast<TInt>(aSum);
}

__device__ __forceinline__ uint8_t round_byte(double aSum)
{
	if (!(aSum > 0.0)) {
		return 0;
	}
	if (aSum >= 255.0) {
		return 255;
	}
	const auto integer = static_cast<unsigned int>(aSum);
	return static_cast<uint8_t>(integer + (aSum - integer >= 0.5));
}

__device__ __forceinline__ void store_converted(char* aPtr, int aType, double aSum)
{
	switch (aType) {
	case IANE_U8: *reinterpret_cast<uint8_t*>(aPtr) = round_byte(aSum); break;
	case IANE_U16: *reinterpret_cast<uint16_t*>(aPtr) = saturate_truncate<uint16_t>(aSum); break;
	case IANE_I32: *reinterpret_cast<int32_t*>(aPtr) = saturate_truncate<int32_t>(aSum); break;
	case IANE_F32: *reinterpret_cast<float*>(aPtr) = static_cast<float>(aSum); break;
	default: *reinterpret_cast<double*>(aPtr) = aSum; break;
	}
}

// FP32 accumulators convert through double so integral destinations get the
// same saturation / rounding rules (the FP32 sum itself is already rounded).
__device__ __forceinline__ void store_converted(char* aPtr, int aType, float aSum)
{
	if (aType == IANE_F32) {
		*reinterpret_cast<float*>(aPtr) = aSum;
	} else {
		store_converted(aPtr, aType, static_cast<double>(aSum));
	}
}

// ---------------------------------------------------------------------------
// Views padded to rank 4 (missing axes get extent 1, stride 0) with 64-bit strides.
struct view4 {
	char* ptr;
	int type;
	int rank;
	int extent[4];
	long long stride[4];
};

inline view4 make_view4(const iane_view_desc* aDesc, const char* aName)
{
	if (aDesc == nullptr) {
		invalid(std::string(aName) + " descriptor is null");
	}
	if (aDesc->rank < 1 || aDesc->rank > IANE_MAX_RANK) {
		fail(IANE_ERR_UNSUPPORTED, std::string(aName) + " rank must be 1.." + std::to_string(IANE_MAX_RANK));
	}
	dtype_size(aDesc->dtype);
	view4 v{};
	v.ptr = static_cast<char*>(aDesc->ptr);
	v.type = aDesc->dtype;
	v.rank = aDesc->rank;
	for (int i = 0; i < 4; ++i) {
		if (i < aDesc->rank) {
			if (aDesc->extent[i] < 0) {
				invalid("Image dimensions must be nonnegative");
			}
			v.extent[i] = aDesc->extent[i];
			v.stride[i] = aDesc->stride_bytes[i];
		} else {
			v.extent[i] = 1;
			v.stride[i] = 0;
		}
	}
	return v;
}

inline long long element_count(const view4& aView)
{
	long long count = 1;
	for (int i = 0; i < 4; ++i) {
		count *= aView.extent[i];
	}
	return count;
}

inline bool is_packed(const view4& aView)
{
	long long stride = dtype_size(aView.type);
	for (int i = 0; i < aView.rank; ++i) {
		if (aView.extent[i] > 1 && aView.stride[i] != stride) {
			return false;
		}
		stride *= aView.extent[i];
	}
	return true;
}

// ---------------------------------------------------------------------------
// Launch-order remap: workgroups are dispatched round-robin over the 8 XCDs of
// an MI355X, each with its own L2. Renumbering so that XCD k receives one
// contiguous range of tiles keeps neighbouring tiles (which share halo rows)
// in the same L2. The map is a bijection for any block count.
__device__ __forceinline__ unsigned xcd_remap(unsigned aBlock, unsigned aCount)
{
	constexpr unsigned xcds = 8;
	const unsigned q = aCount / xcds;
	const unsigned r = aCount % xcds;
	const unsigned xcd = aBlock % xcds;
	const unsigned index = aBlock / xcds;
	return xcd * q + (xcd < r ? xcd : r) + index;
}

// Raw buffer resource of one plane (GFX9 word 3: 32-bit data format). Loads
// and stores through it take a scalar base and a VGPR byte offset, and
// offsets at or past `aBytes` read 0 / are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_resource(const char* aBase, int aBytes)
{
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(aBase), static_cast<short>(0), aBytes, 0x00020000);
}

inline unsigned ceil_div(long long a, long long b) { return static_cast<unsigned>((a + b - 1) / b); }

// A caller-supplied workspace: it must hold `aNeeded` bytes (the size the
// matching *_workspace_bytes query reported) and start on a 256-byte boundary.
inline void check_workspace(const void* aWorkspace, size_t aBytes, size_t aNeeded)
{
	if (aNeeded == 0) {
		return;
	}
	if (!aWorkspace || aBytes < aNeeded) {
		fail(IANE_ERR_LENGTH, "Prepared operation exceeded its scratch reservation");
	}
	if (reinterpret_cast<uintptr_t>(aWorkspace) % 256 != 0) {
		invalid("workspace must be 256-byte aligned");
	}
}

// Split-precision MFMA convolution (conv_mfma.hip). The kernel part (scale and
// Toeplitz table, split_kernel_bytes) can be built once per kernel; the source
// part (scale and padded part planes, split_source_bytes) is scratch per call.
// Sizes depend on the precision (two 16-bit parts per operand for FP16, three
// for BF16).
bool split_applicable(const view4& aSrc, const view4& aKer, const view4& aDst);
size_t split_kernel_bytes(const view4& aKer, int aPrecision);
size_t split_source_bytes(const view4& aSrc, const view4& aKer, int aPrecision);
void split_prepare_kernel(const view4& aKer, int aPrecision, void* aKernelPart, hipStream_t aStream);
void split_convolve_prepared(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, int aPrecision,
	const void* aKernelPart, void* aSourcePart, hipStream_t aStream);

// Direct convolution (conv_direct.hip): tiled, streaming or untiled kernels.
void convolve_direct(const view4& aSrc, const view4& aKer, const view4& aDst, int aBorder, int aPrecision,
	hipStream_t aStream, bool aForceSimple);
void validate_convolution(const view4& aSrc, const view4& aKer, const view4& aDst);

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations
// (lgkmcnt) and then for the other waves, but not for outstanding global loads
// (vmcnt), so register prefetches of the next tile stay in flight across it.
// __syncthreads() carries a workgroup fence that drains every counter.
__device__ __forceinline__ void lds_barrier()
{
	asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

} // namespace iane_hip
