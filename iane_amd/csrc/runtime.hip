// Runtime entry points of the C-ABI: allocation, streams, events, transfers.
// They replace the CUDA calls the reference makes from its resource and
// execution layers (include/iane/resource/cuda/resource_cuda.hpp:26-41,
// include/iane/execution/context.hpp:36-59, include/iane/algorithms/copy.hpp:55-108).
#include "common.hpp"

#include <map>
#include <mutex>

namespace iane_hip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& aMessage) { g_last_error = aMessage; }

namespace {
std::mutex g_tune_mutex;
std::map<std::string, int>& tune_table()
{
	static std::map<std::string, int> table;
	return table;
}
} // namespace

int tune_get(const char* aKey, int aDefault)
{
	std::lock_guard lock(g_tune_mutex);
	const auto it = tune_table().find(aKey);
	return it == tune_table().end() ? aDefault : it->second;
}

namespace {

// Strided element-wise copy between two device views of equal extent and
// element size (rank padded to 4). One thread per element, axis 0 fastest so
// packed axis-0 runs are coalesced.
__global__ void copy_strided_kernel(view4 aSrc, view4 aDst, long long aCount, int aBytes)
{
	const long long index = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (index >= aCount) {
		return;
	}
	long long rest = index;
	long long src = 0;
	long long dst = 0;
	for (int i = 0; i < 4; ++i) {
		const long long c = rest % aSrc.extent[i];
		rest /= aSrc.extent[i];
		src += c * aSrc.stride[i];
		dst += c * aDst.stride[i];
	}
	const char* s = aSrc.ptr + src;
	char* d = aDst.ptr + dst;
	switch (aBytes) {
	case 1: *reinterpret_cast<uint8_t*>(d) = *reinterpret_cast<const uint8_t*>(s); break;
	case 2: *reinterpret_cast<uint16_t*>(d) = *reinterpret_cast<const uint16_t*>(s); break;
	case 4: *reinterpret_cast<uint32_t*>(d) = *reinterpret_cast<const uint32_t*>(s); break;
	default: *reinterpret_cast<uint64_t*>(d) = *reinterpret_cast<const uint64_t*>(s); break;
	}
}

} // namespace
} // namespace iane_hip

using namespace iane_hip;

extern "C" {

const char* iane_hip_last_error(void) { return g_last_error.c_str(); }

int iane_hip_version(void) { return 1; }

int iane_hip_tune_set(const char* aKey, int aValue)
{
	return guarded([&] {
		if (!aKey) {
			invalid("tuning key is null");
		}
		std::lock_guard lock(g_tune_mutex);
		if (aValue < 0) {
			tune_table().erase(aKey);
		} else {
			tune_table()[aKey] = aValue;
		}
	});
}

int iane_hip_tune_get(const char* aKey, int* aValue)
{
	return guarded([&] {
		if (!aKey || !aValue) {
			invalid("tuning key or output is null");
		}
		*aValue = tune_get(aKey, -1);
	});
}

int iane_hip_device_count(int* aCount)
{
	return guarded([&] {
		if (!aCount) {
			invalid("count is null");
		}
		int n = 0;
		const auto status = hipGetDeviceCount(&n);
		*aCount = status == hipSuccess ? n : 0;
	});
}

int iane_hip_set_device(int aDevice)
{
	return guarded([&] { check_hip(hipSetDevice(aDevice), "Select HIP device"); });
}

int iane_hip_malloc(void** aPtr, size_t aBytes)
{
	return guarded([&] {
		if (!aPtr) {
			invalid("pointer output is null");
		}
		*aPtr = nullptr;
		if (aBytes) {
			check_hip(hipMalloc(aPtr, aBytes), "Allocate HIP memory");
		}
	});
}

int iane_hip_free(void* aPtr)
{
	return guarded([&] {
		if (aPtr) {
			check_hip(hipFree(aPtr), "Free HIP memory");
		}
	});
}

int iane_hip_host_malloc(void** aPtr, size_t aBytes)
{
	return guarded([&] {
		if (!aPtr) {
			invalid("pointer output is null");
		}
		*aPtr = nullptr;
		if (aBytes) {
			check_hip(hipHostMalloc(aPtr, aBytes, hipHostMallocDefault), "Allocate pinned host memory");
		}
	});
}

int iane_hip_host_free(void* aPtr)
{
	return guarded([&] {
		if (aPtr) {
			check_hip(hipHostFree(aPtr), "Free pinned host memory");
		}
	});
}

int iane_hip_stream_create(void** aStream)
{
	return guarded([&] {
		if (!aStream) {
			invalid("stream output is null");
		}
		hipStream_t stream{};
		check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "Create HIP stream");
		*aStream = stream;
	});
}

int iane_hip_stream_destroy(void* aStream)
{
	return guarded([&] {
		if (aStream) {
			check_hip(hipStreamDestroy(static_cast<hipStream_t>(aStream)), "Destroy HIP stream");
		}
	});
}

int iane_hip_stream_sync(void* aStream)
{
	return guarded([&] {
		check_hip(hipStreamSynchronize(static_cast<hipStream_t>(aStream)), "HIP stream wait");
	});
}

int iane_hip_device_sync(void)
{
	return guarded([&] { check_hip(hipDeviceSynchronize(), "HIP device wait"); });
}

int iane_hip_memcpy_async(void* aDst, const void* aSrc, size_t aBytes, int aKind, void* aStream)
{
	return guarded([&] {
		if (aKind < 0 || aKind > 3) {
			invalid("unknown copy kind");
		}
		if (aBytes == 0) {
			return;
		}
		check_hip(hipMemcpyAsync(aDst, aSrc, aBytes, static_cast<hipMemcpyKind>(aKind),
			static_cast<hipStream_t>(aStream)), "Copy memory");
	});
}

int iane_hip_memset_async(void* aDst, int aValue, size_t aBytes, void* aStream)
{
	return guarded([&] {
		if (aBytes) {
			check_hip(hipMemsetAsync(aDst, aValue, aBytes, static_cast<hipStream_t>(aStream)), "Fill memory");
		}
	});
}

// Transfer semantics of copy.hpp:55-108: equal extents and element types; a
// single copy when both sides are packed, per-plane 2-D copies when rows are
// contiguous, and (device->device only) a strided kernel otherwise.
int iane_hip_copy_view(const iane_view_desc* aSrc, const iane_view_desc* aDst, int aKind, void* aStream)
{
	return guarded([&] {
		const auto src = make_view4(aSrc, "source");
		const auto dst = make_view4(aDst, "destination");
		if (src.rank != dst.rank) {
			invalid("copy requires views of equal rank");
		}
		for (int i = 0; i < 4; ++i) {
			if (src.extent[i] != dst.extent[i]) {
				invalid("Image dimensions must match");
			}
		}
		if (dtype_size(src.type) != dtype_size(dst.type) || src.type != dst.type) {
			invalid("copy requires equal element types");
		}
		if (aKind < 0 || aKind > 3) {
			invalid("unknown copy kind");
		}
		const auto count = element_count(src);
		if (count == 0) {
			return;
		}
		const int bytes = dtype_size(src.type);
		auto stream = static_cast<hipStream_t>(aStream);
		if (is_packed(src) && is_packed(dst)) {
			check_hip(hipMemcpyAsync(dst.ptr, src.ptr, static_cast<size_t>(count) * bytes,
				static_cast<hipMemcpyKind>(aKind), stream), "Copy image");
			return;
		}
		const bool rows = src.stride[0] == bytes && dst.stride[0] == bytes;
		if (rows) {
			// rows of extent[0] elements, extent[1] rows per plane, planes over axes 2,3
			const size_t width = static_cast<size_t>(src.extent[0]) * bytes;
			for (int w = 0; w < src.extent[3]; ++w) {
				for (int z = 0; z < src.extent[2]; ++z) {
					const char* s = src.ptr + z * src.stride[2] + w * src.stride[3];
					char* d = dst.ptr + z * dst.stride[2] + w * dst.stride[3];
					const size_t spitch = src.extent[1] > 1 ? static_cast<size_t>(src.stride[1]) : width;
					const size_t dpitch = dst.extent[1] > 1 ? static_cast<size_t>(dst.stride[1]) : width;
					if (src.stride[1] < 0 || dst.stride[1] < 0) {
						invalid("copy requires nonnegative row strides");
					}
					check_hip(hipMemcpy2DAsync(d, dpitch, s, spitch, width, src.extent[1],
						static_cast<hipMemcpyKind>(aKind), stream), "Copy image planes");
				}
			}
			return;
		}
		if (aKind != hipMemcpyDeviceToDevice) {
			invalid("Transfers require contiguous rows");
		}
		copy_strided_kernel<<<ceil_div(count, 256), 256, 0, stream>>>(src, dst, count, bytes);
		check_hip(hipGetLastError(), "Strided copy launch");
	});
}

int iane_hip_event_create(void** aEvent)
{
	return guarded([&] {
		if (!aEvent) {
			invalid("event output is null");
		}
		hipEvent_t event{};
		check_hip(hipEventCreate(&event), "Create HIP event");
		*aEvent = event;
	});
}

int iane_hip_event_destroy(void* aEvent)
{
	return guarded([&] {
		if (aEvent) {
			check_hip(hipEventDestroy(static_cast<hipEvent_t>(aEvent)), "Destroy HIP event");
		}
	});
}

int iane_hip_event_record(void* aEvent, void* aStream)
{
	return guarded([&] {
		check_hip(hipEventRecord(static_cast<hipEvent_t>(aEvent), static_cast<hipStream_t>(aStream)),
			"Record HIP event");
	});
}

int iane_hip_event_elapsed_ms(float* aMs, void* aStart, void* aStop)
{
	return guarded([&] {
		check_hip(hipEventSynchronize(static_cast<hipEvent_t>(aStop)), "Wait for HIP event");
		check_hip(hipEventElapsedTime(aMs, static_cast<hipEvent_t>(aStart), static_cast<hipEvent_t>(aStop)),
			"HIP event time");
	});
}

} // extern "C"
