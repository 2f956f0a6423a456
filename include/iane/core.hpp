// Header-only host side of the MI355X IANE backend: coordinates, views,
// owning images, execution policies and completion handles.
//
// Mirrors the reference's public types so that callers keep their code:
//   math::vector / dim_traits / detail::{axis, element_count, advance, coordinate_at}
//     (include/iane/math/{vector,coord_traits,index,offset}.hpp)
//   host_image_view, const_host_image_view, device_image_view, const_device_image_view,
//   host::make_view, device::make_view (include/iane/view/*.hpp)
//   host_image, device_image (include/iane/image/image.hpp, resource/*)
//   cpu_policy, cpu_execution (include/iane/execution/context.hpp:20-31)
//   hip_policy / hip_execution in place of cuda_policy / cuda_execution (:46-59)
// Device work goes through the C-ABI of libiane_hip.so (iane_hip.h); this
// header needs only a C++20 compiler (no HIP compiler).
#pragma once

#include <iane_hip.h>

#include <algorithm>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

namespace iane {

// ------------------------------------------------------------- errors -----
namespace detail {

// Map a C-ABI status to the exception type the reference throws.
inline void check_status(int aStatus)
{
	if (aStatus == IANE_OK) {
		return;
	}
	const std::string message = iane_hip_last_error();
	switch (aStatus) {
	case IANE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(message);
	case IANE_ERR_LENGTH: throw std::length_error(message);
	case IANE_ERR_OVERFLOW: throw std::overflow_error(message);
	default: throw std::runtime_error(message);
	}
}

} // namespace detail

// -------------------------------------------------------------- math ------
namespace math {

template<class T, int N>
struct vector {
	static constexpr int dim = N;
	T mElements[N];

	constexpr vector() : mElements{} {}
	template<class... TArgs>
	requires(sizeof...(TArgs) == N && N > 1 && (std::convertible_to<TArgs, T> && ...))
	constexpr vector(TArgs... aValues) : mElements{static_cast<T>(aValues)...} {}
	// fill constructor: vector(v, std::true_type{}) sets every element to v
	constexpr vector(T aValue, std::true_type) : mElements{}
	{
		for (int i = 0; i < N; ++i) {
			mElements[i] = aValue;
		}
	}
	constexpr T& operator[](int aIndex) { return mElements[aIndex]; }
	constexpr const T& operator[](int aIndex) const { return mElements[aIndex]; }
	friend constexpr bool operator==(const vector& a, const vector& b)
	{
		for (int i = 0; i < N; ++i) {
			if (a[i] != b[i]) {
				return false;
			}
		}
		return true;
	}
};

using vec2i = vector<int, 2>;
using vec3i = vector<int, 3>;
using vec4i = vector<int, 4>;

// rank 1 uses scalar coordinates (coord_traits.hpp:45-60)
template<int N>
struct dim_traits {
	using extents_t = vector<int, N>;
	using coord_t = vector<int, N>;
	using diff_t = vector<int, N>;
};
template<>
struct dim_traits<1> {
	using extents_t = int;
	using coord_t = int;
	using diff_t = int;
};

namespace detail {

template<class TCoordinate>
constexpr decltype(auto) axis(TCoordinate& aCoordinate, int aAxis)
{
	if constexpr (std::is_arithmetic_v<std::remove_cvref_t<TCoordinate>>) {
		return (aCoordinate);
	} else {
		return aCoordinate[aAxis];
	}
}

// index.hpp:24-50
template<int N, class TExtent>
std::size_t element_count(const TExtent& aSize)
{
	bool empty = false;
	for (int i = 0; i < N; ++i) {
		const auto extent = axis(aSize, i);
		if (extent < 0) {
			throw std::invalid_argument("Image dimensions must be nonnegative");
		}
		empty |= extent == 0;
	}
	if (empty) {
		return 0;
	}
	std::size_t count = 1;
	for (int i = 0; i < N; ++i) {
		const auto extent = static_cast<std::size_t>(axis(aSize, i));
		if (count > std::numeric_limits<std::size_t>::max() / extent) {
			throw std::overflow_error("Image element count overflows");
		}
		count *= extent;
	}
	return count;
}

// storage order, axis 0 fastest (index.hpp:53-64)
template<int N, class TCoordinate, class TExtent>
constexpr bool advance(TCoordinate& aPosition, const TExtent& aSize)
{
	for (int i = 0; i < N; ++i) {
		auto& coordinate = axis(aPosition, i);
		if (++coordinate < axis(aSize, i)) {
			return true;
		}
		coordinate = 0;
	}
	return false;
}

template<int N, class TCoordinate, class TExtent>
constexpr TCoordinate coordinate_at(std::size_t aIndex, const TExtent& aSize)
{
	TCoordinate position{};
	for (int i = 0; i < N - 1; ++i) {
		const auto extent = static_cast<std::size_t>(axis(aSize, i));
		axis(position, i) = static_cast<int>(aIndex % extent);
		aIndex /= extent;
	}
	axis(position, N - 1) = static_cast<int>(aIndex);
	return position;
}

} // namespace detail

// offset.hpp:11-27
template<class TVector>
constexpr std::ptrdiff_t offset_in_strided_memory(const TVector& aStrides, const TVector& aCoords)
{
	if constexpr (std::is_arithmetic_v<TVector>) {
		return static_cast<std::ptrdiff_t>(aStrides) * aCoords;
	} else {
		std::ptrdiff_t acc = 0;
		for (int i = 0; i < TVector::dim; ++i) {
			acc += static_cast<std::ptrdiff_t>(aStrides[i]) * aCoords[i];
		}
		return acc;
	}
}

} // namespace math

// ------------------------------------------------------------- layout -----
namespace resource::detail {

// packed_layout.hpp:10-34: element count and packed byte strides
template<class T, int N>
auto packed_layout(const typename math::dim_traits<N>::extents_t& aSize)
{
	const auto count = math::detail::element_count<N>(aSize);
	if (count > std::numeric_limits<std::size_t>::max() / sizeof(T)) {
		throw std::overflow_error("Packed image byte size overflows");
	}
	typename math::dim_traits<N>::extents_t strides{};
	std::size_t stride = sizeof(T);
	for (int i = 0; i < N; ++i) {
		if (stride > static_cast<std::size_t>(std::numeric_limits<int>::max())) {
			throw std::overflow_error("Packed image stride exceeds the supported range");
		}
		math::detail::axis(strides, i) = static_cast<int>(stride);
		if (i + 1 < N) {
			const auto extent = static_cast<std::size_t>(math::detail::axis(aSize, i));
			if (extent != 0 && stride > std::numeric_limits<std::size_t>::max() / extent) {
				throw std::overflow_error("Packed image stride overflows");
			}
			stride *= extent;
		}
	}
	return std::pair{count, strides};
}

} // namespace resource::detail

// -------------------------------------------------------------- views -----
namespace detail {

template<class T, int N>
class view_base {
public:
	using value_type = std::remove_const_t<T>;
	using extents_t = typename math::dim_traits<N>::extents_t;
	using coord_t = typename math::dim_traits<N>::coord_t;
	static constexpr int dimension = N;

	view_base() : mPtr(nullptr), mSize{}, mStrides{} {}
	view_base(T* aPtr, extents_t aSize, extents_t aStrides) : mPtr(aPtr), mSize(aSize), mStrides(aStrides) {}
	extents_t dimensions() const { return mSize; }
	extents_t strides() const { return mStrides; }
	T* pointer() const { return mPtr; }

protected:
	// Pointer, extents and strides of the rank N-1 view with axis tI fixed at aIndex.
	template<int tI>
	auto slice_parts(int aIndex) const
	{
		using lower = typename math::dim_traits<N - 1>::extents_t;
		lower size{};
		lower strides{};
		int j = 0;
		for (int i = 0; i < N; ++i) {
			if (i == tI) {
				continue;
			}
			math::detail::axis(size, j) = math::detail::axis(mSize, i);
			math::detail::axis(strides, j) = math::detail::axis(mStrides, i);
			++j;
		}
		using byte_t = std::conditional_t<std::is_const_v<T>, const char, char>;
		T* first = reinterpret_cast<T*>(reinterpret_cast<byte_t*>(mPtr) +
			static_cast<std::ptrdiff_t>(math::detail::axis(mStrides, tI)) * aIndex);
		return std::tuple{first, size, strides};
	}

	T* at(const coord_t& aCoord) const
	{
		const auto offset = math::offset_in_strided_memory(mStrides, aCoord);
		using byte_t = std::conditional_t<std::is_const_v<T>, const char, char>;
		return reinterpret_cast<T*>(reinterpret_cast<byte_t*>(mPtr) + offset);
	}

	T* mPtr;
	extents_t mSize;
	extents_t mStrides;
};

} // namespace detail

/// Host view: host_image_view.hpp:11-81
template<class T, int N>
class host_image_view : public detail::view_base<T, N> {
	using base = detail::view_base<T, N>;
public:
	struct is_host_view_tag {};
	using typename base::coord_t;
	using typename base::extents_t;
	using base::base;
	T& operator[](const coord_t& aCoord) const { return *this->at(aCoord); }
	host_image_view subview(const coord_t& aCorner, const extents_t& aSize) const
	{
		return {this->at(aCorner), aSize, this->mStrides};
	}
	/// The rank N-1 view with axis tI fixed at aIndex, sharing the storage (host_image_view.hpp:59-75).
	template<int tI>
	requires(N > 1 && tI >= 0 && tI < N)
	host_image_view<T, N - 1> slice(int aIndex) const
	{
		const auto [first, size, strides] = this->template slice_parts<tI>(aIndex);
		return {first, size, strides};
	}
};

template<class T, int N>
class const_host_image_view : public detail::view_base<const T, N> {
	using base = detail::view_base<const T, N>;
public:
	struct is_host_view_tag {};
	using typename base::coord_t;
	using typename base::extents_t;
	using base::base;
	const_host_image_view(host_image_view<T, N> aView)
		: base(aView.pointer(), aView.dimensions(), aView.strides()) {}
	const T& operator[](const coord_t& aCoord) const { return *this->at(aCoord); }
	const_host_image_view subview(const coord_t& aCorner, const extents_t& aSize) const
	{
		return {this->at(aCorner), aSize, this->mStrides};
	}
	/// The rank N-1 view with axis tI fixed at aIndex, sharing the storage (host_image_view.hpp:59-75).
	template<int tI>
	requires(N > 1 && tI >= 0 && tI < N)
	const_host_image_view<T, N - 1> slice(int aIndex) const
	{
		const auto [first, size, strides] = this->template slice_parts<tI>(aIndex);
		return {first, size, strides};
	}
};

/// Device view: device_image_view.hpp:11-60 (no host element access)
template<class T, int N>
class device_image_view : public detail::view_base<T, N> {
	using base = detail::view_base<T, N>;
public:
	struct is_device_view_tag {};
	using typename base::coord_t;
	using typename base::extents_t;
	using base::base;
	device_image_view subview(const coord_t& aCorner, const extents_t& aSize) const
	{
		return {this->at(aCorner), aSize, this->mStrides};
	}
	/// The rank N-1 view with axis tI fixed at aIndex, sharing the storage (host_image_view.hpp:59-75).
	template<int tI>
	requires(N > 1 && tI >= 0 && tI < N)
	device_image_view<T, N - 1> slice(int aIndex) const
	{
		const auto [first, size, strides] = this->template slice_parts<tI>(aIndex);
		return {first, size, strides};
	}
};

template<class T, int N>
class const_device_image_view : public detail::view_base<const T, N> {
	using base = detail::view_base<const T, N>;
public:
	struct is_device_view_tag {};
	using typename base::coord_t;
	using typename base::extents_t;
	using base::base;
	const_device_image_view(device_image_view<T, N> aView)
		: base(aView.pointer(), aView.dimensions(), aView.strides()) {}
	const_device_image_view subview(const coord_t& aCorner, const extents_t& aSize) const
	{
		return {this->at(aCorner), aSize, this->mStrides};
	}
	/// The rank N-1 view with axis tI fixed at aIndex, sharing the storage (host_image_view.hpp:59-75).
	template<int tI>
	requires(N > 1 && tI >= 0 && tI < N)
	const_device_image_view<T, N - 1> slice(int aIndex) const
	{
		const auto [first, size, strides] = this->template slice_parts<tI>(aIndex);
		return {first, size, strides};
	}
};

namespace concepts {

template<class V>
concept HostView = requires { typename V::is_host_view_tag; };
template<class V>
concept DeviceView = requires { typename V::is_device_view_tag; };
template<class V>
concept AnyView = HostView<V> || DeviceView<V>;
template<class V>
concept MemoryView = AnyView<V> && requires(const V& v) { v.pointer(); v.strides(); };

} // namespace concepts

namespace host {
template<class T, int N>
host_image_view<T, N> make_view(T* aPtr, typename math::dim_traits<N>::extents_t aSize)
{
	return {aPtr, aSize, resource::detail::packed_layout<T, N>(aSize).second};
}
template<class T, int N>
const_host_image_view<T, N> make_const_view(const T* aPtr, typename math::dim_traits<N>::extents_t aSize)
{
	return {aPtr, aSize, resource::detail::packed_layout<T, N>(aSize).second};
}
} // namespace host

namespace device {
template<class T, int N>
device_image_view<T, N> make_view(T* aPtr, typename math::dim_traits<N>::extents_t aSize)
{
	return {aPtr, aSize, resource::detail::packed_layout<T, N>(aSize).second};
}
template<class T, int N>
const_device_image_view<T, N> make_const_view(const T* aPtr, typename math::dim_traits<N>::extents_t aSize)
{
	return {aPtr, aSize, resource::detail::packed_layout<T, N>(aSize).second};
}
} // namespace device

// ------------------------------------------------------------- images -----
namespace detail {

struct host_storage {
	static void* allocate(std::size_t aBytes) { return aBytes ? ::operator new(aBytes) : nullptr; }
	static void release(void* aPtr) noexcept { ::operator delete(aPtr); }
};

struct device_storage {
	static void* allocate(std::size_t aBytes)
	{
		void* ptr = nullptr;
		check_status(iane_hip_malloc(&ptr, aBytes));
		return ptr;
	}
	static void release(void* aPtr) noexcept { iane_hip_free(aPtr); }
};

// Owning packed image, move-only (image.hpp:44-151).
template<class T, int N, class Storage, template<class, int> class View, template<class, int> class ConstView>
class basic_image {
public:
	using value_type = T;
	using extents_t = typename math::dim_traits<N>::extents_t;
	using view_t = View<T, N>;
	using const_view_t = ConstView<T, N>;
	static constexpr int dimension = N;
	static constexpr int dim = N;

	basic_image() = default;
	explicit basic_image(extents_t aSize) { resize(aSize); }
	template<class... TArgs>
	requires(sizeof...(TArgs) == N && N > 1)
	explicit basic_image(TArgs... aSize) : basic_image(extents_t(aSize...)) {}
	basic_image(const basic_image&) = delete;
	basic_image& operator=(const basic_image&) = delete;
	basic_image(basic_image&& aOther) noexcept { swap(aOther); }
	basic_image& operator=(basic_image&& aOther) noexcept
	{
		swap(aOther);
		return *this;
	}
	~basic_image() { Storage::release(mPtr); }

	/// Release the storage and become empty (image.hpp:139-145).
	void reset() noexcept
	{
		Storage::release(mPtr);
		mPtr = nullptr;
		mSize = extents_t{};
		mStrides = extents_t{};
	}

	void resize(extents_t aSize)
	{
		const auto [count, strides] = resource::detail::packed_layout<T, N>(aSize);
		void* fresh = Storage::allocate(count * sizeof(T));
		Storage::release(mPtr);
		mPtr = static_cast<T*>(fresh);
		mSize = aSize;
		mStrides = strides;
	}
	T* pointer() const { return mPtr; }
	extents_t dimensions() const { return mSize; }
	extents_t strides() const { return mStrides; }
	view_t view() const { return view_t(mPtr, mSize, mStrides); }
	const_view_t const_view() const { return const_view_t(mPtr, mSize, mStrides); }

private:
	void swap(basic_image& aOther) noexcept
	{
		std::swap(mPtr, aOther.mPtr);
		std::swap(mSize, aOther.mSize);
		std::swap(mStrides, aOther.mStrides);
	}
	T* mPtr = nullptr;
	extents_t mSize{};
	extents_t mStrides{};
};

} // namespace detail

template<class T, int N>
using host_image = detail::basic_image<T, N, detail::host_storage, host_image_view, const_host_image_view>;
template<class T, int N>
using device_image = detail::basic_image<T, N, detail::device_storage, device_image_view, const_device_image_view>;

// ------------------------------------------------------- execution --------
/// context.hpp:20-23: host traversal with at most `threads` workers.
struct cpu_policy {
	std::size_t threads = 1;
	std::size_t grain_size = 4096;
};

/// context.hpp:26-31
struct cpu_execution {
	void wait() const noexcept {}
	operator cpu_policy() const noexcept { return {}; }
};

/// Queue HIP work on a borrowed stream (the cuda_policy of context.hpp:46-49).
struct hip_policy {
	void* stream = nullptr;
};

/// Completion handle: wait() synchronises the stream (context.hpp:51-59).
struct hip_execution {
	void* stream = nullptr;
	operator hip_policy() const noexcept { return {stream}; }
	void wait() const { detail::check_status(iane_hip_stream_sync(stream)); }
};

template<concepts::AnyView V>
using view_policy = std::conditional_t<concepts::DeviceView<V>, hip_policy, cpu_policy>;
template<concepts::AnyView V>
using view_execution = std::conditional_t<concepts::DeviceView<V>, hip_execution, cpu_execution>;

// ------------------------------------------------------- descriptors ------
namespace detail {

template<class T>
constexpr int element_code()
{
	using U = std::remove_const_t<T>;
	if constexpr (std::same_as<U, std::uint8_t>) {
		return IANE_U8;
	} else if constexpr (std::same_as<U, std::uint16_t>) {
		return IANE_U16;
	} else if constexpr (std::same_as<U, std::int32_t>) {
		return IANE_I32;
	} else if constexpr (std::same_as<U, float>) {
		return IANE_F32;
	} else if constexpr (std::same_as<U, double>) {
		return IANE_F64;
	} else {
		return -1;
	}
}

template<class T>
inline constexpr bool abi_element = element_code<T>() >= 0;

// Describe a memory view for the C-ABI (pointer, extents, byte strides).
template<concepts::MemoryView V>
iane_view_desc describe(const V& aView, int aCode = element_code<typename V::value_type>())
{
	iane_view_desc d{};
	d.ptr = const_cast<void*>(static_cast<const void*>(aView.pointer()));
	d.rank = V::dimension;
	d.dtype = aCode;
	const auto size = aView.dimensions();
	const auto strides = aView.strides();
	for (int i = 0; i < V::dimension; ++i) {
		d.extent[i] = math::detail::axis(size, i);
		d.stride_bytes[i] = math::detail::axis(strides, i);
	}
	return d;
}

} // namespace detail

// ------------------------------------------------------- texture views ----
/// Texel types of texture images (include/iane/view/texture_image_view.hpp:13-18).
template<class T>
concept texture_element = std::same_as<T, float> || std::same_as<T, signed char> || std::same_as<T, unsigned char> ||
	std::same_as<T, short> || std::same_as<T, unsigned short> || std::same_as<T, int> || std::same_as<T, unsigned int>;

/**
 * Read-only texture view (texture_image_view.hpp:20-98). gfx950 offers no
 * image/texture sampling to HIP, so the texels live in device memory and the
 * view reads them as a const device view with the reference's lookup rules:
 * an integer lookup returns the texel without interpolation or normalisation,
 * coordinates are local to the view, and a subview keeps the storage and moves
 * its origin (a crop is an independent image for the border rule). Like the
 * reference it is a readable device view, but neither a memory view nor a
 * writable one.
 */
template<texture_element T, int N = 2>
requires(N >= 1 && N <= 3)
class texture_image_view {
public:
	using value_type = T;
	using coord_t = typename math::dim_traits<N>::coord_t;
	using extents_t = typename math::dim_traits<N>::extents_t;
	struct is_device_view_tag {};
	static constexpr int dimension = N;

	explicit texture_image_view(const_device_image_view<T, N> aTexels) : mTexels(aTexels) {}
	/// Logical extent in texels.
	extents_t dimensions() const { return mTexels.dimensions(); }
	/// In-bounds region whose coordinates start at zero (texture_image_view.hpp:74-80).
	texture_image_view subview(const coord_t& aCorner, const extents_t& aSize) const
	{
		return texture_image_view(mTexels.subview(aCorner, aSize));
	}
	/// The texels of this view as device memory (how this backend stores a texture).
	const_device_image_view<T, N> texels() const { return mTexels; }

private:
	const_device_image_view<T, N> mTexels;
};

namespace detail {

// A texture view goes to the C-ABI as the device memory view of its texels.
template<class T, int N>
iane_view_desc describe(const texture_image_view<T, N>& aView, int aCode = element_code<T>())
{
	return describe(aView.texels(), aCode);
}

} // namespace detail

} // namespace iane
