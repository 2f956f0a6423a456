// Texture images: texture_image and the uploads and copies of the reference
// (include/iane/resource/cuda/texture_image.hpp:13-176, docs/cuda_texture_images.md).
//
// A texture image owns a packed device allocation of its texels (gfx950 has no
// image/texture sampling in HIP; reading texels as memory is also what the
// LDS-tiled convolution kernels want). The owner hands out texture_image_view
// (core.hpp); convolve accepts such a view as its source and applies the
// border rule at the view's logical boundary, and copy materialises a view
// into device memory. Uploads take host or device views of the same rank,
// extents and element type with contiguous rows; rows and 3-D slices may be
// padded, and a slice spacing that is not a whole number of rows is copied
// slice by slice.
#pragma once

#include <iane/run_on.hpp>

namespace iane {

/**
 * Move-only owner of a 1-, 2- or 3-D texture image (texture_image.hpp:13-98).
 * Keep it alive until all uploads and reads finish. An empty extent allocates
 * nothing. @throws std::invalid_argument for negative extents.
 */
template<texture_element T, int N = 2>
requires(N >= 1 && N <= 3)
class texture_image {
public:
	using extents_t = typename math::dim_traits<N>::extents_t;
	using const_view_t = texture_image_view<T, N>;
	static constexpr int dimension = N;

	explicit texture_image(extents_t aSize) : mTexels(aSize) {}

	/// Logical extent in texels.
	extents_t dimensions() const { return mTexels.dimensions(); }
	/// Borrow a read-only view until the allocation is released.
	const_view_t const_view() const { return const_view_t(mTexels.const_view()); }
	/// Device storage of the texels, the interoperability handle of this backend
	/// (the reference returns its CUDA array); nullptr for an empty image.
	const T* array() const { return mTexels.pointer(); }

	/// Upload `aSource` (copy(source, texture, policy), texture_image.hpp:100-176).
	template<concepts::MemoryView TSource>
	requires(TSource::dimension == N && std::same_as<typename TSource::value_type, T>)
	hip_execution upload(TSource aSource, hip_policy aPolicy)
	{
		using math::detail::axis;
		const auto size = dimensions();
		const auto source_size = aSource.dimensions();
		for (int i = 0; i < N; ++i) {
			if (axis(source_size, i) != axis(size, i)) {
				throw std::invalid_argument("Texture upload requires matching dimensions");
			}
		}
		if (math::detail::element_count<N>(size) == 0) {
			return {aPolicy.stream};
		}
		const auto strides = aSource.strides();
		const auto row_bytes = static_cast<std::size_t>(axis(size, 0)) * sizeof(T);
		if (axis(strides, 0) != static_cast<int>(sizeof(T))) {
			throw std::invalid_argument("Texture upload requires contiguous rows");
		}
		if constexpr (N >= 2) {
			if (strides[1] < 0 || static_cast<std::size_t>(strides[1]) < row_bytes) {
				throw std::invalid_argument("Texture upload requires a valid row pitch");
			}
		}
		if constexpr (N == 3) {
			const auto pitch = static_cast<std::size_t>(strides[1]);
			const auto height = static_cast<std::size_t>(size[1]);
			// the occupied rows of a slice, checked without multiplying pitch by height
			if (strides[2] < 0 || static_cast<std::size_t>(strides[2]) < row_bytes ||
				(static_cast<std::size_t>(strides[2]) - row_bytes) / pitch < height - 1) {
				throw std::invalid_argument("Texture upload requires a valid slice stride");
			}
		}
		return copy(aSource, mTexels.view(), aPolicy);
	}

private:
	device_image<T, N> mTexels;
};

/// copy(source view, texture, policy): texture_image.hpp:100-176.
template<concepts::MemoryView TSource, texture_element T, int N>
requires(TSource::dimension == N && std::same_as<typename TSource::value_type, T>)
hip_execution copy(TSource aSource, texture_image<T, N>& aDestination, hip_policy aPolicy = {})
{
	return aDestination.upload(aSource, aPolicy);
}

/// copy(texture view, memory view, policy): materialise the texels of a view.
template<texture_element T, int N, concepts::MemoryView TDst>
requires(TDst::dimension == N && std::same_as<typename TDst::value_type, T>)
hip_execution copy(texture_image_view<T, N> aSource, TDst aDestination, hip_policy aPolicy = {})
{
	return copy(aSource.texels(), aDestination, aPolicy);
}

} // namespace iane
