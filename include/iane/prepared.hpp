// Prepared algorithms with one shared scratch allocation and the FFT drop-ins.
//
// execution_context / scratch_space / make_execution_context mirror
// include/iane/execution/prepared.hpp:18-180 of the reference (prepare reserves
// the op's scratch, the context grows one allocation, operations are bound to
// their context, growth and destruction wait for submitted device work).
// fft_forward / fft_inverse / rfft / irfft / rfft_extent / fft_complex mirror
// include/iane/algorithms/fft/{fft.hpp:196-231, rfft.hpp:107-161, layout.hpp:15-21,
// complex.hpp:6-9} for device views of rank 1..3, on the Stockham FFT of
// libiane_hip.so. (The reference's host FFT is FFTW; host views are not
// provided here.)
#pragma once

#include <iane/core.hpp>

#include <cstring>

namespace iane {

/// complex.hpp:6-9
struct alignas(2 * sizeof(float)) fft_complex {
	float real;
	float imag;
};

/// layout.hpp:15-21
inline math::vec2i rfft_extent(math::vec2i aRealExtent)
{
	if (aRealExtent[0] <= 0 || aRealExtent[1] <= 0) {
		throw std::invalid_argument("Real FFT requires positive dimensions");
	}
	return math::vec2i{aRealExtent[0] / 2 + 1, aRealExtent[1]};
}

namespace detail {

// Carve regions out of the context's current allocation (prepared.hpp:18-57).
template<class TPolicy>
class scratch_space {
public:
	scratch_space(void* aData, std::size_t aSize) : mData(static_cast<char*>(aData)), mSize(aSize) {}
	void* bytes(std::size_t aSize, std::size_t aAlignment = alignof(std::max_align_t))
	{
		if (aSize == 0) {
			return nullptr;
		}
		if (aAlignment == 0 || aAlignment > alignof(std::max_align_t) || (aAlignment & (aAlignment - 1)) != 0) {
			throw std::invalid_argument("Unsupported scratch alignment");
		}
		const auto address = reinterpret_cast<std::uintptr_t>(mData);
		const auto pad = (aAlignment - address % aAlignment) % aAlignment;
		if (mData == nullptr || pad + aSize > mSize) {
			throw std::length_error("Prepared operation exceeded its scratch reservation");
		}
		char* result = mData + pad;
		mData = result + aSize;
		mSize -= pad + aSize;
		return result;
	}

private:
	char* mData;
	std::size_t mSize;
};

template<class TPolicy>
struct scratch_deleter {
	void operator()(void* aPointer) const noexcept
	{
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			iane_hip_free(aPointer);
		} else {
			::operator delete(aPointer);
		}
	}
};

} // namespace detail

/**
 * Prepare and run algorithms with one shared scratch allocation
 * (prepared.hpp:84-180). Calls on a context must be serialised by the
 * caller; a HIP context borrows a stream that must outlive it.
 */
template<class TPolicy>
class execution_context {
	template<class TState>
	class prepared_operation {
		friend class execution_context;
		prepared_operation(const execution_context* aOwner, TState&& aState)
			: mOwner(aOwner), mState(std::move(aState)) {}
		const execution_context* mOwner;
		TState mState;
	};

public:
	explicit execution_context(TPolicy aPolicy) : mPolicy(aPolicy) {}
	execution_context(const execution_context&) = delete;
	execution_context& operator=(const execution_context&) = delete;
	~execution_context()
	{
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			if (mSubmitted) {
				iane_hip_stream_sync(mPolicy.stream);
			}
		}
	}

	template<class TAlgorithm, class TSrc, class TDst>
	auto prepare(TAlgorithm aAlgorithm, TSrc aSource, TDst aDestination)
	{
		auto state = aAlgorithm.prepare(aSource, aDestination, mPolicy);
		reserve(state.scratch_bytes());
		return prepared_operation<decltype(state)>{this, std::move(state)};
	}

	template<class TState, class TSrc, class TDst>
	auto run(prepared_operation<TState>& aOperation, TSrc aSource, TDst aDestination)
	{
		if (aOperation.mOwner != this) {
			throw std::invalid_argument("Prepared operation belongs to a different execution context");
		}
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			mSubmitted = true;
		}
		detail::scratch_space<TPolicy> scratch(mScratch.get(), mCapacity);
		return aOperation.mState.run(aSource, aDestination, scratch, mPolicy);
	}

	std::size_t scratch_capacity() const noexcept { return mCapacity; }

private:
	void reserve(std::size_t aSize)
	{
		if (aSize <= mCapacity) {
			return;
		}
		void* storage = nullptr;
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			if (mSubmitted) {
				hip_execution{mPolicy.stream}.wait();
				mSubmitted = false;
			}
			detail::check_status(iane_hip_malloc(&storage, aSize));
		} else {
			storage = ::operator new(aSize);
		}
		mScratch.reset(storage);
		mCapacity = aSize;
	}

	TPolicy mPolicy;
	std::unique_ptr<void, detail::scratch_deleter<TPolicy>> mScratch;
	std::size_t mCapacity{};
	bool mSubmitted{};
};

inline auto make_execution_context(cpu_policy aPolicy = {}) { return execution_context<cpu_policy>(aPolicy); }
inline auto make_execution_context(hip_policy aPolicy) { return execution_context<hip_policy>(aPolicy); }

// ------------------------------------------------------------------ FFT ---
namespace detail {

struct fft_plan_deleter {
	void operator()(iane_fft_plan* aPlan) const noexcept { iane_hip_fft_plan_destroy(aPlan); }
};

template<class V>
std::size_t check_packed_fft_view(const V& aView)
{
	using T = typename V::value_type;
	const auto [count, strides] = resource::detail::packed_layout<T, V::dimension>(aView.dimensions());
	const auto actual = aView.strides();
	for (int i = 0; i < V::dimension; ++i) {
		if (math::detail::axis(actual, i) != math::detail::axis(strides, i)) {
			throw std::invalid_argument("FFT requires packed views");
		}
	}
	const auto address = reinterpret_cast<std::uintptr_t>(aView.pointer());
	if (address == 0 || address % alignof(fft_complex) != 0) {
		throw std::invalid_argument("FFT requires aligned input and output buffers");
	}
	return count * sizeof(T);
}

// A prepared FFT: a plan for one shape, reusable on new packed buffers (fft.hpp:174-193).
template<int Kind, int N>
struct fft_state {
	typename math::dim_traits<N>::extents_t real_size;
	std::unique_ptr<iane_fft_plan, fft_plan_deleter> plan;
	std::size_t workspace{};

	std::size_t scratch_bytes() const { return workspace; }

	template<class TSrc, class TDst>
	hip_execution run(TSrc aSource, TDst aDestination, scratch_space<hip_policy>& aScratch, hip_policy aPolicy)
	{
		check_packed_fft_view(aSource);
		check_packed_fft_view(aDestination);
		const auto real = Kind == IANE_FFT_C2R ? aDestination.dimensions() : aSource.dimensions();
		if (!(real == real_size)) {
			throw std::invalid_argument("FFT views must match the prepared shape");
		}
		void* ws = workspace ? aScratch.bytes(workspace, alignof(fft_complex)) : nullptr;
		check_status(iane_hip_fft_exec(plan.get(), aSource.pointer(), const_cast<void*>(static_cast<const void*>(
			aDestination.pointer())), ws, aPolicy.stream));
		return hip_execution{aPolicy.stream};
	}
};

template<int Kind>
struct fft_algorithm {
	template<class TSrc, class TDst>
	requires(concepts::DeviceView<TSrc> && concepts::DeviceView<TDst> && TSrc::dimension == TDst::dimension &&
		TSrc::dimension <= 3)
	auto prepare(TSrc aSource, TDst aDestination, hip_policy) const
	{
		constexpr int N = TSrc::dimension;
		using in_t = typename TSrc::value_type;
		using out_t = typename TDst::value_type;
		if constexpr (Kind == IANE_FFT_R2C) {
			static_assert(std::same_as<in_t, float> && std::same_as<out_t, fft_complex>);
		} else if constexpr (Kind == IANE_FFT_C2R) {
			static_assert(std::same_as<in_t, fft_complex> && std::same_as<out_t, float>);
		} else {
			static_assert(std::same_as<in_t, fft_complex> && std::same_as<out_t, fft_complex>);
		}
		const auto real = Kind == IANE_FFT_C2R ? aDestination.dimensions() : aSource.dimensions();
		const auto other = Kind == IANE_FFT_C2R ? aSource.dimensions() : aDestination.dimensions();
		int32_t extent[3] = {1, 1, 1};
		for (int i = 0; i < N; ++i) {
			extent[i] = math::detail::axis(real, i);
			if (extent[i] <= 0) {
				throw std::invalid_argument(Kind >= IANE_FFT_R2C ? "Real FFT requires positive dimensions"
				                                                 : "FFT requires positive dimensions");
			}
			const int expect = (Kind >= IANE_FFT_R2C && i == 0) ? extent[0] / 2 + 1 : extent[i];
			if (math::detail::axis(other, i) != expect) {
				throw std::invalid_argument(Kind >= IANE_FFT_R2C
					? "Real FFT requires matching real and packed spectrum extents"
					: "Source and destination require matching nonnegative dimensions");
			}
		}
		const auto in_bytes = check_packed_fft_view(aSource);
		const auto out_bytes = check_packed_fft_view(aDestination);
		const auto src = reinterpret_cast<std::uintptr_t>(aSource.pointer());
		const auto dst = reinterpret_cast<std::uintptr_t>(aDestination.pointer());
		if (src <= dst ? dst - src < in_bytes : src - dst < out_bytes) {
			throw std::invalid_argument("FFT requires disjoint input and output buffers");
		}
		iane_fft_plan* raw = nullptr;
		check_status(iane_hip_fft_plan_create(&raw, N, extent, Kind));
		fft_state<Kind, N> state{real, std::unique_ptr<iane_fft_plan, fft_plan_deleter>(raw), 0};
		check_status(iane_hip_fft_plan_workspace_bytes(raw, &state.workspace));
		return state;
	}
};

} // namespace detail

/// out-of-place packed complex FFT with the negative exponent (fft.hpp:228-229)
inline constexpr detail::fft_algorithm<IANE_FFT_C2C_FORWARD> fft_forward{};
/// unnormalised inverse of fft_forward (fft.hpp:230-231)
inline constexpr detail::fft_algorithm<IANE_FFT_C2C_INVERSE> fft_inverse{};
/// float pixels to the packed half spectrum along axis 0 (rfft.hpp:152-154)
inline constexpr detail::fft_algorithm<IANE_FFT_R2C> rfft{};
/// unnormalised inverse real FFT; the input spectrum is preserved (rfft.hpp:155-161)
inline constexpr detail::fft_algorithm<IANE_FFT_C2R> irfft{};

} // namespace iane
