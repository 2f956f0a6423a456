// Prepared algorithms with one shared workspace and the FFT drop-ins.
//
// execution_context / prepared_operation / scratch_space / make_execution_context
// keep the interface of include/iane/execution/prepared.hpp:18-180 of the
// reference: prepare() reserves the operation's scratch in the context's one
// workspace, run() hands the operation that workspace, an operation runs only
// on its own context, and growth and destruction wait for the last device
// submission.
// fft_forward / fft_inverse / rfft / irfft / rfft_extent / fft_complex mirror
// include/iane/algorithms/fft/{fft.hpp:196-231, rfft.hpp:107-161, layout.hpp:15-21,
// complex.hpp:6-9} for views of rank 1..3: device views on the Stockham FFT of
// libiane_hip.so, host views on the header's own FFT (include/iane/host_fft.hpp,
// in the role of the reference's FFTW plans).
#pragma once

#include <iane/core.hpp>
#include <iane/host_fft.hpp>

#include <array>
#include <atomic>
#include <cstring>
#include <memory>
#include <optional>

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

/**
 * Bump allocator over one borrowed byte range: the view of the context's
 * workspace that one prepared operation sees while it runs. Regions are
 * handed out front to back; nothing is freed (the range is reused by the next
 * submission on the same stream).
 */
template<class TPolicy>
class scratch_space {
public:
	scratch_space(void* aBase, std::size_t aLength)
		: mBase(reinterpret_cast<std::uintptr_t>(aBase)), mCursor(0), mLength(aBase ? aLength : 0) {}

	/// A region of `aSize` bytes aligned to the power of two `aAlignment`
	/// (at most alignof(max_align_t)); empty requests return null.
	void* bytes(std::size_t aSize, std::size_t aAlignment = alignof(std::max_align_t))
	{
		if (aSize == 0) {
			return nullptr;
		}
		const bool power_of_two = aAlignment != 0 && (aAlignment & (aAlignment - 1)) == 0;
		if (!power_of_two || aAlignment > alignof(std::max_align_t)) {
			throw std::invalid_argument("Unsupported scratch alignment");
		}
		const std::uintptr_t first = (mBase + mCursor + aAlignment - 1) & ~static_cast<std::uintptr_t>(aAlignment - 1);
		const std::size_t offset = static_cast<std::size_t>(first - mBase);
		if (mBase == 0 || offset > mLength || mLength - offset < aSize) {
			throw std::length_error("Prepared operation exceeded its scratch reservation");
		}
		mCursor = offset + aSize;
		return reinterpret_cast<void*>(first);
	}

	/// A packed image of the backend's kind carved out of the range.
	template<class TElement, int tDim>
	auto image(const typename math::dim_traits<tDim>::extents_t& aSize)
	{
		static_assert(std::is_trivially_copyable_v<TElement>);
		const auto count = resource::detail::packed_layout<TElement, tDim>(aSize).first;
		auto* first = static_cast<TElement*>(bytes(count * sizeof(TElement), alignof(TElement)));
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			return device::make_view<TElement, tDim>(first, aSize);
		} else {
			return host::make_view<TElement, tDim>(first, aSize);
		}
	}

	std::size_t remaining() const noexcept { return mLength - mCursor; }

private:
	std::uintptr_t mBase;
	std::size_t mCursor;
	std::size_t mLength;
};

// Where a context's workspace comes from: host heap or device memory.
template<class TPolicy>
struct workspace_allocator {
	static void* acquire(std::size_t aBytes) { return ::operator new(aBytes); }
	static void release(void* aBlock) noexcept { ::operator delete(aBlock); }
};

template<>
struct workspace_allocator<hip_policy> {
	static void* acquire(std::size_t aBytes)
	{
		void* block = nullptr;
		check_status(iane_hip_malloc(&block, aBytes));
		return block;
	}
	static void release(void* aBlock) noexcept { iane_hip_free(aBlock); }
};

// Identity of a context, so an operation can refuse to run on another one.
inline std::uint64_t next_context_serial() noexcept
{
	static std::atomic<std::uint64_t> counter{0};
	return ++counter;
}

} // namespace detail

/**
 * Prepared algorithm state bound to the execution_context that reserved its
 * scratch (the reference's prepared_operation, include/iane/execution/prepared.hpp:88-98).
 * Only its context can run it.
 */
template<class TState>
class prepared_operation {
public:
	prepared_operation(prepared_operation&&) noexcept = default;
	prepared_operation& operator=(prepared_operation&&) noexcept = default;

private:
	template<class>
	friend class execution_context;
	prepared_operation(std::uint64_t aContext, TState&& aState) : mContext(aContext), mPrepared(std::move(aState)) {}
	std::uint64_t mContext;
	TState mPrepared;
};

/**
 * Prepares algorithms and runs them with one workspace shared by all of them
 * (the reference's execution_context, include/iane/execution/prepared.hpp:84-180).
 *
 * The workspace only grows: preparing an algorithm that needs more than the
 * current capacity replaces it. A HIP context enqueues on the borrowed stream
 * and reuses the workspace by stream order; before it replaces or frees the
 * workspace it waits for the last submission it made. Submissions must come
 * from one host thread at a time; overlapping work needs separate contexts.
 */
template<class TPolicy>
class execution_context {
	using allocator = detail::workspace_allocator<TPolicy>;
	static constexpr bool device_backend = std::same_as<TPolicy, hip_policy>;

public:
	explicit execution_context(TPolicy aPolicy) : mPolicy(aPolicy), mSerial(detail::next_context_serial()) {}
	execution_context(const execution_context&) = delete;
	execution_context& operator=(const execution_context&) = delete;
	~execution_context()
	{
		drain_noexcept();
		allocator::release(mWorkspace);
	}

	/// Prepare `aAlgorithm` for views shaped like the prototypes and make room for its scratch.
	template<class TAlgorithm, class TSrc, class TDst>
	auto prepare(TAlgorithm aAlgorithm, TSrc aSource, TDst aDestination)
	{
		auto state = aAlgorithm.prepare(aSource, aDestination, mPolicy);
		grow(state.scratch_bytes());
		return prepared_operation<decltype(state)>(mSerial, std::move(state));
	}

	/// Run a prepared operation on views of its prepared shape; returns the completion handle.
	template<class TState, class TSrc, class TDst>
	auto run(prepared_operation<TState>& aOperation, TSrc aSource, TDst aDestination)
	{
		if (aOperation.mContext != mSerial) {
			throw std::invalid_argument("Prepared operation belongs to a different execution context");
		}
		detail::scratch_space<TPolicy> scratch(mWorkspace, mBytes);
		auto done = aOperation.mPrepared.run(aSource, aDestination, scratch, mPolicy);
		if constexpr (device_backend) {
			mInFlight = done;
		}
		return done;
	}

	/// Bytes of the shared workspace currently held.
	std::size_t scratch_capacity() const noexcept { return mBytes; }

private:
	void grow(std::size_t aBytes)
	{
		if (aBytes <= mBytes) {
			return;
		}
		if constexpr (device_backend) {
			if (mInFlight) {
				mInFlight->wait();
				mInFlight.reset();
			}
		}
		void* larger = allocator::acquire(aBytes);
		allocator::release(mWorkspace);
		mWorkspace = larger;
		mBytes = aBytes;
	}

	void drain_noexcept() noexcept
	{
		if constexpr (device_backend) {
			if (mInFlight) {
				iane_hip_stream_sync(mInFlight->stream);
			}
		}
	}

	TPolicy mPolicy;
	std::uint64_t mSerial;
	void* mWorkspace = nullptr;
	std::size_t mBytes = 0;
	std::optional<hip_execution> mInFlight; // last device submission (HIP contexts only)
};

inline auto make_execution_context(cpu_policy aPolicy = {}) { return execution_context<cpu_policy>(aPolicy); }
inline auto make_execution_context(hip_policy aPolicy) { return execution_context<hip_policy>(aPolicy); }

// ------------------------------------------------------------------ FFT ---
namespace detail {

struct fft_plan_deleter {
	void operator()(iane_fft_plan* aPlan) const noexcept { iane_hip_fft_plan_destroy(aPlan); }
};

template<class V>
std::size_t check_packed_fft_view(const V& aView, std::size_t aAlignment = alignof(fft_complex))
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
	if (address == 0 || address % aAlignment != 0) {
		throw std::invalid_argument("FFT requires aligned input and output buffers");
	}
	return count * sizeof(T);
}

// The checks of fft.hpp:29-66 and rfft.hpp:29-45 for both backends: element
// types, positive extents, the packed half spectrum of real transforms, packed
// aligned views (host real buffers need float alignment, as for FFTW), and
// disjoint buffers. Returns the real extents (1 past the rank).
template<int Kind, class TSrc, class TDst>
std::array<int32_t, 3> check_fft_views(const TSrc& aSource, const TDst& aDestination, std::size_t aRealAlignment)
{
	constexpr int N = TSrc::dimension;
	using in_t = typename TSrc::value_type;
	using out_t = typename TDst::value_type;
	if constexpr (Kind == IANE_FFT_R2C) {
		static_assert(std::same_as<std::remove_const_t<in_t>, float> && std::same_as<out_t, fft_complex>);
	} else if constexpr (Kind == IANE_FFT_C2R) {
		static_assert(std::same_as<std::remove_const_t<in_t>, fft_complex> && std::same_as<out_t, float>);
	} else {
		static_assert(std::same_as<std::remove_const_t<in_t>, fft_complex> && std::same_as<out_t, fft_complex>);
	}
	const auto real = Kind == IANE_FFT_C2R ? aDestination.dimensions() : aSource.dimensions();
	const auto other = Kind == IANE_FFT_C2R ? aSource.dimensions() : aDestination.dimensions();
	std::array<int32_t, 3> extent{1, 1, 1};
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
	const auto in_bytes = check_packed_fft_view(aSource, Kind == IANE_FFT_R2C ? aRealAlignment : alignof(fft_complex));
	const auto out_bytes = check_packed_fft_view(aDestination,
		Kind == IANE_FFT_C2R ? aRealAlignment : alignof(fft_complex));
	const auto src = reinterpret_cast<std::uintptr_t>(aSource.pointer());
	const auto dst = reinterpret_cast<std::uintptr_t>(aDestination.pointer());
	if (src <= dst ? dst - src < in_bytes : src - dst < out_bytes) {
		throw std::invalid_argument("FFT requires disjoint input and output buffers");
	}
	return extent;
}

// A prepared host FFT: the line transforms of each axis, reusable on new packed
// buffers of the prepared shape. C2R works on a private copy of its input
// spectrum in the context's scratch (rfft.hpp:70-82), so the caller's spectrum
// stays as it was.
template<int Kind, int N>
struct host_fft_state {
	std::array<int32_t, 3> extent;  // real extents
	std::array<std::shared_ptr<const host_fft_line>, 3> lines;
	std::size_t copy_bytes = 0;

	std::size_t scratch_bytes() const { return copy_bytes; }

	template<class TSrc, class TDst>
	cpu_execution run(TSrc aSource, TDst aDestination, scratch_space<cpu_policy>& aScratch, cpu_policy) const
	{
		if (check_fft_views<Kind>(aSource, aDestination, alignof(float)) != extent) {
			throw std::invalid_argument("FFT views must match the prepared shape");
		}
		constexpr bool real = Kind >= IANE_FFT_R2C;
		const int n0 = extent[0];
		int work[3] = {real ? n0 / 2 + 1 : n0, extent[1], extent[2]};
		const std::size_t rows = static_cast<std::size_t>(work[1]) * work[2];
		const std::size_t count = static_cast<std::size_t>(work[0]) * rows;
		std::vector<host_complex> data(count);
		if constexpr (Kind == IANE_FFT_C2R) {
			auto* copy = static_cast<fft_complex*>(aScratch.bytes(copy_bytes, alignof(fft_complex)));
			std::memcpy(copy, aSource.pointer(), copy_bytes);
			for (std::size_t i = 0; i < count; ++i) {
				data[i] = host_complex(copy[i].real, copy[i].imag);
			}
			for (int axis = N - 1; axis >= 1; --axis) {
				host_fft_axis(data, work, N, axis, *lines[axis]);
			}
			// axis 0: the full Hermitian line of each half spectrum, inverse, real part
			std::vector<host_complex> full(static_cast<std::size_t>(n0)), out(static_cast<std::size_t>(n0));
			float* dst = aDestination.pointer();
			for (std::size_t r = 0; r < rows; ++r) {
				const host_complex* half = data.data() + r * work[0];
				for (int k = 0; k < n0; ++k) {
					full[k] = k < work[0] ? half[k] : std::conj(half[n0 - k]);
				}
				lines[0]->transform(full.data(), 1, out.data());
				for (int t = 0; t < n0; ++t) {
					dst[r * n0 + t] = static_cast<float>(out[t].real());
				}
			}
		} else {
			if constexpr (Kind == IANE_FFT_R2C) {
				const float* src = aSource.pointer();
				std::vector<host_complex> in(static_cast<std::size_t>(n0)), out(static_cast<std::size_t>(n0));
				for (std::size_t r = 0; r < rows; ++r) {
					for (int t = 0; t < n0; ++t) {
						in[t] = host_complex(src[r * n0 + t], 0.0);
					}
					lines[0]->transform(in.data(), 1, out.data());
					std::copy(out.begin(), out.begin() + work[0], data.begin() + static_cast<std::ptrdiff_t>(r * work[0]));
				}
			} else {
				const fft_complex* src = aSource.pointer();
				for (std::size_t i = 0; i < count; ++i) {
					data[i] = host_complex(src[i].real, src[i].imag);
				}
				host_fft_axis(data, work, N, 0, *lines[0]);
			}
			for (int axis = 1; axis < N; ++axis) {
				host_fft_axis(data, work, N, axis, *lines[axis]);
			}
			fft_complex* dst = aDestination.pointer();
			for (std::size_t i = 0; i < count; ++i) {
				dst[i] = fft_complex{static_cast<float>(data[i].real()), static_cast<float>(data[i].imag())};
			}
		}
		return {};
	}
};

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
		auto extent = check_fft_views<Kind>(aSource, aDestination, alignof(fft_complex));
		const auto real = Kind == IANE_FFT_C2R ? aDestination.dimensions() : aSource.dimensions();
		iane_fft_plan* raw = nullptr;
		check_status(iane_hip_fft_plan_create(&raw, N, extent.data(), Kind));
		fft_state<Kind, N> state{real, std::unique_ptr<iane_fft_plan, fft_plan_deleter>(raw), 0};
		check_status(iane_hip_fft_plan_workspace_bytes(raw, &state.workspace));
		return state;
	}

	template<class TSrc, class TDst>
	requires(concepts::HostView<TSrc> && concepts::HostView<TDst> && concepts::MemoryView<TSrc> &&
		concepts::MemoryView<TDst> && TSrc::dimension == TDst::dimension && TSrc::dimension <= 3)
	auto prepare(TSrc aSource, TDst aDestination, cpu_policy) const
	{
		constexpr int N = TSrc::dimension;
		host_fft_state<Kind, N> state{check_fft_views<Kind>(aSource, aDestination, alignof(float)), {}, 0};
		constexpr bool inverse = Kind == IANE_FFT_C2C_INVERSE || Kind == IANE_FFT_C2R;
		for (int axis = 0; axis < N; ++axis) {
			state.lines[axis] = std::make_shared<const host_fft_line>(state.extent[axis], inverse);
		}
		if constexpr (Kind == IANE_FFT_C2R) {
			state.copy_bytes = static_cast<std::size_t>(state.extent[0] / 2 + 1) * state.extent[1] * state.extent[2] *
				sizeof(fft_complex);
		}
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
