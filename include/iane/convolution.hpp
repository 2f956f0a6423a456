// iane::convolve drop-in (include/iane/algorithms/convolution.hpp:97-117 of the reference).
//
// Device views go to iane_hip_convolve (LDS-tiled HIP kernels on gfx950);
// host views run the reference's CPU semantics here: taps in kernel storage
// order, FP64 fma per tap for arithmetic pairs (convolution.hpp:36-49), the
// border rule per axis (locator/image_locator.hpp:40-56), the output
// conversion (:51-75), and the cpu_policy partition over std::jthread workers
// (algorithms/for_each.hpp:75-134). Custom element types accumulate in the
// product type on the host path only.
#pragma once

#include <iane/prepared.hpp>

#include <cmath>
#include <exception>
#include <thread>
#include <vector>

namespace iane {

/// locator/image_locator.hpp:14
enum class border_mode { clamp, zero };

/// Arithmetic of a device convolution (no reference counterpart; the
/// reference always accumulates in FP64, which is `reference`).
enum class convolution_precision { reference = IANE_PREC_REFERENCE, fp32 = IANE_PREC_FP32,
	split_bf16 = IANE_PREC_SPLIT_BF16, split_fp16 = IANE_PREC_SPLIT_FP16 };

/// convolution.hpp:16-18 (+ the device arithmetic, default = the reference's)
struct convolution_options {
	border_mode border = border_mode::clamp;
	convolution_precision precision = convolution_precision::reference;
};

namespace detail {

template<class S, class K>
inline constexpr bool arithmetic_convolution = std::is_arithmetic_v<S> && std::is_arithmetic_v<K>;

template<class S, class K>
using convolution_sum = std::conditional_t<arithmetic_convolution<S, K>, double,
	std::remove_cvref_t<decltype(std::declval<S>() * std::declval<K>())>>;

// Output conversion of convolution.hpp:51-75.
template<class R, class Sum>
R convert_result(Sum aSum)
{
	if constexpr (std::is_arithmetic_v<Sum> && std::same_as<R, std::uint8_t>) {
		if (!(aSum > 0.0)) {
			return 0;
		}
		if (aSum >= 255.0) {
			return 255;
		}
		const auto integer = static_cast<unsigned int>(aSum);
		return static_cast<R>(integer + (aSum - integer >= 0.5));
	} else if constexpr (std::is_arithmetic_v<Sum> && std::integral<R>) {
		constexpr auto lower = std::numeric_limits<R>::lowest();
		constexpr auto upper = std::numeric_limits<R>::max();
		if (aSum != aSum) {
			return {};
		}
		if (aSum <= static_cast<double>(lower)) {
			return lower;
		}
		if (aSum >= static_cast<double>(upper)) {
			return upper;
		}
		return static_cast<R>(aSum);
	} else {
		return static_cast<R>(aSum);
	}
}

template<class TSrc, class TKernel, class TDst>
void check_convolution(const TSrc& aSource, const TKernel& aKernel, const TDst& aDestination)
{
	const auto size = aKernel.dimensions();
	for (int i = 0; i < TKernel::dimension; ++i) {
		const auto extent = math::detail::axis(size, i);
		if (extent <= 0 || extent % 2 == 0) {
			throw std::invalid_argument("Convolution requires positive odd kernel dimensions");
		}
	}
	const auto out = aDestination.dimensions();
	const auto in = aSource.dimensions();
	for (int i = 0; i < TDst::dimension; ++i) {
		if (math::detail::axis(out, i) < 0 || math::detail::axis(in, i) != math::detail::axis(out, i)) {
			throw std::invalid_argument("Source and destination require matching nonnegative dimensions");
		}
	}
}

// One output sample (convolution.hpp:23-77 with the border of image_locator.hpp:40-56).
template<class TSrc, class TKernel, class TDst>
typename TDst::value_type convolve_one(const TSrc& aSource, const TKernel& aKernel, border_mode aBorder,
	const typename TDst::coord_t& aPosition)
{
	using S = typename TSrc::value_type;
	using K = typename TKernel::value_type;
	constexpr int N = TKernel::dimension;
	const auto ksize = aKernel.dimensions();
	const auto ssize = aSource.dimensions();
	convolution_sum<S, K> sum{};
	typename TKernel::coord_t q{};
	do {
		typename TSrc::coord_t sample{};
		bool zero = false;
		for (int a = 0; a < N; ++a) {
			const long long c = static_cast<long long>(math::detail::axis(aPosition, a)) +
				math::detail::axis(ksize, a) / 2 - math::detail::axis(q, a);
			const int extent = math::detail::axis(ssize, a);
			if (c >= 0 && c < extent) {
				math::detail::axis(sample, a) = static_cast<int>(c);
			} else if (aBorder == border_mode::zero || extent <= 0) {
				zero = true;
				break;
			} else {
				math::detail::axis(sample, a) = c < 0 ? 0 : extent - 1;
			}
		}
		const S value = zero ? S{} : aSource[sample];
		if constexpr (arithmetic_convolution<S, K>) {
			sum = std::fma(static_cast<double>(value), static_cast<double>(aKernel[q]), sum);
		} else {
			sum = sum + value * aKernel[q];
		}
	} while (math::detail::advance<N>(q, ksize));
	return convert_result<typename TDst::value_type>(sum);
}

// The per-output callable of the host traversal (each worker gets a copy).
template<class TSrc, class TKernel, class TDst>
struct convolve_visitor {
	TSrc source;
	TKernel kernel;
	TDst destination;
	border_mode border;
	void operator()(const typename TDst::coord_t& aPosition) const
	{
		destination[aPosition] = convolve_one<TSrc, TKernel, TDst>(source, kernel, border, aPosition);
	}
};

// for_each.hpp:75-134: serial storage order for one worker, otherwise
// contiguous balanced task ranges over std::jthread, first exception rethrown.
template<class TDst, class F>
void visit_host(const TDst& aView, F aFunction, cpu_policy aPolicy)
{
	if (aPolicy.threads == 0 || aPolicy.grain_size == 0) {
		throw std::invalid_argument("CPU policy requires positive threads and grain size");
	}
	constexpr int N = TDst::dimension;
	const auto size = aView.dimensions();
	const auto count = math::detail::element_count<N>(size);
	if (count == 0) {
		return;
	}
	const auto tasks = 1 + (count - 1) / aPolicy.grain_size;
	const auto workers = std::min(aPolicy.threads, tasks);
	if (workers == 1) {
		typename TDst::coord_t position{};
		do {
			aFunction(position);
		} while (math::detail::advance<N>(position, size));
		return;
	}
	std::vector<std::exception_ptr> errors(workers);
	auto run = [&](std::size_t aWorker, F& aCallback) {
		try {
			const auto base = tasks / workers;
			const auto extra = tasks % workers;
			const auto first = aWorker * base + std::min(aWorker, extra);
			const auto last = first + base + (aWorker < extra);
			for (auto task = first; task < last; ++task) {
				const auto begin = task * aPolicy.grain_size;
				const auto length = std::min(aPolicy.grain_size, count - begin);
				auto position = math::detail::coordinate_at<N, typename TDst::coord_t>(begin, size);
				for (std::size_t i = 0; i < length; ++i) {
					aCallback(position);
					math::detail::advance<N>(position, size);
				}
			}
		} catch (...) {
			errors[aWorker] = std::current_exception();
		}
	};
	{
		std::vector<std::jthread> threads;
		threads.reserve(workers - 1);
		for (std::size_t w = 1; w < workers; ++w) {
			threads.emplace_back([&, w, callback = aFunction]() mutable { run(w, callback); });
		}
		run(0, aFunction);
	}
	for (const auto& error : errors) {
		if (error) {
			std::rethrow_exception(error);
		}
	}
}

} // namespace detail

/**
 * Convolve a source with an odd-sized, centred kernel of matching rank
 * (convolution.hpp:81-117): the sample at p - d is multiplied by the
 * coefficient at offset d. Device views enqueue on the policy's stream and
 * return a handle; host views complete before returning.
 * @throws std::invalid_argument for non-odd kernels or mismatched extents.
 */
template<class TSrc, class TKernel, class TDst>
requires(concepts::AnyView<TSrc> && concepts::AnyView<TKernel> && concepts::AnyView<TDst> &&
	TSrc::dimension == TKernel::dimension && TSrc::dimension == TDst::dimension &&
	concepts::DeviceView<TSrc> == concepts::DeviceView<TKernel> &&
	concepts::DeviceView<TSrc> == concepts::DeviceView<TDst>)
auto convolve(TSrc aSource, TKernel aKernel, TDst aDestination, convolution_options aOptions = {},
	view_policy<TDst> aPolicy = {})
{
	detail::check_convolution(aSource, aKernel, aDestination);
	if constexpr (concepts::DeviceView<TDst>) {
		static_assert(detail::abi_element<typename TSrc::value_type> && detail::abi_element<typename TKernel::value_type> &&
			detail::abi_element<typename TDst::value_type>,
			"device convolution supports u8, u16, i32, f32 and f64 elements");
		const auto s = detail::describe(aSource);
		const auto k = detail::describe(aKernel);
		const auto d = detail::describe(aDestination);
		std::size_t scratch = 0;
		detail::check_status(iane_hip_convolve_workspace_bytes(&s, &k, &d, static_cast<int>(aOptions.border),
			static_cast<int>(aOptions.precision), &scratch));
		if (scratch != 0) {
			throw std::invalid_argument(
				"Split-precision convolution needs scratch: prepare iane::convolution in an execution_context");
		}
		detail::check_status(iane_hip_convolve(&s, &k, &d, static_cast<int>(aOptions.border),
			static_cast<int>(aOptions.precision), nullptr, 0, aPolicy.stream));
		return hip_execution{aPolicy.stream};
	} else {
		detail::visit_host(aDestination, detail::convolve_visitor<TSrc, TKernel, TDst>{aSource, aKernel, aDestination,
			aOptions.border}, aPolicy);
		return cpu_execution{};
	}
}

namespace detail {

// Prepared convolution with a fixed kernel (no reference counterpart: the
// reference's convolve needs no scratch). The split-precision MFMA path takes
// its padded planes and Toeplitz table from the context's workspace.
template<class TKernel>
struct convolution_state {
	TKernel kernel;
	convolution_options options;
	std::size_t workspace;
	iane_view_desc source_shape;

	std::size_t scratch_bytes() const { return workspace; }

	template<class TSrc, class TDst, class TPolicy>
	auto run(TSrc aSource, TDst aDestination, scratch_space<TPolicy>& aScratch, TPolicy aPolicy)
	{
		if constexpr (std::same_as<TPolicy, hip_policy>) {
			const auto s = describe(aSource);
			const auto k = describe(kernel);
			const auto d = describe(aDestination);
			for (int i = 0; i < TSrc::dimension; ++i) {
				if (s.extent[i] != source_shape.extent[i]) {
					throw std::invalid_argument("Convolution views must match the prepared shape");
				}
			}
			// the region is reserved 256 bytes long so its start can move to a 256-byte boundary
			void* ws = nullptr;
			std::size_t usable = 0;
			if (workspace) {
				const auto raw = reinterpret_cast<std::uintptr_t>(aScratch.bytes(workspace));
				ws = reinterpret_cast<void*>((raw + 255) & ~static_cast<std::uintptr_t>(255));
				usable = workspace - 256;
			}
			check_status(iane_hip_convolve(&s, &k, &d, static_cast<int>(options.border),
				static_cast<int>(options.precision), ws, usable, aPolicy.stream));
			return hip_execution{aPolicy.stream};
		} else {
			return convolve(aSource, kernel, aDestination, options, aPolicy);
		}
	}
};

template<class TKernel>
struct convolution_algorithm {
	TKernel kernel;
	convolution_options options;

	template<class TSrc, class TDst, class TPolicy>
	auto prepare(TSrc aSource, TDst aDestination, TPolicy) const
	{
		check_convolution(aSource, kernel, aDestination);
		std::size_t bytes = 0;
		iane_view_desc shape{};
		if constexpr (concepts::DeviceView<TDst>) {
			const auto s = describe(aSource);
			const auto k = describe(kernel);
			const auto d = describe(aDestination);
			check_status(iane_hip_convolve_workspace_bytes(&s, &k, &d, static_cast<int>(options.border),
				static_cast<int>(options.precision), &bytes));
			shape = s;
		}
		// the scratch region is carved at max_align_t; the library wants 256-byte starts
		return convolution_state<TKernel>{kernel, options, bytes ? bytes + 256 : 0, shape};
	}
};

} // namespace detail

/**
 * A convolution with a fixed kernel as a prepared algorithm:
 * `context.prepare(iane::convolution(kernel, options), src, dst)` reserves the
 * scratch the chosen arithmetic needs (split-precision MFMA only) in the
 * context's workspace, `context.run(op, src, dst)` convolves views of the
 * prepared shape. No reference counterpart.
 */
template<class TKernel>
requires(concepts::AnyView<TKernel>)
auto convolution(TKernel aKernel, convolution_options aOptions = {})
{
	return detail::convolution_algorithm<TKernel>{aKernel, aOptions};
}

/**
 * Separable device convolution: the kernel is the outer product of one
 * odd-length f32/f64 tap vector per axis (no reference counterpart). The
 * single-kernel path needs no scratch; shapes it does not cover run one pass
 * per axis through two intermediate images taken from `aScratch` (a region of
 * at least separable_scratch_bytes), or, when none is given, refused.
 */
template<class TSrc, class TTaps, class TDst, std::size_t R>
requires(concepts::DeviceView<TSrc> && concepts::DeviceView<TDst> && TTaps::dimension == 1 && R == TSrc::dimension)
std::size_t separable_scratch_bytes(TSrc aSource, const TTaps (&aTaps)[R], TDst aDestination,
	convolution_options aOptions = {border_mode::clamp, convolution_precision::fp32})
{
	const auto s = detail::describe(aSource);
	const auto d = detail::describe(aDestination);
	iane_view_desc taps[R];
	for (std::size_t i = 0; i < R; ++i) {
		taps[i] = detail::describe(aTaps[i]);
	}
	std::size_t bytes = 0;
	detail::check_status(iane_hip_convolve_separable_workspace_bytes(&s, taps, static_cast<int>(R), &d,
		static_cast<int>(aOptions.border), static_cast<int>(aOptions.precision), &bytes));
	return bytes;
}

template<class TSrc, class TTaps, class TDst, std::size_t R>
requires(concepts::DeviceView<TSrc> && concepts::DeviceView<TDst> && TTaps::dimension == 1 && R == TSrc::dimension)
hip_execution convolve_separable(TSrc aSource, const TTaps (&aTaps)[R], TDst aDestination,
	convolution_options aOptions = {border_mode::clamp, convolution_precision::fp32}, hip_policy aPolicy = {},
	void* aScratch = nullptr, std::size_t aScratchBytes = 0)
{
	const auto s = detail::describe(aSource);
	const auto d = detail::describe(aDestination);
	iane_view_desc taps[R];
	for (std::size_t i = 0; i < R; ++i) {
		taps[i] = detail::describe(aTaps[i]);
	}
	detail::check_status(iane_hip_convolve_separable(&s, taps, static_cast<int>(R), &d,
		static_cast<int>(aOptions.border), static_cast<int>(aOptions.precision), aScratch, aScratchBytes,
		aPolicy.stream));
	return hip_execution{aPolicy.stream};
}

} // namespace iane
