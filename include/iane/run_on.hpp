// Transfers and the measurement harness.
//
// copy: host<->device and device<->device transfer of equal views
// (include/iane/algorithms/copy.hpp:55-108 of the reference: one copy when both
// sides are packed, per-plane 2-D copies when rows are contiguous).
// run_on / run_prepared_on: the harness of execution/run_on.hpp:86-237 (same
// options, staging rules and log lines: backend=... size=... / setup_ms /
// upload_ms / prepare_ms / compute_ms min= median= max= / download_ms).
#pragma once

#include <iane/convolution.hpp>
#include <iane/prepared.hpp>

#include <algorithm>
#include <chrono>
#include <iostream>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace iane {

template<class TSrc, class TDst>
requires(concepts::MemoryView<TSrc> && concepts::MemoryView<TDst> &&
	(concepts::DeviceView<TSrc> || concepts::DeviceView<TDst>))
auto copy(TSrc aSource, TDst aDestination, hip_policy aPolicy = {})
{
	static_assert(std::same_as<typename TSrc::value_type, typename TDst::value_type>);
	constexpr int kind = concepts::DeviceView<TSrc> ? (concepts::DeviceView<TDst> ? 3 : 2) : 1;
	using T = typename TSrc::value_type;
	constexpr int code = detail::abi_element<T> ? detail::element_code<T>() :
		(sizeof(T) == 8 ? IANE_C64 : (sizeof(T) == 4 ? IANE_F32 : (sizeof(T) == 2 ? IANE_U16 : IANE_U8)));
	static_assert(detail::abi_element<T> || sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8);
	const auto s = detail::describe(aSource, code);
	const auto d = detail::describe(aDestination, code);
	detail::check_status(iane_hip_copy_view(&s, &d, kind, aPolicy.stream));
	return hip_execution{aPolicy.stream};
}

/// Backend chosen at run time (run_on.hpp:21-30; backend::cuda becomes backend::hip).
enum class backend { cpu, hip };

/// Backend, device and measurement settings (run_on.hpp:24-30).
struct run_options {
	iane::backend backend = backend::cpu;
	int device = 0;
	int warmup = 0;
	int runs = 1;
};

namespace detail {

// Milliseconds on the steady clock since construction.
class stopwatch {
public:
	double ms() const { return std::chrono::duration<double, std::milli>(clock::now() - mStart).count(); }

private:
	using clock = std::chrono::steady_clock;
	clock::time_point mStart = clock::now();
};

// Order statistics of the timed runs; the median of an even count is the
// mean of the two middle samples.
struct timing_summary {
	double least;
	double median;
	double most;
};

inline timing_summary summarize(std::vector<double> aSamples)
{
	const auto n = aSamples.size();
	const auto upper = aSamples.begin() + static_cast<std::ptrdiff_t>(n / 2);
	std::nth_element(aSamples.begin(), upper, aSamples.end());
	const double upper_middle = *upper;
	const double lower_middle = n % 2 == 1 ? upper_middle : *std::max_element(aSamples.begin(), upper);
	const auto [least, most] = std::minmax_element(aSamples.begin(), aSamples.end());
	return {*least, 0.5 * (lower_middle + upper_middle), *most};
}

// Untimed warm-up calls, then `runs` calls each timed through completion
// (the callable waits); logs "compute_ms min= median= max=".
template<class TCallAndWait>
void log_compute(const run_options& aOptions, TCallAndWait&& aCallAndWait)
{
	for (int w = 0; w < aOptions.warmup; ++w) {
		aCallAndWait();
	}
	std::vector<double> times(static_cast<std::size_t>(aOptions.runs));
	for (auto& t : times) {
		const stopwatch clock;
		aCallAndWait();
		t = clock.ms();
	}
	const auto summary = summarize(std::move(times));
	std::cout << "compute_ms min=" << summary.least << " median=" << summary.median << " max=" << summary.most
	          << '\n';
}

// "backend=<cpu|hip>[ device=D] size=AxBx.. warmup=W runs=R"
template<int tDim, class TExtents>
void log_header(const run_options& aOptions, const TExtents& aSize)
{
	const bool hip = aOptions.backend == backend::hip;
	std::cout << "backend=" << (hip ? "hip" : "cpu");
	if (hip) {
		std::cout << " device=" << aOptions.device;
	}
	for (int a = 0; a < tDim; ++a) {
		std::cout << (a == 0 ? " size=" : "x") << math::detail::axis(aSize, a);
	}
	std::cout << " warmup=" << aOptions.warmup << " runs=" << aOptions.runs << '\n';
}

// Time an operation: a plain one is invoked per run; a prepared one is built
// once (logged as prepare_ms through completion of the factory's work) and its
// callable is invoked per run and destroyed before returning.
template<bool tPrepared, class TExecution, class TInvoke>
void time_operation(const run_options& aOptions, const TExecution& aExecution, TInvoke&& aInvoke)
{
	if constexpr (tPrepared) {
		const stopwatch clock;
		auto prepared = aInvoke();
		aExecution.wait();
		std::cout << "prepare_ms=" << clock.ms() << '\n';
		log_compute(aOptions, [&] { prepared().wait(); });
	} else {
		log_compute(aOptions, [&] { aInvoke().wait(); });
	}
}

template<class T>
inline constexpr bool device_element = abi_element<T> || std::same_as<T, fft_complex>;

// A private stream on the selected device for one HIP run; on exit it is
// drained and destroyed.
class run_stream {
public:
	explicit run_stream(int aDevice)
	{
		check_status(iane_hip_set_device(aDevice));
		check_status(iane_hip_stream_create(&mHandle));
	}
	run_stream(const run_stream&) = delete;
	run_stream& operator=(const run_stream&) = delete;
	~run_stream()
	{
		iane_hip_stream_sync(mHandle);
		iane_hip_stream_destroy(mHandle);
	}
	hip_execution execution() const { return {mHandle}; }

private:
	void* mHandle = nullptr;
};

// One host input with its device copy.
template<class THost>
struct staged_input {
	THost host;
	device_image<typename THost::value_type, THost::dimension> device;

	explicit staged_input(THost aHost) : host(aHost), device(aHost.dimensions()) {}
	void upload(hip_execution aExecution) const { copy(host, device.view(), aExecution); }
	auto view() const { return device.const_view(); }
};

template<class TDst, class... TSrc>
auto run_extent(const std::tuple<TSrc...>& aSources, const TDst& aDestination)
{
	if constexpr (sizeof...(TSrc) == 0) {
		return std::pair{aDestination.dimensions(), std::integral_constant<int, TDst::dimension>{}};
	} else {
		using first = std::tuple_element_t<0, std::tuple<TSrc...>>;
		return std::pair{std::get<0>(aSources).dimensions(), std::integral_constant<int, first::dimension>{}};
	}
}

template<bool tPrepared, class TDst, class TOperation, class... TSrc>
void run_on_backend(const run_options& aOptions, const std::tuple<TSrc...>& aSources, TDst aDestination,
	TOperation& aOperation)
{
	if (aOptions.warmup < 0 || aOptions.runs <= 0 || aOptions.device < 0) {
		throw std::invalid_argument("Invalid run options");
	}
	const auto [extent, rank] = run_extent(aSources, aDestination);
	log_header<decltype(rank)::value>(aOptions, extent);

	if (aOptions.backend == backend::cpu) {
		cpu_execution host;
		time_operation<tPrepared>(aOptions, host, [&]() -> decltype(auto) {
			return std::apply([&](const auto&... aInput) { return aOperation(aInput..., aDestination, host); },
				aSources);
		});
		return;
	}

	constexpr bool supported = (device_element<typename TSrc::value_type> && ...) &&
		device_element<typename TDst::value_type>;
	if constexpr (!supported) {
		// custom element types (pixel structs, ...) run on the host backend only
		throw std::invalid_argument("HIP execution supports u8, u16, i32, f32, f64 and fft_complex elements");
	} else {
		// setup: device, stream and device images for every input and the output
		const stopwatch setup_clock;
		const run_stream stream(aOptions.device);
		auto staged = std::apply([](const auto&... aInput) { return std::tuple{staged_input(aInput)...}; }, aSources);
		device_image<typename TDst::value_type, TDst::dimension> result(aDestination.dimensions());
		auto execution = stream.execution();
		std::cout << "setup_ms=" << setup_clock.ms();
		if constexpr (sizeof...(TSrc) > 0) {
			const stopwatch upload_clock;
			std::apply([&](const auto&... aInput) { (aInput.upload(execution), ...); }, staged);
			execution.wait();
			std::cout << " upload_ms=" << upload_clock.ms();
		}
		std::cout << '\n';
		time_operation<tPrepared>(aOptions, execution, [&]() -> decltype(auto) {
			return std::apply([&](const auto&... aInput) { return aOperation(aInput.view()..., result.view(), execution); },
				staged);
		});
		const stopwatch download_clock;
		copy(result.const_view(), aDestination, execution).wait();
		std::cout << "download_ms=" << download_clock.ms() << '\n';
	}
}

} // namespace detail

/**
 * Run `aOperation(inputs..., output, execution)` on the host views, or on
 * device copies of them (run_on.hpp:178-201). The HIP backend stages the
 * inputs once, runs warm-up and timed runs, then downloads the output; the
 * destination is ready on return. Logs the reference's timing lines.
 * @throws std::invalid_argument for negative warm-up, non-positive runs or device.
 */
template<class TDst, class TOperation, class... TSrc>
requires(concepts::HostView<TDst> && (concepts::HostView<TSrc> && ...))
void run_on(const run_options& aOptions, std::tuple<TSrc...> aSources, TDst aDestination, TOperation&& aOperation)
{
	detail::run_on_backend<false>(aOptions, aSources, aDestination, aOperation);
}

/**
 * Like run_on, but `aPrepare(inputs..., output, execution)` runs once and
 * returns a callable that produces the output and its completion handle
 * (run_on.hpp:203-221); its time is logged as prepare_ms.
 */
template<class TDst, class TPrepare, class... TSrc>
requires(concepts::HostView<TDst> && (concepts::HostView<TSrc> && ...))
void run_prepared_on(const run_options& aOptions, std::tuple<TSrc...> aSources, TDst aDestination, TPrepare&& aPrepare)
{
	detail::run_on_backend<true>(aOptions, aSources, aDestination, aPrepare);
}

/// One input (run_on.hpp:223-229).
template<class TSrc, class TDst, class TOperation>
requires(concepts::HostView<TSrc> && concepts::HostView<TDst>)
void run_on(const run_options& aOptions, TSrc aSource, TDst aDestination, TOperation&& aOperation)
{
	run_on(aOptions, std::tuple{aSource}, aDestination, std::forward<TOperation>(aOperation));
}

/// No input: a generator writing the destination (run_on.hpp:231-237).
template<class TDst, class TOperation>
requires(concepts::HostView<TDst>)
void run_on(const run_options& aOptions, TDst aDestination, TOperation&& aOperation)
{
	run_on(aOptions, std::tuple<>{}, aDestination, std::forward<TOperation>(aOperation));
}

} // namespace iane
