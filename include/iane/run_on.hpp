// Transfers and the measurement harness.
//
// copy: host<->device and device<->device transfer of equal views
// (include/iane/algorithms/copy.hpp:55-108 of the reference: one copy when both
// sides are packed, per-plane 2-D copies when rows are contiguous).
// run_on / run_prepared_on: execution/run_on.hpp:86-237 — stage host inputs on
// the device (or run on the host), time warm-up + samples through completion
// on the steady clock, and print the same log lines
// (backend=... size=... / setup_ms / upload_ms / prepare_ms / compute_ms min= median= max= / download_ms).
#pragma once

#include <iane/convolution.hpp>
#include <iane/prepared.hpp>

#include <chrono>
#include <iostream>
#include <numeric>
#include <optional>
#include <tuple>
#include <vector>

namespace iane {

template<class TSrc, class TDst>
requires(concepts::MemoryView<TSrc> && concepts::MemoryView<TDst>)
auto copy(TSrc aSource, TDst aDestination, hip_policy aPolicy = {})
{
	static_assert(std::same_as<typename TSrc::value_type, typename TDst::value_type>);
	static_assert(concepts::DeviceView<TSrc> || concepts::DeviceView<TDst>, "host-to-host copy is a plain loop");
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

/// run_on.hpp:21-30 (backend::cuda becomes backend::hip)
enum class backend { cpu, hip };

struct run_options {
	iane::backend backend = backend::cpu;
	int device = 0;
	int warmup = 0;
	int runs = 1;
};

namespace detail {

template<class F>
double elapsed_ms(F&& aOperation)
{
	const auto start = std::chrono::steady_clock::now();
	aOperation();
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
}

template<class F>
void run_samples(const run_options& aOptions, F&& aOperation)
{
	for (int i = 0; i < aOptions.warmup; ++i) {
		aOperation();
	}
	std::vector<double> samples;
	samples.reserve(aOptions.runs);
	for (int i = 0; i < aOptions.runs; ++i) {
		samples.push_back(elapsed_ms(aOperation));
	}
	std::sort(samples.begin(), samples.end());
	const auto middle = samples.size() / 2;
	const auto median = std::midpoint(samples[(samples.size() - 1) / 2], samples[middle]);
	std::cout << "compute_ms min=" << samples.front() << " median=" << median << " max=" << samples.back() << '\n';
}

template<class T>
inline constexpr bool device_element = abi_element<T> || std::same_as<T, fft_complex>;

struct stream_owner {
	void* stream = nullptr;
	~stream_owner()
	{
		if (stream) {
			iane_hip_stream_sync(stream);
			iane_hip_stream_destroy(stream);
		}
	}
};

template<bool tPrepare, class TDst, class TOperation, class... TSrc>
void run_on_impl(const run_options& aOptions, std::tuple<TSrc...>& aSources, TDst aDestination, TOperation&& aOperation)
{
	if (aOptions.warmup < 0 || aOptions.runs <= 0 || aOptions.device < 0) {
		throw std::invalid_argument("Invalid run options");
	}
	const auto size = [&] {
		if constexpr (sizeof...(TSrc) == 0) {
			return aDestination.dimensions();
		} else {
			return std::get<0>(aSources).dimensions();
		}
	}();
	constexpr int dimension = TDst::dimension;
	const bool use_hip = aOptions.backend == backend::hip;
	std::cout << "backend=" << (use_hip ? "hip" : "cpu");
	if (use_hip) {
		std::cout << " device=" << aOptions.device;
	}
	std::cout << " size=" << math::detail::axis(size, 0);
	for (int i = 1; i < dimension; ++i) {
		std::cout << 'x' << math::detail::axis(size, i);
	}
	std::cout << " warmup=" << aOptions.warmup << " runs=" << aOptions.runs << '\n';
	const auto run = [&](auto& aExecution, auto aInvoke) {
		if constexpr (tPrepare) {
			const auto start = std::chrono::steady_clock::now();
			auto operation = aInvoke();
			aExecution.wait();
			std::cout << "prepare_ms="
			          << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count() << '\n';
			run_samples(aOptions, [&] { operation().wait(); });
		} else {
			run_samples(aOptions, [&] { aInvoke().wait(); });
		}
	};
	if (!use_hip) {
		cpu_execution execution;
		std::apply([&](const auto&... aInput) {
			run(execution, [&]() -> decltype(auto) { return aOperation(aInput..., aDestination, execution); });
		}, aSources);
		return;
	}
	// Device images hold the element types the HIP kernels take (custom
	// element types such as pixel structs run on the host backend only).
	constexpr bool device_elements = (device_element<typename TSrc::value_type> && ...) &&
		device_element<typename TDst::value_type>;
	if constexpr (!device_elements) {
		throw std::invalid_argument("HIP execution supports u8, u16, i32, f32, f64 and fft_complex elements");
	} else {
	std::tuple<device_image<typename TSrc::value_type, TSrc::dimension>...> inputs;
	device_image<typename TDst::value_type, TDst::dimension> output;
	stream_owner queue;
	const auto setup = elapsed_ms([&] {
		check_status(iane_hip_set_device(aOptions.device));
		check_status(iane_hip_stream_create(&queue.stream));
		[&]<std::size_t... I>(std::index_sequence<I...>) {
			(std::get<I>(inputs).resize(std::get<I>(aSources).dimensions()), ...);
		}(std::index_sequence_for<TSrc...>{});
		output.resize(aDestination.dimensions());
	});
	hip_execution execution{queue.stream};
	std::cout << "setup_ms=" << setup;
	if constexpr (sizeof...(TSrc) > 0) {
		const auto upload = elapsed_ms([&] {
			[&]<std::size_t... I>(std::index_sequence<I...>) {
				(copy(std::get<I>(aSources), std::get<I>(inputs).view(), execution), ...);
			}(std::index_sequence_for<TSrc...>{});
			execution.wait();
		});
		std::cout << " upload_ms=" << upload;
	}
	std::cout << '\n';
	std::apply([&](const auto&... aInput) {
		run(execution, [&]() -> decltype(auto) { return aOperation(aInput.const_view()..., output.view(), execution); });
	}, inputs);
	const auto download = elapsed_ms([&] { copy(output.const_view(), aDestination, execution).wait(); });
	std::cout << "download_ms=" << download << '\n';
	}
}

} // namespace detail

/// run_on.hpp:196-201
template<class TDst, class TOperation, class... TSrc>
requires(concepts::HostView<TDst> && (concepts::HostView<TSrc> && ...))
void run_on(const run_options& aOptions, std::tuple<TSrc...> aSources, TDst aDestination, TOperation&& aOperation)
{
	detail::run_on_impl<false>(aOptions, aSources, aDestination, std::forward<TOperation>(aOperation));
}

/// run_on.hpp:215-221
template<class TDst, class TPrepare, class... TSrc>
requires(concepts::HostView<TDst> && (concepts::HostView<TSrc> && ...))
void run_prepared_on(const run_options& aOptions, std::tuple<TSrc...> aSources, TDst aDestination, TPrepare&& aPrepare)
{
	detail::run_on_impl<true>(aOptions, aSources, aDestination, std::forward<TPrepare>(aPrepare));
}

template<class TSrc, class TDst, class TOperation>
requires(concepts::HostView<TSrc> && concepts::HostView<TDst>)
void run_on(const run_options& aOptions, TSrc aSource, TDst aDestination, TOperation&& aOperation)
{
	run_on(aOptions, std::tuple{aSource}, aDestination, std::forward<TOperation>(aOperation));
}

} // namespace iane
