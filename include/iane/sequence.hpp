// Scan and reduce prepared algorithms: reduce / reduce_along / inclusive_scan /
// exclusive_scan / inclusive_scan_along / exclusive_scan_along of the reference
// (include/iane/algorithms/{reduce.hpp:7-36, scan.hpp:7-46}, built on
// algorithms/detail/sequence.hpp:255-301), used through execution_context:
//
//   auto context = iane::make_execution_context(iane::hip_policy{stream});
//   auto sum = context.prepare(iane::reduce(std::plus<double>{}, 0.0), src, total);
//   context.run(sum, src, total).wait();
//
// The initial value sets the accumulator and destination element type; it is
// applied once per line and is the result of an empty reduction. Shapes: scans
// keep the source extents; a whole reduction (or one of a rank-1 view) writes a
// rank-1 view of extent 1; reduce_along<A> removes axis A.
//
// Device views run on libiane_hip.so, whose C ABI carries an operator code
// rather than a functor: std::plus, std::multiplies (typed or transparent),
// iane::minimum and iane::maximum (std::min / std::max semantics) are accepted.
// Integer sums and products wrap. Host views run any pure associative functor
// with the reference's sequential fold (serial; the reference's threaded CPU
// partition is grouping-only).
#pragma once

#include <iane/prepared.hpp>

#include <functional>

namespace iane {

/// std::min as an operator for scans and reductions.
struct minimum {
	static constexpr int code = IANE_OP_MIN;
	template<class T>
	constexpr T operator()(const T& a, const T& b) const { return b < a ? b : a; }
};

/// std::max as an operator for scans and reductions.
struct maximum {
	static constexpr int code = IANE_OP_MAX;
	template<class T>
	constexpr T operator()(const T& a, const T& b) const { return a < b ? b : a; }
};

namespace detail {

template<class T, template<class...> class Template>
inline constexpr bool specialization_of = false;
template<template<class...> class Template, class... Args>
inline constexpr bool specialization_of<Template<Args...>, Template> = true;

template<class TOp>
constexpr int sequence_op_code()
{
	if constexpr (requires { TOp::code; }) {
		return TOp::code;
	} else if constexpr (specialization_of<TOp, std::plus>) {
		return IANE_OP_SUM;
	} else if constexpr (specialization_of<TOp, std::multiplies>) {
		return IANE_OP_PRODUCT;
	} else {
		return -1;
	}
}

enum class sequence_kind { inclusive = IANE_SEQ_INCLUSIVE, exclusive = IANE_SEQ_EXCLUSIVE, reduction = IANE_SEQ_REDUCE };

struct sequence_plan_deleter {
	void operator()(iane_sequence_plan* aPlan) const noexcept { iane_hip_sequence_plan_destroy(aPlan); }
};

// Device state: a plan for one shape (sequence.hpp:195-248).
template<class TValue>
struct device_sequence_state {
	std::shared_ptr<iane_sequence_plan> plan;
	std::size_t workspace{};

	std::size_t scratch_bytes() const { return workspace; }

	template<class TSrc, class TDst>
	hip_execution run(TSrc aSource, TDst aDestination, scratch_space<hip_policy>& aScratch, hip_policy aPolicy)
	{
		const auto s = describe(aSource);
		const auto d = describe(aDestination);
		void* ws = workspace ? aScratch.bytes(workspace, alignof(double)) : nullptr;
		check_status(iane_hip_sequence_run(plan.get(), &s, &d, ws, aPolicy.stream));
		return hip_execution{aPolicy.stream};
	}
};

// Host state: the reference's sequential fold per line.
template<int tAxis, sequence_kind tKind, class TOp, class TValue>
struct host_sequence_state {
	TOp op;
	TValue init;

	std::size_t scratch_bytes() const { return 0; }

	template<class TSrc, class TDst>
	cpu_execution run(TSrc aSource, TDst aDestination, scratch_space<cpu_policy>&, cpu_policy)
	{
		constexpr int N = TSrc::dimension;
		const auto extent = aSource.dimensions();
		const auto count = math::detail::element_count<N>(extent);
		using coord_t = typename TSrc::coord_t;
		if constexpr (tAxis < 0) {
			TValue total = init;
			for (std::size_t i = 0; i < count; ++i) {
				const auto p = math::detail::coordinate_at<N, coord_t>(i, extent);
				if constexpr (tKind == sequence_kind::exclusive) {
					aDestination[p] = total;
				}
				total = op(total, static_cast<TValue>(aSource[p]));
				if constexpr (tKind == sequence_kind::inclusive) {
					aDestination[p] = total;
				}
			}
			if constexpr (tKind == sequence_kind::reduction) {
				aDestination[typename TDst::coord_t{}] = total;
			}
		} else {
			const auto length = static_cast<std::size_t>(math::detail::axis(extent, tAxis));
			auto line_extent = extent;
			math::detail::axis(line_extent, tAxis) = 1;
			const auto lines = math::detail::element_count<N>(line_extent);
			for (std::size_t line = 0; line < lines; ++line) {
				auto p = math::detail::coordinate_at<N, coord_t>(line, line_extent);
				TValue total = init;
				for (std::size_t i = 0; i < length; ++i) {
					math::detail::axis(p, tAxis) = static_cast<int>(i);
					if constexpr (tKind == sequence_kind::exclusive) {
						aDestination[p] = total;
					}
					total = op(total, static_cast<TValue>(aSource[p]));
					if constexpr (tKind == sequence_kind::inclusive) {
						aDestination[p] = total;
					}
				}
				if constexpr (tKind == sequence_kind::reduction) {
					typename TDst::coord_t q{};
					if constexpr (N > 1) {
						for (int a = 0, b = 0; a < N; ++a) {
							if (a != tAxis) {
								math::detail::axis(q, b++) = math::detail::axis(p, a);
							}
						}
					}
					aDestination[q] = total;
				}
			}
		}
		return {};
	}
};

template<int tAxis, sequence_kind tKind, class TOp, class TValue>
struct sequence_algorithm {
	TOp op;
	TValue init;

	template<class TSrc, class TDst, class TPolicy>
	requires(concepts::AnyView<TSrc> && concepts::AnyView<TDst> && (tAxis < TSrc::dimension) &&
		std::same_as<typename TDst::value_type, TValue> &&
		(TDst::dimension == (tKind != sequence_kind::reduction ? TSrc::dimension :
			(tAxis < 0 || TSrc::dimension == 1 ? 1 : TSrc::dimension - 1))))
	auto prepare(TSrc aSource, TDst aDestination, TPolicy) const
	{
		// destination extents (sequence.hpp:276-293)
		const auto extent = aSource.dimensions();
		const auto out = aDestination.dimensions();
		for (int i = 0; i < TDst::dimension; ++i) {
			int expected = 1;
			if constexpr (tKind != sequence_kind::reduction) {
				expected = math::detail::axis(extent, i);
			} else if constexpr (!(tAxis < 0 || TSrc::dimension == 1)) {
				expected = math::detail::axis(extent, i + (i >= tAxis));
			}
			if (math::detail::axis(out, i) != expected) {
				throw std::invalid_argument("Sequence views must match the prepared shape");
			}
		}
		if constexpr (concepts::DeviceView<TDst>) {
			static_assert(concepts::DeviceView<TSrc>, "device scans and reductions need device views");
			constexpr int code = sequence_op_code<TOp>();
			static_assert(code >= 0, "device scans and reductions take std::plus, std::multiplies, "
				"iane::minimum or iane::maximum");
			static_assert(abi_element<TValue> && abi_element<typename TSrc::value_type>,
				"device scans and reductions support u8, u16, i32, f32 and f64 elements");
			const auto s = describe(aSource);
			const auto d = describe(aDestination);
			iane_sequence_plan* raw = nullptr;
			check_status(iane_hip_sequence_plan_create(&raw, static_cast<int>(tKind), tAxis, code,
				element_code<TValue>(), &init, &s, &d));
			device_sequence_state<TValue> state{std::shared_ptr<iane_sequence_plan>(raw, sequence_plan_deleter{}), 0};
			check_status(iane_hip_sequence_plan_workspace_bytes(raw, &state.workspace));
			return state;
		} else {
			return host_sequence_state<tAxis, tKind, TOp, TValue>{op, init};
		}
	}
};

} // namespace detail

/// A reduction of the whole source into a rank-1 view of extent 1 (reduce.hpp:7-22).
template<class TOp, class TValue>
auto reduce(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<-1, detail::sequence_kind::reduction, TOp, TValue>{aOp, aInit};
}

/// Independent reductions along axis tAxis; the output drops that axis (reduce.hpp:24-36).
template<int tAxis, class TOp, class TValue>
requires(tAxis >= 0)
auto reduce_along(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<tAxis, detail::sequence_kind::reduction, TOp, TValue>{aOp, aInit};
}

/// Inclusive prefix scan of the whole source in storage order (scan.hpp:7-24).
template<class TOp, class TValue>
auto inclusive_scan(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<-1, detail::sequence_kind::inclusive, TOp, TValue>{aOp, aInit};
}

/// Exclusive prefix scan of the whole source (scan.hpp:26-31).
template<class TOp, class TValue>
auto exclusive_scan(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<-1, detail::sequence_kind::exclusive, TOp, TValue>{aOp, aInit};
}

/// Independent inclusive scans along axis tAxis (scan.hpp:33-38).
template<int tAxis, class TOp, class TValue>
requires(tAxis >= 0)
auto inclusive_scan_along(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<tAxis, detail::sequence_kind::inclusive, TOp, TValue>{aOp, aInit};
}

/// Independent exclusive scans along axis tAxis (scan.hpp:40-45).
template<int tAxis, class TOp, class TValue>
requires(tAxis >= 0)
auto exclusive_scan_along(TOp aOp, TValue aInit)
{
	return detail::sequence_algorithm<tAxis, detail::sequence_kind::exclusive, TOp, TValue>{aOp, aInit};
}

} // namespace iane
