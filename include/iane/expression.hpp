// Traversal, neighbourhood locators, lazy views and same-backend copies.
//
// Interface of the reference's
//   locator/image_locator.hpp:14-62        image_locator, border_mode
//   algorithms/for_each.hpp:183-210        for_each, for_each_position, for_each_locator
//   algorithms/transform.hpp:64-102        transform, transform_position, transform_locator
//   view/procedural.hpp:93-150             generate_view, constant_view, map_view, zip_view
//   view/algebra.hpp:52-151                view arithmetic and math functions
//   algorithms/copy.hpp:28-47              copy of a readable view or expression
//
// Host views run any callable with the reference's semantics on the cpu_policy
// traversal of convolution.hpp (detail::visit_host). Device views have no
// device compiler behind them here, so their per-pixel functions are written
// with the pointwise vocabulary of this header (placeholders _1 .. _4, float
// constants, + - * /, min, max, comparisons, select, sqrt, exp, log, ...):
// such a function is an expression object that the host evaluates directly
// and that a device expression serialises into a register program for the
// library's pointwise kernel (iane_hip_pointwise). Device values are FP32.
#pragma once

#include <iane/run_on.hpp>

#include <cmath>
#include <cstring>
#include <tuple>
#include <vector>

namespace iane {

// ------------------------------------------------------------ locator -----
/**
 * Reads samples at offsets from a position of a view (image_locator.hpp:14-56):
 * an in-range coordinate is used as is; out of range, the zero border returns
 * value_type{} for the whole sample and the clamp border clamps that axis. The
 * view's extent is the boundary (a crop is an independent image).
 */
template<class TView>
class image_locator {
public:
	using value_type = typename TView::value_type;
	using coord_t = typename TView::coord_t;
	using offset_t = coord_t;
	static constexpr int dimension = TView::dimension;

	image_locator(TView aSource, coord_t aPosition, border_mode aBorder)
		: mSource(aSource), mCentre(aPosition), mBorder(aBorder) {}

	coord_t position() const { return mCentre; }

	value_type operator[](const offset_t& aOffset) const
	{
		const auto extent = mSource.dimensions();
		coord_t at{};
		for (int a = 0; a < dimension; ++a) {
			const long long c = static_cast<long long>(math::detail::axis(mCentre, a)) + math::detail::axis(aOffset, a);
			const int n = math::detail::axis(extent, a);
			if (c >= 0 && c < n) {
				math::detail::axis(at, a) = static_cast<int>(c);
			} else if (mBorder == border_mode::zero || n <= 0) {
				return value_type{};
			} else {
				math::detail::axis(at, a) = c < 0 ? 0 : n - 1;
			}
		}
		return mSource[at];
	}

private:
	TView mSource;
	coord_t mCentre;
	border_mode mBorder;
};

// ------------------------------------------------- pointwise functions ----
namespace pointwise {

/// Base of every pointwise function object (placeholders, constants, operations).
struct expression {};

template<class T>
concept function = std::derived_from<std::remove_cvref_t<T>, expression>;

namespace detail {

// The float semantics shared by the host evaluation and the device kernel.
inline float apply(int aCode, float x, float y, float z)
{
	switch (aCode) {
	case IANE_PW_ADD: return x + y;
	case IANE_PW_SUB: return x - y;
	case IANE_PW_MUL: return x * y;
	case IANE_PW_DIV: return x / y;
	case IANE_PW_MIN: return y < x ? y : x;
	case IANE_PW_MAX: return x < y ? y : x;
	case IANE_PW_NEG: return -x;
	case IANE_PW_ABS: return std::fabs(x);
	case IANE_PW_SQRT: return std::sqrt(x);
	case IANE_PW_EXP: return std::exp(x);
	case IANE_PW_LOG: return std::log(x);
	case IANE_PW_SIN: return std::sin(x);
	case IANE_PW_COS: return std::cos(x);
	case IANE_PW_LOG1P: return std::log1p(x);
	case IANE_PW_LOG2: return std::log2(x);
	case IANE_PW_LOG10: return std::log10(x);
	case IANE_PW_EXP2: return std::exp2(x);
	case IANE_PW_POW: return std::pow(x, y);
	case IANE_PW_LT: return x < y ? 1.0f : 0.0f;
	case IANE_PW_LE: return x <= y ? 1.0f : 0.0f;
	case IANE_PW_GT: return x > y ? 1.0f : 0.0f;
	case IANE_PW_GE: return x >= y ? 1.0f : 0.0f;
	case IANE_PW_EQ: return x == y ? 1.0f : 0.0f;
	case IANE_PW_NE: return x != y ? 1.0f : 0.0f;
	case IANE_PW_SELECT: return x != 0.0f ? y : z;
	default: return std::fma(x, y, z);
	}
}

} // namespace detail

/// Appends register operations to a program (at most IANE_PW_MAX_OPS of them).
class program_builder {
public:
	explicit program_builder(std::vector<iane_pw_op>& aOps) : mOps(aOps) {}
	int push(int aCode, int aA = 0, int aB = 0, int aC = 0, float aValue = 0.0f)
	{
		if (mOps.size() >= IANE_PW_MAX_OPS) {
			throw std::length_error("Pointwise expression exceeds 32 operations");
		}
		const int dst = static_cast<int>(mOps.size());
		mOps.push_back(iane_pw_op{aCode, dst, aA, aB, aC, aValue});
		return dst;
	}

private:
	std::vector<iane_pw_op>& mOps;
};

/// The I-th argument of the function (zero-based), converted to float.
template<int I>
struct placeholder : expression {
	static constexpr int arity = I + 1;
	template<class... V>
	float operator()(const V&... aValues) const
	{
		static_assert(I < static_cast<int>(sizeof...(V)), "placeholder beyond the number of operands");
		return static_cast<float>(std::get<I>(std::forward_as_tuple(aValues...)));
	}
	int emit(program_builder&, const int* aArgs) const { return aArgs[I]; }
};

/// A float constant; as a function it ignores its arguments.
struct constant : expression {
	static constexpr int arity = 0;
	float value;
	explicit constant(float aValue) : value(aValue) {}
	template<class... V>
	float operator()(const V&...) const { return value; }
	int emit(program_builder& aProgram, const int*) const { return aProgram.push(IANE_PW_CONST, 0, 0, 0, value); }
};

template<class T>
auto lift(const T& aValue)
{
	if constexpr (function<T>) {
		return aValue;
	} else {
		static_assert(std::is_arithmetic_v<T>, "pointwise operands are functions or arithmetic constants");
		return constant(static_cast<float>(aValue));
	}
}

template<class T>
using lifted = decltype(lift(std::declval<T>()));

template<int tCode, class A, class B = constant, class C = constant>
struct operation : expression {
	static constexpr int arity = std::max({A::arity, B::arity, C::arity});
	A a;
	B b;
	C c;
	operation(A aA, B aB = constant(0.0f), C aC = constant(0.0f)) : a(aA), b(aB), c(aC) {}

	template<class... V>
	float operator()(const V&... aValues) const
	{
		return detail::apply(tCode, a(aValues...), b(aValues...), c(aValues...));
	}
	int emit(program_builder& aProgram, const int* aArgs) const
	{
		const int ra = a.emit(aProgram, aArgs);
		constexpr bool unary = tCode >= IANE_PW_NEG && tCode < IANE_PW_POW;
		constexpr bool ternary = tCode == IANE_PW_SELECT || tCode == IANE_PW_FMA;
		const int rb = unary ? ra : b.emit(aProgram, aArgs);
		const int rc = ternary ? c.emit(aProgram, aArgs) : ra;
		return aProgram.push(tCode, ra, rb, rc);
	}
};

template<class A, class B>
concept operands = (function<A> || function<B>) && (function<A> || std::is_arithmetic_v<A>) &&
	(function<B> || std::is_arithmetic_v<B>);

#define IANE_POINTWISE_BINARY(tOp, tCode)                                                     \
	template<class A, class B>                                                                \
	requires operands<A, B>                                                                   \
	auto operator tOp(const A& aA, const B& aB)                                               \
	{                                                                                         \
		return operation<tCode, lifted<A>, lifted<B>>(lift(aA), lift(aB));                    \
	}
IANE_POINTWISE_BINARY(+, IANE_PW_ADD)
IANE_POINTWISE_BINARY(-, IANE_PW_SUB)
IANE_POINTWISE_BINARY(*, IANE_PW_MUL)
IANE_POINTWISE_BINARY(/, IANE_PW_DIV)
IANE_POINTWISE_BINARY(<, IANE_PW_LT)
IANE_POINTWISE_BINARY(<=, IANE_PW_LE)
IANE_POINTWISE_BINARY(>, IANE_PW_GT)
IANE_POINTWISE_BINARY(>=, IANE_PW_GE)
IANE_POINTWISE_BINARY(==, IANE_PW_EQ)
IANE_POINTWISE_BINARY(!=, IANE_PW_NE)
#undef IANE_POINTWISE_BINARY

#define IANE_POINTWISE_CALL2(tName, tCode)                                                    \
	template<class A, class B>                                                                \
	requires operands<A, B>                                                                   \
	auto tName(const A& aA, const B& aB)                                                      \
	{                                                                                         \
		return operation<tCode, lifted<A>, lifted<B>>(lift(aA), lift(aB));                    \
	}
IANE_POINTWISE_CALL2(min, IANE_PW_MIN)
IANE_POINTWISE_CALL2(max, IANE_PW_MAX)
IANE_POINTWISE_CALL2(pow, IANE_PW_POW)
#undef IANE_POINTWISE_CALL2

#define IANE_POINTWISE_CALL1(tName, tCode)                                                    \
	template<function A>                                                                      \
	auto tName(const A& aA)                                                                   \
	{                                                                                         \
		return operation<tCode, A>(aA);                                                       \
	}
IANE_POINTWISE_CALL1(abs, IANE_PW_ABS)
IANE_POINTWISE_CALL1(sqrt, IANE_PW_SQRT)
IANE_POINTWISE_CALL1(exp, IANE_PW_EXP)
IANE_POINTWISE_CALL1(log, IANE_PW_LOG)
IANE_POINTWISE_CALL1(sin, IANE_PW_SIN)
IANE_POINTWISE_CALL1(cos, IANE_PW_COS)
IANE_POINTWISE_CALL1(log1p, IANE_PW_LOG1P)
IANE_POINTWISE_CALL1(log2, IANE_PW_LOG2)
IANE_POINTWISE_CALL1(log10, IANE_PW_LOG10)
IANE_POINTWISE_CALL1(exp2, IANE_PW_EXP2)
#undef IANE_POINTWISE_CALL1

template<function A>
auto operator-(const A& aA)
{
	return operation<IANE_PW_NEG, A>(aA);
}

/// aCondition != 0 ? aThen : aElse (a C++ conditional on floats).
template<class C, class T, class E>
requires(function<C> || function<T> || function<E>)
auto select(const C& aCondition, const T& aThen, const E& aElse)
{
	return operation<IANE_PW_SELECT, lifted<C>, lifted<T>, lifted<E>>(lift(aCondition), lift(aThen), lift(aElse));
}

/// std::fma(a, b, c) in float.
template<class A, class B, class C>
requires(function<A> || function<B> || function<C>)
auto fma(const A& aA, const B& aB, const C& aC)
{
	return operation<IANE_PW_FMA, lifted<A>, lifted<B>, lifted<C>>(lift(aA), lift(aB), lift(aC));
}

/// A function returning `aValue` whatever its arguments.
inline constant value(float aValue) { return constant(aValue); }

inline constexpr placeholder<0> _1{};
inline constexpr placeholder<1> _2{};
inline constexpr placeholder<2> _3{};
inline constexpr placeholder<3> _4{};

} // namespace pointwise

// ----------------------------------------------------- device expression --
/**
 * A lazy device view: a register program over device memory views, all of
 * the same extent, evaluated by copy() (or transform) into device memory.
 * value_type is float. Built by constant_view, map_view, zip_view and the view
 * algebra on device views; not indexable on the host.
 */
template<int N>
class device_expression {
public:
	struct is_device_view_tag {};
	using value_type = float;
	using extents_t = typename math::dim_traits<N>::extents_t;
	using coord_t = typename math::dim_traits<N>::coord_t;
	static constexpr int dimension = N;

	explicit device_expression(extents_t aSize) : mSize(aSize) {}
	extents_t dimensions() const { return mSize; }

	/// A device memory view as an expression (an input load).
	template<concepts::MemoryView V>
	requires(concepts::DeviceView<V> && V::dimension == N)
	static device_expression load(const V& aView)
	{
		device_expression e(aView.dimensions());
		e.mInputs.push_back(detail::describe(aView));
		pointwise::program_builder program(e.mOps);
		e.mResult = program.push(IANE_PW_INPUT, 0);
		return e;
	}

	static device_expression fill(extents_t aSize, float aValue)
	{
		device_expression e(aSize);
		pointwise::program_builder program(e.mOps);
		e.mResult = program.push(IANE_PW_CONST, 0, 0, 0, aValue);
		return e;
	}

	/// The coordinate along aAxis of each element (as float): the position
	/// for_each_position hands its function, for device expressions.
	static device_expression coordinate(extents_t aSize, int aAxis)
	{
		if (aAxis < 0 || aAxis >= N) {
			throw std::invalid_argument("Coordinate axis out of range");
		}
		device_expression e(aSize);
		pointwise::program_builder program(e.mOps);
		e.mResult = program.push(IANE_PW_COORD, aAxis);
		return e;
	}

	/// Point sample of a rank-R device view (a texture view's texels) at the
	/// view-local coordinates the expressions give, one per axis of the view;
	/// the sampled view may have any extent.
	template<int R>
	static device_expression sample(const iane_view_desc& aSource, const device_expression* const (&aCoords)[R])
	{
		static_assert(R >= 1 && R <= 3, "point samples read views of rank 1..3");
		device_expression e(aCoords[0]->dimensions());
		int regs[3] = {0, 0, 0};
		for (int i = 0; i < R; ++i) {
			const auto size = aCoords[i]->dimensions();
			for (int k = 0; k < N; ++k) {
				if (math::detail::axis(size, k) != math::detail::axis(e.mSize, k)) {
					throw std::invalid_argument("Zip operands require matching nonnegative dimensions");
				}
			}
			regs[i] = e.append(*aCoords[i]);
		}
		const int input = e.add_input(aSource);
		pointwise::program_builder program(e.mOps);
		e.mResult = program.push(IANE_PW_SAMPLE, input, regs[0], R >= 2 ? regs[1] : regs[0],
			static_cast<float>(R >= 3 ? regs[2] : regs[0]));
		return e;
	}

	/// aFunction(operands[0], operands[1], ...) per pixel.
	template<pointwise::function F>
	static device_expression apply(extents_t aSize, const F& aFunction, std::initializer_list<const device_expression*> aOperands)
	{
		device_expression e(aSize);
		std::vector<int> args;
		for (const device_expression* operand : aOperands) {
			args.push_back(e.append(*operand));
		}
		if (static_cast<int>(args.size()) < F::arity) {
			throw std::invalid_argument("Pointwise function reads more operands than it was given");
		}
		pointwise::program_builder program(e.mOps);
		e.mResult = aFunction.emit(program, args.data());
		return e;
	}

	/// True when the expression is one constant (its value in *aValue).
	bool constant_value(float* aValue) const
	{
		if (mOps.size() == 1 && mOps[0].code == IANE_PW_CONST) {
			*aValue = mOps[0].value;
			return true;
		}
		return false;
	}

	/// Evaluate into a device memory view of the same extent.
	template<concepts::MemoryView TDst>
	requires(concepts::DeviceView<TDst> && TDst::dimension == N)
	hip_execution evaluate(TDst aDestination, hip_policy aPolicy) const
	{
		iane_pw_program program{};
		program.n_inputs = static_cast<int32_t>(mInputs.size());
		program.n_ops = static_cast<int32_t>(mOps.size());
		program.result = mResult;
		std::copy(mInputs.begin(), mInputs.end(), program.inputs);
		std::copy(mOps.begin(), mOps.end(), program.ops);
		const auto d = detail::describe(aDestination);
		detail::check_status(iane_hip_pointwise(&program, &d, aPolicy.stream));
		return hip_execution{aPolicy.stream};
	}

private:
	// The index of an input view (shared when an earlier input describes the
	// same view).
	int add_input(const iane_view_desc& aView)
	{
		for (std::size_t i = 0; i < mInputs.size(); ++i) {
			if (std::memcmp(&mInputs[i], &aView, sizeof(aView)) == 0) {
				return static_cast<int>(i);
			}
		}
		if (mInputs.size() >= IANE_PW_MAX_INPUTS) {
			throw std::length_error("Pointwise expression exceeds 8 input views");
		}
		mInputs.push_back(aView);
		return static_cast<int>(mInputs.size()) - 1;
	}

	// Copy another expression's inputs and operations in; returns its result
	// register here.
	int append(const device_expression& aOther)
	{
		std::vector<int> input_map;
		for (const auto& in : aOther.mInputs) {
			input_map.push_back(add_input(in));
		}
		const int base = static_cast<int>(mOps.size());
		pointwise::program_builder program(mOps);
		for (const auto& op : aOther.mOps) {
			switch (op.code) {
			case IANE_PW_INPUT: program.push(IANE_PW_INPUT, input_map[op.a]); break;
			case IANE_PW_COORD: program.push(IANE_PW_COORD, op.a); break;
			case IANE_PW_CONST: program.push(IANE_PW_CONST, 0, 0, 0, op.value); break;
			case IANE_PW_SAMPLE:
				program.push(IANE_PW_SAMPLE, input_map[op.a], op.b + base, op.c + base,
					static_cast<float>(static_cast<int>(op.value) + base));
				break;
			default: program.push(op.code, op.a + base, op.b + base, op.c + base, op.value); break;
			}
		}
		return aOther.mResult + base;
	}

	extents_t mSize;
	std::vector<iane_view_desc> mInputs;
	std::vector<iane_pw_op> mOps;
	int mResult = 0;
};

namespace detail {

template<class T>
inline constexpr bool is_device_expression = false;
template<int N>
inline constexpr bool is_device_expression<device_expression<N>> = true;

template<class V>
concept device_operand = concepts::DeviceView<V> && (concepts::MemoryView<V> || is_device_expression<V>);

template<class V>
auto as_expression(const V& aView)
{
	if constexpr (is_device_expression<V>) {
		return aView;
	} else {
		static_assert(concepts::MemoryView<V> && abi_element<typename V::value_type>,
			"device expressions read u8, u16, i32, f32 or f64 memory views");
		return device_expression<V::dimension>::load(aView);
	}
}

template<class A, class B>
void check_zip_extents(const A& aLeft, const B& aRight)
{
	const auto left = aLeft.dimensions();
	const auto right = aRight.dimensions();
	for (int i = 0; i < A::dimension; ++i) {
		if (math::detail::axis(left, i) < 0 || math::detail::axis(left, i) != math::detail::axis(right, i)) {
			throw std::invalid_argument("Zip operands require matching nonnegative dimensions");
		}
	}
}

template<class>
inline constexpr bool always_false = false;

template<class TValue>
struct constant_pixel {
	TValue value;
	template<class TCoordinate>
	TValue operator()(const TCoordinate&) const { return value; }
};

} // namespace detail

// ------------------------------------------------------ host procedural ---
/// A read-only host view computed from coordinates (procedural.hpp:29-48).
template<int N, class F>
class generated_view {
public:
	struct is_host_view_tag {};
	using coord_t = typename math::dim_traits<N>::coord_t;
	using extents_t = typename math::dim_traits<N>::extents_t;
	using value_type = std::remove_cvref_t<std::invoke_result_t<const F&, const coord_t&>>;
	static constexpr int dimension = N;

	generated_view(extents_t aSize, F aFunction) : mSize(aSize), mFunction(std::move(aFunction)) {}
	extents_t dimensions() const { return mSize; }
	value_type operator[](const coord_t& aPosition) const { return mFunction(aPosition); }

private:
	extents_t mSize;
	F mFunction;
};

/**
 * generate_view (procedural.hpp:93-106): values from a coordinate function over
 * the domain's extent. Host domains only (a device domain has no device
 * compiler for an arbitrary coordinate function).
 */
template<class TDomain, class TFunction>
requires(concepts::AnyView<TDomain>)
auto generate_view(const TDomain& aDomain, TFunction aFunction)
{
	static_assert(concepts::HostView<TDomain>, "generate_view over device views needs a device compiler: use constant_view, "
		"map_view, zip_view or the view algebra with iane::pointwise functions");
	return generated_view<TDomain::dimension, TFunction>(aDomain.dimensions(), std::move(aFunction));
}

/**
 * The coordinate along aAxis at every position of the domain, as float (no
 * reference counterpart: the device form of the position for_each_position
 * passes to its function, for expressions such as select(coordinate == 7, ...)).
 */
template<class TDomain>
requires(concepts::DeviceView<TDomain>)
auto coordinate_view(const TDomain& aDomain, int aAxis)
{
	return device_expression<TDomain::dimension>::coordinate(aDomain.dimensions(), aAxis);
}

/**
 * texture_image_view::sample (view/texture_image_view.hpp:56-70) for device
 * expressions: the point sample of aTexture at view-local coordinates measured
 * from the outer corner, one coordinate expression (or device view) per texture
 * axis, evaluated over their common domain. Per axis the coordinate is clamped
 * to the edge texel centres [0.5, extent - 0.5] and texel floor(coordinate) is
 * read; an empty view or a nonfinite coordinate gives 0.
 */
template<class T, int R, class... TCoord>
requires(sizeof...(TCoord) == R && (concepts::DeviceView<TCoord> && ...))
auto sample(const texture_image_view<T, R>& aTexture, const TCoord&... aCoords)
{
	static_assert(detail::abi_element<T>, "point samples read u8, u16, i32, f32 or f64 texels");
	const auto coords = std::make_tuple(detail::as_expression(aCoords)...);
	return std::apply([&](const auto&... aExpr) {
		using expression_t = std::remove_cvref_t<decltype(std::get<0>(coords))>;
		const expression_t* pointers[R] = {&aExpr...};
		return expression_t::template sample<R>(detail::describe(aTexture), pointers);
	}, coords);
}

/// constant_view (procedural.hpp:108-117): the same value at every coordinate.
template<class TDomain, class TValue>
requires(concepts::AnyView<TDomain>)
auto constant_view(const TDomain& aDomain, TValue aValue)
{
	if constexpr (concepts::DeviceView<TDomain>) {
		return device_expression<TDomain::dimension>::fill(aDomain.dimensions(), static_cast<float>(aValue));
	} else {
		return generated_view<TDomain::dimension, detail::constant_pixel<TValue>>(aDomain.dimensions(),
			detail::constant_pixel<TValue>{aValue});
	}
}

/// map_view (procedural.hpp:119-131): aFunction(source pixel), lazily.
template<class TSource, class TFunction>
requires(concepts::AnyView<TSource>)
auto map_view(TSource aSource, TFunction aFunction)
{
	if constexpr (concepts::DeviceView<TSource>) {
		static_assert(pointwise::function<TFunction>, "device map_view takes an iane::pointwise function (_1, ...)");
		const auto operand = detail::as_expression(aSource);
		return device_expression<TSource::dimension>::apply(aSource.dimensions(), aFunction, {&operand});
	} else {
		using coord_t = typename TSource::coord_t;
		return generate_view(aSource, [aSource, aFunction](const coord_t& aPosition) {
			return aFunction(aSource[aPosition]);
		});
	}
}

/// zip_view (procedural.hpp:133-150): aFunction(left pixel, right pixel), lazily.
template<class TLeft, class TRight, class TFunction>
requires(concepts::AnyView<TLeft> && concepts::AnyView<TRight> && TLeft::dimension == TRight::dimension &&
	concepts::DeviceView<TLeft> == concepts::DeviceView<TRight>)
auto zip_view(TLeft aLeft, TRight aRight, TFunction aFunction)
{
	detail::check_zip_extents(aLeft, aRight);
	if constexpr (concepts::DeviceView<TLeft>) {
		static_assert(pointwise::function<TFunction>, "device zip_view takes an iane::pointwise function (_1, _2, ...)");
		const auto left = detail::as_expression(aLeft);
		const auto right = detail::as_expression(aRight);
		return device_expression<TLeft::dimension>::apply(aLeft.dimensions(), aFunction, {&left, &right});
	} else {
		using coord_t = typename TLeft::coord_t;
		return generate_view(aLeft, [aLeft, aRight, aFunction](const coord_t& aPosition) {
			return aFunction(aLeft[aPosition], aRight[aPosition]);
		});
	}
}

// ------------------------------------------------------------ algebra -----
namespace detail {

template<class V>
concept algebra_view = concepts::AnyView<V> && !pointwise::function<V>;

template<class L, class R>
concept algebra_operands = (algebra_view<L> && algebra_view<R> && L::dimension == R::dimension &&
	concepts::DeviceView<L> == concepts::DeviceView<R>) || (algebra_view<L> && std::is_arithmetic_v<R>) ||
	(std::is_arithmetic_v<L> && algebra_view<R>);

// C++ operators on host pixels (the pixel type's own promotion rules).
template<int tCode>
struct host_pixel_operation {
	template<class A, class B>
	auto operator()(const A& a, const B& b) const
	{
		if constexpr (tCode == IANE_PW_ADD) {
			return a + b;
		} else if constexpr (tCode == IANE_PW_SUB) {
			return a - b;
		} else if constexpr (tCode == IANE_PW_MUL) {
			return a * b;
		} else if constexpr (tCode == IANE_PW_DIV) {
			return a / b;
		} else {
			return std::pow(a, b);
		}
	}
};

template<int tCode, class L, class R>
auto arithmetic_view(const L& aLeft, const R& aRight)
{
	if constexpr ((algebra_view<L> && concepts::DeviceView<L>) || (algebra_view<R> && concepts::DeviceView<R>)) {
		const auto f = pointwise::operation<tCode, pointwise::placeholder<0>, pointwise::placeholder<1>>(
			pointwise::placeholder<0>{}, pointwise::placeholder<1>{});
		if constexpr (!algebra_view<R>) {
			return zip_view(aLeft, constant_view(aLeft, aRight), f);
		} else if constexpr (!algebra_view<L>) {
			return zip_view(constant_view(aRight, aLeft), aRight, f);
		} else {
			return zip_view(aLeft, aRight, f);
		}
	} else {
		if constexpr (!algebra_view<R>) {
			return map_view(aLeft, [aRight](const auto& a) { return host_pixel_operation<tCode>{}(a, aRight); });
		} else if constexpr (!algebra_view<L>) {
			return map_view(aRight, [aLeft](const auto& b) { return host_pixel_operation<tCode>{}(aLeft, b); });
		} else {
			return zip_view(aLeft, aRight, host_pixel_operation<tCode>{});
		}
	}
}

template<int tCode, class V>
auto math_view(const V& aView)
{
	if constexpr (concepts::DeviceView<V>) {
		const auto f = pointwise::operation<tCode, pointwise::placeholder<0>>(pointwise::placeholder<0>{});
		return map_view(aView, f);
	} else {
		return map_view(aView, [](const auto& a) {
			if constexpr (tCode == IANE_PW_NEG) {
				return -a;
			} else if constexpr (tCode == IANE_PW_LOG) {
				return std::log(a);
			} else if constexpr (tCode == IANE_PW_LOG1P) {
				return std::log1p(a);
			} else if constexpr (tCode == IANE_PW_LOG2) {
				return std::log2(a);
			} else if constexpr (tCode == IANE_PW_LOG10) {
				return std::log10(a);
			} else if constexpr (tCode == IANE_PW_EXP) {
				return std::exp(a);
			} else if constexpr (tCode == IANE_PW_EXP2) {
				return std::exp2(a);
			} else if constexpr (tCode == IANE_PW_SQRT) {
				return std::sqrt(a);
			} else if constexpr (tCode == IANE_PW_SIN) {
				return std::sin(a);
			} else {
				return std::cos(a);
			}
		});
	}
}

} // namespace detail

/// Pointwise arithmetic of views, or of a view and an arithmetic scalar (algebra.hpp:64-117).
template<class L, class R>
requires detail::algebra_operands<L, R>
auto operator+(const L& aLeft, const R& aRight) { return detail::arithmetic_view<IANE_PW_ADD>(aLeft, aRight); }
template<class L, class R>
requires detail::algebra_operands<L, R>
auto operator-(const L& aLeft, const R& aRight) { return detail::arithmetic_view<IANE_PW_SUB>(aLeft, aRight); }
template<class L, class R>
requires detail::algebra_operands<L, R>
auto operator*(const L& aLeft, const R& aRight) { return detail::arithmetic_view<IANE_PW_MUL>(aLeft, aRight); }
template<class L, class R>
requires detail::algebra_operands<L, R>
auto operator/(const L& aLeft, const R& aRight) { return detail::arithmetic_view<IANE_PW_DIV>(aLeft, aRight); }
/// std::pow of corresponding pixels, or with a scalar (algebra.hpp:137-145).
template<class L, class R>
requires detail::algebra_operands<L, R>
auto pow(const L& aBase, const R& aExponent) { return detail::arithmetic_view<IANE_PW_POW>(aBase, aExponent); }
/// Unary negation (algebra.hpp:52-57).
template<detail::algebra_view V>
auto operator-(const V& aView) { return detail::math_view<IANE_PW_NEG>(aView); }

/// Pointwise math functions (algebra.hpp:119-131).
template<detail::algebra_view V> auto log(const V& aView) { return detail::math_view<IANE_PW_LOG>(aView); }
template<detail::algebra_view V> auto log1p(const V& aView) { return detail::math_view<IANE_PW_LOG1P>(aView); }
template<detail::algebra_view V> auto log2(const V& aView) { return detail::math_view<IANE_PW_LOG2>(aView); }
template<detail::algebra_view V> auto log10(const V& aView) { return detail::math_view<IANE_PW_LOG10>(aView); }
template<detail::algebra_view V> auto exp(const V& aView) { return detail::math_view<IANE_PW_EXP>(aView); }
template<detail::algebra_view V> auto exp2(const V& aView) { return detail::math_view<IANE_PW_EXP2>(aView); }
template<detail::algebra_view V> auto sqrt(const V& aView) { return detail::math_view<IANE_PW_SQRT>(aView); }
template<detail::algebra_view V> auto sin(const V& aView) { return detail::math_view<IANE_PW_SIN>(aView); }
template<detail::algebra_view V> auto cos(const V& aView) { return detail::math_view<IANE_PW_COS>(aView); }

// ----------------------------------------------------------- traversal ----
namespace detail {

template<class TSrc, class TDst>
auto matching_dimensions(const TSrc& aSource, const TDst& aDestination)
{
	const auto size = aDestination.dimensions();
	const auto input = aSource.dimensions();
	for (int i = 0; i < TDst::dimension; ++i) {
		if (math::detail::axis(size, i) < 0 || math::detail::axis(input, i) != math::detail::axis(size, i)) {
			throw std::invalid_argument("Source and destination require matching nonnegative dimensions");
		}
	}
	return size;
}

// Per-coordinate callables of the host traversal (each worker holds a copy).
template<class TView, class F>
struct visit_pixel {
	TView view;
	F function;
	void operator()(const typename TView::coord_t& aPosition) { function(view[aPosition], aPosition); }
};

template<class TView, class F>
struct visit_locator {
	TView view;
	F function;
	border_mode border;
	void operator()(const typename TView::coord_t& aPosition) { function(image_locator<TView>{view, aPosition, border}); }
};

template<int tKind, class TSrc, class TDst, class F>
struct assign_pixel {
	TSrc source;
	TDst destination;
	F function;
	border_mode border;
	void operator()(const typename TDst::coord_t& aPosition)
	{
		if constexpr (tKind == 0) {
			destination[aPosition] = function(source[aPosition]);
		} else if constexpr (tKind == 1) {
			destination[aPosition] = function(source[aPosition], aPosition);
		} else {
			destination[aPosition] = function(image_locator<TSrc>{source, aPosition, border});
		}
	}
};

template<class V>
void device_traversal_unavailable()
{
	static_assert(always_false<V>, "this traversal runs arbitrary callables, which need a device compiler: on device "
		"views use copy / transform / for_each with iane::pointwise functions, or convolve");
}

} // namespace detail

/**
 * for_each_position (for_each.hpp:183-197): aFunction(pixel, position) for
 * every coordinate of a host view, in storage order for one worker; the
 * cpu_policy partition otherwise (first exception rethrown after joining).
 */
template<class TView, class TFunction>
requires(concepts::AnyView<TView>)
auto for_each_position(TView aView, TFunction aFunction, view_policy<TView> aPolicy = {})
{
	if constexpr (concepts::DeviceView<TView>) {
		detail::device_traversal_unavailable<TView>();
		return hip_execution{aPolicy.stream};
	} else {
		detail::visit_host(aView, detail::visit_pixel<TView, TFunction>{aView, aFunction}, aPolicy);
		return cpu_execution{};
	}
}

/**
 * for_each (for_each.hpp:199-204): aFunction(pixel) for every pixel. On a
 * device memory view the function is an iane::pointwise function of the pixel
 * value and the pixel is replaced by its result.
 */
template<class TView, class TFunction>
requires(concepts::AnyView<TView>)
auto for_each(TView aView, TFunction aFunction, view_policy<TView> aPolicy = {})
{
	if constexpr (concepts::DeviceView<TView>) {
		static_assert(pointwise::function<TFunction>, "device for_each takes an iane::pointwise function of the pixel");
		return copy(map_view(aView, aFunction), aView, aPolicy);
	} else {
		return for_each_position(aView, [aFunction](auto&& aPixel, const auto&) mutable {
			aFunction(static_cast<decltype(aPixel)&&>(aPixel));
		}, aPolicy);
	}
}

/// for_each_locator (for_each.hpp:206-210): aFunction(locator) for every position (host views).
template<class TView, class TFunction>
requires(concepts::AnyView<TView>)
auto for_each_locator(TView aView, TFunction aFunction, border_mode aBorder = border_mode::clamp,
	view_policy<TView> aPolicy = {})
{
	if constexpr (concepts::DeviceView<TView>) {
		detail::device_traversal_unavailable<TView>();
		return hip_execution{aPolicy.stream};
	} else {
		detail::visit_host(aView, detail::visit_locator<TView, TFunction>{aView, aFunction, aBorder}, aPolicy);
		return cpu_execution{};
	}
}

/**
 * transform (transform.hpp:64-77): dst[p] = aFunction(src[p]). Device views take
 * an iane::pointwise function and run as one pointwise launch on the stream.
 * @throws std::invalid_argument for mismatched extents.
 */
template<class TSrc, class TDst, class TFunction>
requires(concepts::AnyView<TSrc> && concepts::AnyView<TDst> && TSrc::dimension == TDst::dimension &&
	concepts::DeviceView<TSrc> == concepts::DeviceView<TDst>)
auto transform(TSrc aSource, TDst aDestination, TFunction aFunction, view_policy<TDst> aPolicy = {})
{
	detail::matching_dimensions(aSource, aDestination);
	if constexpr (concepts::DeviceView<TDst>) {
		static_assert(pointwise::function<TFunction>, "device transform takes an iane::pointwise function (_1, ...)");
		return copy(map_view(aSource, aFunction), aDestination, aPolicy);
	} else {
		detail::visit_host(aDestination, detail::assign_pixel<0, TSrc, TDst, TFunction>{aSource, aDestination, aFunction,
			border_mode::clamp}, aPolicy);
		return cpu_execution{};
	}
}

/// transform_position (transform.hpp:79-86): dst[p] = aFunction(src[p], p) (host views).
template<class TSrc, class TDst, class TFunction>
requires(concepts::AnyView<TSrc> && concepts::AnyView<TDst> && TSrc::dimension == TDst::dimension)
auto transform_position(TSrc aSource, TDst aDestination, TFunction aFunction, view_policy<TDst> aPolicy = {})
{
	detail::matching_dimensions(aSource, aDestination);
	if constexpr (concepts::DeviceView<TDst>) {
		detail::device_traversal_unavailable<TDst>();
		return hip_execution{aPolicy.stream};
	} else {
		detail::visit_host(aDestination, detail::assign_pixel<1, TSrc, TDst, TFunction>{aSource, aDestination, aFunction,
			border_mode::clamp}, aPolicy);
		return cpu_execution{};
	}
}

/**
 * transform_locator (transform.hpp:88-102): dst[p] = aFunction(image_locator{src, p,
 * border}) for host views. (On device views the neighbourhood operation this
 * path provides is convolve.)
 */
template<class TSrc, class TDst, class TFunction>
requires(concepts::AnyView<TSrc> && concepts::AnyView<TDst> && TSrc::dimension == TDst::dimension)
auto transform_locator(TSrc aSource, TDst aDestination, TFunction aFunction, border_mode aBorder = border_mode::clamp,
	view_policy<TDst> aPolicy = {})
{
	detail::matching_dimensions(aSource, aDestination);
	if constexpr (concepts::DeviceView<TDst>) {
		detail::device_traversal_unavailable<TDst>();
		return hip_execution{aPolicy.stream};
	} else {
		detail::visit_host(aDestination, detail::assign_pixel<2, TSrc, TDst, TFunction>{aSource, aDestination, aFunction,
			aBorder}, aPolicy);
		return cpu_execution{};
	}
}

// --------------------------------------------------------------- copy -----
/// copy (copy.hpp:28-33): materialise a readable host view or expression into a host view.
template<class TSrc, class TDst>
requires(concepts::HostView<TSrc> && concepts::HostView<TDst> && TSrc::dimension == TDst::dimension)
cpu_execution copy(TSrc aSource, TDst aDestination, cpu_policy aPolicy = {})
{
	return transform(aSource, aDestination, [](const auto& aValue) { return aValue; }, aPolicy);
}

/// copy (copy.hpp:35-47): evaluate a device expression into a device memory view on the stream.
template<int N, class TDst>
requires(concepts::DeviceView<TDst> && concepts::MemoryView<TDst> && TDst::dimension == N)
hip_execution copy(const device_expression<N>& aSource, TDst aDestination, hip_policy aPolicy = {})
{
	detail::matching_dimensions(aSource, aDestination);
	return aSource.evaluate(aDestination, aPolicy);
}

/**
 * convolve with a device constant_view source (the sensitivity step of the
 * reference's RL loop, examples/rl_deconvolution/main.cu:62-63): the constant
 * image is not materialised. Other device expressions must be copied into an
 * image first.
 */
template<int N, class TKernel, class TDst>
requires(concepts::DeviceView<TKernel> && concepts::DeviceView<TDst> && TKernel::dimension == N &&
	TDst::dimension == N)
hip_execution convolve(const device_expression<N>& aSource, TKernel aKernel, TDst aDestination,
	convolution_options aOptions = {}, hip_policy aPolicy = {})
{
	detail::check_convolution(aSource, aKernel, aDestination);
	float value = 0.0f;
	if (!aSource.constant_value(&value)) {
		throw std::invalid_argument("Device convolution reads memory views or a constant_view: copy the expression into "
			"an image first");
	}
	const auto k = detail::describe(aKernel);
	const auto d = detail::describe(aDestination);
	detail::check_status(iane_hip_convolve_constant(static_cast<double>(value), &k, &d,
		static_cast<int>(aOptions.border), static_cast<int>(aOptions.precision), aPolicy.stream));
	return hip_execution{aPolicy.stream};
}

} // namespace iane
