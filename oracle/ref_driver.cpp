// TEST INFRASTRUCTURE — C entry points over the reference's own header-only
// CPU path, compiled from the headers where they lie (-I /root/reference/include)
// into oracle/_ref/libiane_ref.so by oracle/Makefile. Used only to write golden
// fixtures, to cross-check the C oracle, and as the timed CPU baseline.
//
//   ref_convolve  : iane::convolve (include/iane/algorithms/convolution.hpp:97-117)
//                   with cpu_policy{threads, grain} (include/iane/execution/context.hpp:20-23)
//   ref_rl        : the Richardson–Lucy loop of examples/rl_deconvolution/main.cu:51-81
//                   driven through the reference's convolve / zip_view / copy on float
//                   images (the example itself needs Boost and cannot be built here).
#include <iane/algorithms/convolution.hpp>
#include <iane/algorithms/copy.hpp>
#include <iane/algorithms/transform.hpp>
#include <iane/image/image.hpp>
#include <iane/resource/host/resource_host.hpp>
#include <iane/view/procedural.hpp>

#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>

namespace {

enum { T_U8 = 0, T_U16 = 1, T_I32 = 2, T_F32 = 3, T_F64 = 4 };

struct ref_view {
	void* ptr;
	int rank;
	int dtype;
	int extent[4];
	int stride[4]; // bytes
};

thread_local std::string g_error;

template<class T, int N>
auto make(const ref_view& v)
{
	using view_t = iane::host_image_view<T, N>;
	typename view_t::extents_t size{};
	typename view_t::extents_t strides{};
	for (int i = 0; i < N; ++i) {
		iane::math::detail::axis(size, i) = v.extent[i];
		iane::math::detail::axis(strides, i) = v.stride[i];
	}
	return view_t(static_cast<T*>(v.ptr), size, strides);
}

template<int N, class S, class K, class D>
void run(const ref_view& s, const ref_view& k, const ref_view& d, int border, size_t threads, size_t grain)
{
	const auto src = make<S, N>(s);
	const auto ker = make<K, N>(k);
	const auto dst = make<D, N>(d);
	iane::convolve(src, ker, dst,
		iane::convolution_options{border == 0 ? iane::border_mode::clamp : iane::border_mode::zero},
		iane::cpu_policy{threads, grain}).wait();
}

template<int N, class S, class K>
void by_dst(const ref_view& s, const ref_view& k, const ref_view& d, int b, size_t t, size_t g)
{
	switch (d.dtype) {
	case T_U8: run<N, S, K, std::uint8_t>(s, k, d, b, t, g); break;
	case T_U16: run<N, S, K, std::uint16_t>(s, k, d, b, t, g); break;
	case T_I32: run<N, S, K, std::int32_t>(s, k, d, b, t, g); break;
	case T_F32: run<N, S, K, float>(s, k, d, b, t, g); break;
	case T_F64: run<N, S, K, double>(s, k, d, b, t, g); break;
	default: throw std::invalid_argument("dst type");
	}
}

template<int N, class S>
void by_ker(const ref_view& s, const ref_view& k, const ref_view& d, int b, size_t t, size_t g)
{
	switch (k.dtype) {
	case T_I32: by_dst<N, S, std::int32_t>(s, k, d, b, t, g); break;
	case T_F32: by_dst<N, S, float>(s, k, d, b, t, g); break;
	case T_F64: by_dst<N, S, double>(s, k, d, b, t, g); break;
	default: throw std::invalid_argument("kernel type (i32, f32, f64)");
	}
}

template<int N>
void by_src(const ref_view& s, const ref_view& k, const ref_view& d, int b, size_t t, size_t g)
{
	switch (s.dtype) {
	case T_U8: by_ker<N, std::uint8_t>(s, k, d, b, t, g); break;
	case T_U16: by_ker<N, std::uint16_t>(s, k, d, b, t, g); break;
	case T_I32: by_ker<N, std::int32_t>(s, k, d, b, t, g); break;
	case T_F32: by_ker<N, float>(s, k, d, b, t, g); break;
	case T_F64: by_ker<N, double>(s, k, d, b, t, g); break;
	default: throw std::invalid_argument("src type");
	}
}

// RL functors with the semantics of examples/rl_deconvolution/main.cu:11-49, float observed.
struct initial_estimate {
	float epsilon;
	float operator()(float aObserved) const { return aObserved > epsilon ? aObserved : epsilon; }
};
struct observation_ratio {
	float epsilon;
	float operator()(float aObserved, float aPredicted) const
	{
		return aObserved / (aPredicted > epsilon ? aPredicted : epsilon);
	}
};
struct edge_correction {
	float epsilon;
	float operator()(float aBackprojection, float aSensitivity) const
	{
		return aSensitivity > 0.0f ? aBackprojection / (aSensitivity > epsilon ? aSensitivity : epsilon) : 0.0f;
	}
};
struct multiply_estimate {
	float operator()(float aEstimate, float aCorrection) const { return aEstimate * aCorrection; }
};

template<int N>
void rl(const ref_view& obs, const ref_view& psf, const ref_view& flipped, const ref_view& out, int iterations,
	float epsilon, size_t threads)
{
	const auto observed = make<float, N>(obs);
	const auto kernel = make<float, N>(psf);
	const auto adjoint = make<float, N>(flipped);
	const auto estimate = make<float, N>(out);
	const auto size = observed.dimensions();
	iane::host_image<float, N> predicted(size);
	iane::host_image<float, N> correction(size);
	iane::host_image<float, N> sensitivity(size);
	const iane::cpu_policy policy{threads, 4096};
	constexpr iane::convolution_options options{.border = iane::border_mode::zero};
	iane::convolve(iane::constant_view(sensitivity.view(), 1.0f), adjoint, sensitivity.view(), options, policy);
	iane::transform(observed, estimate, initial_estimate{epsilon}, policy);
	for (int i = 0; i < iterations; ++i) {
		iane::convolve(estimate, kernel, predicted.view(), options, policy);
		iane::copy(iane::zip_view(observed, predicted.const_view(), observation_ratio{epsilon}), predicted.view(),
			policy);
		iane::convolve(predicted.const_view(), adjoint, correction.view(), options, policy);
		const auto factor = iane::zip_view(correction.const_view(), sensitivity.const_view(), edge_correction{epsilon});
		iane::copy(iane::zip_view(estimate, factor, multiply_estimate{}), estimate, policy);
	}
}

template<class F>
int guarded(F&& f)
{
	try {
		f();
		return 0;
	} catch (const std::invalid_argument& e) {
		g_error = e.what();
		return 1;
	} catch (const std::exception& e) {
		g_error = e.what();
		return 3;
	}
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

int ref_convolve(const ref_view* s, const ref_view* k, const ref_view* d, int border, size_t threads, size_t grain)
{
	return guarded([&] {
		switch (s->rank) {
		case 1: by_src<1>(*s, *k, *d, border, threads, grain); break;
		case 2: by_src<2>(*s, *k, *d, border, threads, grain); break;
		case 3: by_src<3>(*s, *k, *d, border, threads, grain); break;
		case 4: by_src<4>(*s, *k, *d, border, threads, grain); break;
		default: throw std::invalid_argument("rank");
		}
	});
}

int ref_rl(const ref_view* obs, const ref_view* psf, const ref_view* flipped, const ref_view* out, int iterations,
	float epsilon, size_t threads)
{
	return guarded([&] {
		switch (obs->rank) {
		case 2: rl<2>(*obs, *psf, *flipped, *out, iterations, epsilon, threads); break;
		case 3: rl<3>(*obs, *psf, *flipped, *out, iterations, epsilon, threads); break;
		default: throw std::invalid_argument("rank");
		}
	});
}

} // extern "C"
