// C++ drop-in tests: the reference's own test cases restated against
// include/iane (host views on the CPU, device views through libiane_hip.so).
//
//   test_api host     -> host path only (no GPU needed)
//   test_api device   -> device path; exits 77 when no GPU is visible
//                        (test/unit/algorithms/cuda_test_main.hpp:12-23)
//
// Cases: test/unit/algorithms/convolution_cases.hpp:12-34 (3^N pattern, kernel
// [2,-1,0] on the last axis, zero border, N = 2, 3, 4), test_convolution.cpp:10-50
// (scalar coordinates; flip + crop, both borders), test_convolution_cuda.cu:53-86
// (device equals host exactly), test_for_each.cpp:43-56 (policy validation),
// prepared_cases.hpp:73-123 and real_fft_cases.hpp:12-77 (FFT), and the RL loop of
// examples/rl_deconvolution/main.cu:51-81 (device DIRECT/reference == host loop, bitwise),
// the texture-image cases of test_convolution_cuda.cu:10-163, and the scan /
// reduce case of test/unit/prepared/sequence_cases.hpp:31-117.
#include <iane/iane.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <numbers>
#include <numeric>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(...)                                                                             \
	do {                                                                                       \
		++g_checks;                                                                            \
		if (!(__VA_ARGS__)) {                                                                  \
			++g_failures;                                                                      \
			std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #__VA_ARGS__); \
		}                                                                                      \
	} while (0)

template<class E, class F>
void check_throws(F&& f, const char* what)
{
	++g_checks;
	try {
		f();
	} catch (const E&) {
		return;
	} catch (...) {
	}
	++g_failures;
	std::fprintf(stderr, "expected exception not thrown: %s\n", what);
}

// algorithm_test_views.hpp:12-25
struct pixel {
	int first{};
	int second{};
	pixel operator+(pixel o) const { return {first + o.first, second + o.second}; }
	pixel operator*(int f) const { return {first * f, second * f}; }
};

template<int N>
auto cube(int v)
{
	return typename iane::math::dim_traits<N>::extents_t(v, std::true_type{});
}

// convolution_cases.hpp:12-34, host path, custom element type
template<int N>
void pattern_host()
{
	iane::host_image<pixel, N> input(cube<N>(3));
	int count = 1;
	for (int i = 0; i < N; ++i) {
		count *= 3;
	}
	for (int i = 0; i < count; ++i) {
		input.pointer()[i] = {i + 1, -i - 1};
	}
	auto kshape = cube<N>(1);
	kshape[N - 1] = 3;
	iane::host_image<int, N> kernel(kshape);
	kernel.pointer()[0] = 2;
	kernel.pointer()[1] = -1;
	kernel.pointer()[2] = 0;
	iane::host_image<pixel, N> output(input.dimensions());
	iane::run_on(iane::run_options{}, std::tuple{input.const_view(), kernel.const_view()}, output.view(),
		[](auto s, auto k, auto d, auto& e) { return iane::convolve(s, k, d, {iane::border_mode::zero}, e); });
	const int plane = count / 3;
	for (int index = 0; index < count; ++index) {
		const int value = index + 1;
		const int next = index / plane == 2 ? 0 : value + plane;
		CHECK(output.pointer()[index].first == 2 * next - value);
		CHECK(output.pointer()[index].second == value - 2 * next);
	}
}

// the same case on the device, integer components (device element types are arithmetic)
template<int N>
void pattern_device()
{
	iane::host_image<int, N> input(cube<N>(3));
	int count = 1;
	for (int i = 0; i < N; ++i) {
		count *= 3;
	}
	for (int i = 0; i < count; ++i) {
		input.pointer()[i] = i + 1;
	}
	auto kshape = cube<N>(1);
	kshape[N - 1] = 3;
	iane::host_image<int, N> kernel(kshape);
	kernel.pointer()[0] = 2;
	kernel.pointer()[1] = -1;
	kernel.pointer()[2] = 0;
	iane::host_image<int, N> output(input.dimensions());
	iane::run_on(iane::run_options{.backend = iane::backend::hip}, std::tuple{input.const_view(), kernel.const_view()},
		output.view(), [](auto s, auto k, auto d, auto& e) { return iane::convolve(s, k, d, {iane::border_mode::zero}, e); });
	const int plane = count / 3;
	for (int index = 0; index < count; ++index) {
		const int value = index + 1;
		const int next = index / plane == 2 ? 0 : value + plane;
		CHECK(output.pointer()[index] == 2 * next - value);
	}
}

void scalar_coordinates()
{
	iane::host_image<int, 1> source(3), kernel(3), output(3);
	for (int i = 0; i < 3; ++i) {
		source.pointer()[i] = i + 1;
	}
	kernel.pointer()[0] = 1;
	kernel.pointer()[1] = 0;
	kernel.pointer()[2] = 0;
	iane::convolve(source.const_view(), kernel.const_view(), output.view(), {iane::border_mode::zero});
	CHECK(output.pointer()[0] == 2);
	CHECK(output.pointer()[1] == 3);
	CHECK(output.pointer()[2] == 0);
}

// test_convolution.cpp:25-50 on host and device views
void crop_flip(bool device)
{
	using coord_t = iane::math::vec2i;
	iane::host_image<std::uint8_t, 2> source(4, 2), result(2, 2);
	iane::host_image<float, 2> kernel(3, 1);
	for (int y = 0; y < 2; ++y) {
		for (int x = 0; x < 4; ++x) {
			source.view()[coord_t{x, y}] = static_cast<std::uint8_t>(10 * x + 100 * y);
		}
	}
	kernel.pointer()[0] = 1.0f;
	kernel.pointer()[1] = 0.0f;
	kernel.pointer()[2] = 0.0f;
	const std::uint8_t clamp[] = {20, 20, 120, 120};
	const std::uint8_t zero[] = {20, 0, 120, 0};
	for (const auto border : {iane::border_mode::clamp, iane::border_mode::zero}) {
		if (device) {
			iane::device_image<std::uint8_t, 2> dsrc(source.dimensions()), dres(result.dimensions());
			iane::device_image<float, 2> dker(kernel.dimensions());
			iane::copy(source.const_view(), dsrc.view()).wait();
			iane::copy(kernel.const_view(), dker.view()).wait();
			iane::convolve(dsrc.const_view().subview(coord_t{1, 0}, coord_t{2, 2}), dker.const_view(), dres.view(),
				{border}).wait();
			iane::copy(dres.const_view(), result.view()).wait();
		} else {
			iane::convolve(source.const_view().subview(coord_t{1, 0}, coord_t{2, 2}), kernel.const_view(), result.view(),
				{border});
		}
		const auto* expected = border == iane::border_mode::clamp ? clamp : zero;
		for (int i = 0; i < 4; ++i) {
			CHECK(result.pointer()[i] == expected[i]);
		}
	}
}

// test_convolution_cuda.cu:53-86: 5x4 source i-6, 3x3 kernel (i-3)/8, crop (1,1) 3x3; device == host exactly
void device_equals_host()
{
	using coord_t = iane::math::vec2i;
	iane::host_image<float, 2> source(coord_t{5, 4}), kernel(coord_t{3, 3}), reference(coord_t{3, 3}), output(coord_t{3, 3});
	for (int i = 0; i < 20; ++i) {
		source.pointer()[i] = static_cast<float>(i - 6);
	}
	for (int i = 0; i < 9; ++i) {
		kernel.pointer()[i] = static_cast<float>(i - 3) / 8.0f;
	}
	iane::device_image<float, 2> dsrc(source.dimensions()), dker(kernel.dimensions()), dout(output.dimensions());
	iane::copy(source.const_view(), dsrc.view()).wait();
	iane::copy(kernel.const_view(), dker.view()).wait();
	for (const auto border : {iane::border_mode::clamp, iane::border_mode::zero}) {
		iane::convolve(source.const_view().subview(coord_t{1, 1}, coord_t{3, 3}), kernel.const_view(), reference.view(),
			{.border = border});
		iane::convolve(dsrc.const_view().subview(coord_t{1, 1}, coord_t{3, 3}), dker.const_view(), dout.view(),
			{.border = border});
		iane::copy(dout.const_view(), output.view()).wait();
		for (int i = 0; i < 9; ++i) {
			CHECK(output.pointer()[i] == reference.pointer()[i]);
		}
	}
}

void policy_validation()
{
	iane::host_image<float, 2> a(4, 4), k(3, 3), b(4, 4);
	check_throws<std::invalid_argument>([&] {
		iane::convolve(a.const_view(), k.const_view(), b.view(), {}, iane::cpu_policy{0, 4});
	}, "threads == 0");
	check_throws<std::invalid_argument>([&] {
		iane::convolve(a.const_view(), k.const_view(), b.view(), {}, iane::cpu_policy{2, 0});
	}, "grain == 0");
	iane::host_image<float, 2> even(2, 3);
	check_throws<std::invalid_argument>([&] { iane::convolve(a.const_view(), even.const_view(), b.view()); }, "even kernel");
	iane::host_image<float, 2> c(4, 5);
	check_throws<std::invalid_argument>([&] { iane::convolve(a.const_view(), k.const_view(), c.view()); }, "extent mismatch");
	// a threaded run equals the serial one
	for (int i = 0; i < 16; ++i) {
		a.pointer()[i] = static_cast<float>(i * i % 7);
	}
	for (int i = 0; i < 9; ++i) {
		k.pointer()[i] = static_cast<float>(i) / 9.0f;
	}
	iane::host_image<float, 2> serial(4, 4);
	iane::convolve(a.const_view(), k.const_view(), serial.view());
	iane::convolve(a.const_view(), k.const_view(), b.view(), {}, iane::cpu_policy{3, 4});
	CHECK(std::memcmp(serial.pointer(), b.pointer(), 16 * sizeof(float)) == 0);
}

// prepared_cases.hpp:73-123 on the device through run_on
void fft_known_answers()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{4, 2};
	iane::host_image<iane::fft_complex, 2> input(size), output(size), round(size);
	for (int i = 0; i < 8; ++i) {
		input.pointer()[i] = {0.0f, 0.0f};
	}
	input.pointer()[1] = {1.0f, 0.0f};
	{
		const iane::hip_policy policy{};
		iane::device_image<iane::fft_complex, 2> source(size), destination(size), restored(size);
		iane::copy(input.const_view(), source.view(), policy).wait();
		auto context = iane::make_execution_context(policy);
		auto forward = context.prepare(iane::fft_forward, source.const_view(), destination.view());
		auto inverse = context.prepare(iane::fft_inverse, destination.const_view(), restored.view());
		context.run(forward, source.const_view(), destination.view());
		context.run(inverse, destination.const_view(), restored.view());
		iane::copy(destination.const_view(), output.view(), policy);
		iane::copy(restored.const_view(), round.view(), policy).wait();
		CHECK(context.scratch_capacity() == 0);
	}
	const float real[] = {1, 0, -1, 0};
	const float imag[] = {0, -1, 0, 1};
	for (int i = 0; i < 8; ++i) {
		CHECK(std::abs(output.pointer()[i].real - real[i % 4]) < 1e-5f);
		CHECK(std::abs(output.pointer()[i].imag - imag[i % 4]) < 1e-5f);
		CHECK(std::abs(round.pointer()[i].real - (i == 1 ? 8.0f : 0.0f)) < 1e-5f);
		CHECK(std::abs(round.pointer()[i].imag) < 1e-5f);
	}
	// aliasing is rejected (fft.hpp:44-53)
	iane::device_image<iane::fft_complex, 2> d(size);
	auto context = iane::make_execution_context(iane::hip_policy{});
	check_throws<std::invalid_argument>([&] { context.prepare(iane::fft_forward, d.const_view(), d.view()); }, "aliased FFT");
}

// real_fft_cases.hpp:12-77 for sizes 4x3, 5x3, 6x4
void real_fft_known_answers(iane::math::vec2i aSize)
{
	using coord_t = iane::math::vec2i;
	const auto packed_size = iane::rfft_extent(aSize);
	const int count = aSize[0] * aSize[1];
	iane::host_image<float, 2> input(aSize), output(aSize);
	iane::host_image<iane::fft_complex, 2> before(packed_size), after(packed_size);
	for (int i = 0; i < count; ++i) {
		input.pointer()[i] = static_cast<float>((i * 7) % 11 - 4);
	}
	{
		const iane::hip_policy policy{};
		iane::device_image<float, 2> source(aSize), destination(aSize);
		iane::device_image<iane::fft_complex, 2> packed(packed_size);
		iane::copy(input.const_view(), source.view(), policy).wait();
		auto exec = iane::make_execution_context(policy);
		auto forward = exec.prepare(iane::rfft, source.const_view(), packed.view());
		auto inverse = exec.prepare(iane::irfft, packed.const_view(), destination.view());
		exec.run(forward, source.const_view(), packed.view());
		iane::copy(packed.const_view(), before.view(), policy);
		exec.run(inverse, packed.const_view(), destination.view());
		iane::copy(packed.const_view(), after.view(), policy);
		iane::copy(destination.const_view(), output.view(), policy).wait();
	}
	for (int i = 0; i < count; ++i) {
		CHECK(std::abs(output.pointer()[i] - count * static_cast<float>((i * 7) % 11 - 4)) < 2e-4f);
	}
	for (int ky = 0; ky < aSize[1]; ++ky) {
		for (int kx = 0; kx < packed_size[0]; ++kx) {
			double re = 0, im = 0;
			for (int y = 0; y < aSize[1]; ++y) {
				for (int x = 0; x < aSize[0]; ++x) {
					const double phase = -2 * std::numbers::pi *
						(static_cast<double>(kx * x) / aSize[0] + static_cast<double>(ky * y) / aSize[1]);
					const double v = input.const_view()[coord_t{x, y}];
					re += v * std::cos(phase);
					im += v * std::sin(phase);
				}
			}
			const auto saved = before.const_view()[coord_t{kx, ky}];
			const auto kept = after.const_view()[coord_t{kx, ky}];
			CHECK(std::abs(saved.real - re) < 3e-5);
			CHECK(std::abs(saved.imag - im) < 3e-5);
			CHECK(std::memcmp(&saved, &kept, sizeof(saved)) == 0);
		}
	}
}

// The same FFT cases on host views (the reference's FFTW path): prepared_cases.hpp:73-123
// known answers with a round trip, test_prepared.cpp:41-61 (a plan reused on new buffers,
// aliasing rejected at run time, no scratch), and real_fft_cases.hpp:12-77.
void host_fft_known_answers()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{4, 2};
	iane::host_image<iane::fft_complex, 2> input(size), output(size), round(size);
	for (int i = 0; i < 8; ++i) {
		input.pointer()[i] = {0.0f, 0.0f};
	}
	input.pointer()[1] = {1.0f, 0.0f};
	auto context = iane::make_execution_context();
	auto forward = context.prepare(iane::fft_forward, input.const_view(), output.view());
	auto inverse = context.prepare(iane::fft_inverse, output.const_view(), round.view());
	context.run(forward, input.const_view(), output.view()).wait();
	context.run(inverse, output.const_view(), round.view()).wait();
	CHECK(context.scratch_capacity() == 0);
	const float real[] = {1, 0, -1, 0};
	const float imag[] = {0, -1, 0, 1};
	for (int i = 0; i < 8; ++i) {
		CHECK(std::abs(output.pointer()[i].real - real[i % 4]) < 1e-5f);
		CHECK(std::abs(output.pointer()[i].imag - imag[i % 4]) < 1e-5f);
		CHECK(std::abs(round.pointer()[i].real - (i == 1 ? 8.0f : 0.0f)) < 1e-5f);
		CHECK(std::abs(round.pointer()[i].imag) < 1e-5f);
	}
	// a plan prepared on prototypes runs on new buffers; aliasing is refused at run time
	iane::host_image<iane::fft_complex, 2> prototype_in(size), prototype_out(size), twos(size), spectrum(size);
	for (int i = 0; i < 8; ++i) {
		twos.pointer()[i] = {2.0f, 0.0f};
	}
	auto plan = context.prepare(iane::fft_forward, prototype_in.const_view(), prototype_out.view());
	check_throws<std::invalid_argument>([&] { context.run(plan, twos.const_view(), twos.view()); }, "aliased host FFT");
	context.run(plan, twos.const_view(), spectrum.view()).wait();
	CHECK(spectrum.pointer()[0].real == 16.0f);
	for (int i = 1; i < 8; ++i) {
		CHECK(spectrum.pointer()[i].real == 0.0f && spectrum.pointer()[i].imag == 0.0f);
	}
	CHECK(context.scratch_capacity() == 0);
	iane::host_image<iane::fft_complex, 2> other(coord_t{4, 3});
	check_throws<std::invalid_argument>([&] { context.run(plan, other.const_view(), other.view()); }, "shape");
}

void host_real_fft_known_answers(iane::math::vec2i aSize)
{
	using coord_t = iane::math::vec2i;
	const auto packed_size = iane::rfft_extent(aSize);
	const int count = aSize[0] * aSize[1];
	iane::host_image<float, 2> input(aSize), output(aSize);
	iane::host_image<iane::fft_complex, 2> packed(packed_size), before(packed_size);
	for (int i = 0; i < count; ++i) {
		input.pointer()[i] = static_cast<float>((i * 7) % 11 - 4);
	}
	auto exec = iane::make_execution_context();
	auto forward = exec.prepare(iane::rfft, input.const_view(), packed.view());
	auto inverse = exec.prepare(iane::irfft, packed.const_view(), output.view());
	CHECK(exec.scratch_capacity() == static_cast<std::size_t>(packed_size[0] * packed_size[1]) * sizeof(iane::fft_complex));
	exec.run(forward, input.const_view(), packed.view()).wait();
	std::memcpy(before.pointer(), packed.pointer(), sizeof(iane::fft_complex) * packed_size[0] * packed_size[1]);
	exec.run(inverse, packed.const_view(), output.view()).wait();
	for (int i = 0; i < count; ++i) {
		CHECK(std::abs(output.pointer()[i] - count * static_cast<float>((i * 7) % 11 - 4)) < 2e-4f);
		CHECK(input.pointer()[i] == static_cast<float>((i * 7) % 11 - 4));
	}
	for (int ky = 0; ky < aSize[1]; ++ky) {
		for (int kx = 0; kx < packed_size[0]; ++kx) {
			double re = 0, im = 0;
			for (int y = 0; y < aSize[1]; ++y) {
				for (int x = 0; x < aSize[0]; ++x) {
					const double phase = -2 * std::numbers::pi *
						(static_cast<double>(kx * x) / aSize[0] + static_cast<double>(ky * y) / aSize[1]);
					const double v = input.const_view()[coord_t{x, y}];
					re += v * std::cos(phase);
					im += v * std::sin(phase);
				}
			}
			const auto saved = before.const_view()[coord_t{kx, ky}];
			const auto kept = packed.const_view()[coord_t{kx, ky}];
			CHECK(std::abs(saved.real - re) < 3e-5);
			CHECK(std::abs(saved.imag - im) < 3e-5);
			CHECK(std::memcmp(&saved, &kept, sizeof(saved)) == 0);
		}
	}
}

// Host FFT lines of every kind of length (powers of two, smooth, primes up to 64
// on the generic radix, larger primes and lengths with a large prime factor on
// Bluestein's form) and 2-D / 3-D shapes against the direct float64 DFT: max error
// within 2e-6 of max|X|, and forward then inverse returns N x the input.
void host_fft_against_dft()
{
	const auto check_shape = [](const int* aExtent, int aRank) {
		const int n0 = aExtent[0], n1 = aRank > 1 ? aExtent[1] : 1, n2 = aRank > 2 ? aExtent[2] : 1;
		const int count = n0 * n1 * n2;
		std::vector<iane::fft_complex> x(count), X(count), back(count);
		for (int i = 0; i < count; ++i) {
			x[i] = {static_cast<float>(std::sin(0.37 * i + 0.1) + 0.25), static_cast<float>(std::cos(1.7 * i) * 0.5)};
		}
		auto ctx = iane::make_execution_context();
		double err = 0, peak = 0, round = 0;
		const auto dft = [&](int k0, int k1, int k2) {
			std::complex<double> sum = 0;
			for (int t2 = 0; t2 < n2; ++t2) {
				for (int t1 = 0; t1 < n1; ++t1) {
					for (int t0 = 0; t0 < n0; ++t0) {
						const double phase = -2 * std::numbers::pi * (static_cast<double>(k0) * t0 / n0 +
							static_cast<double>(k1) * t1 / n1 + static_cast<double>(k2) * t2 / n2);
						const auto& v = x[(t2 * n1 + t1) * n0 + t0];
						sum += std::complex<double>(v.real, v.imag) * std::polar(1.0, phase);
					}
				}
			}
			return sum;
		};
		if (aRank == 1) {
			auto src = iane::host::make_const_view<iane::fft_complex, 1>(x.data(), n0);
			auto dst = iane::host::make_view<iane::fft_complex, 1>(X.data(), n0);
			auto inv = iane::host::make_view<iane::fft_complex, 1>(back.data(), n0);
			auto f = ctx.prepare(iane::fft_forward, src, dst);
			auto b = ctx.prepare(iane::fft_inverse, iane::host::make_const_view<iane::fft_complex, 1>(X.data(), n0), inv);
			ctx.run(f, src, dst).wait();
			ctx.run(b, iane::host::make_const_view<iane::fft_complex, 1>(X.data(), n0), inv).wait();
		} else if (aRank == 2) {
			const iane::math::vec2i s{n0, n1};
			auto src = iane::host::make_const_view<iane::fft_complex, 2>(x.data(), s);
			auto dst = iane::host::make_view<iane::fft_complex, 2>(X.data(), s);
			auto inv = iane::host::make_view<iane::fft_complex, 2>(back.data(), s);
			auto f = ctx.prepare(iane::fft_forward, src, dst);
			auto b = ctx.prepare(iane::fft_inverse, iane::host::make_const_view<iane::fft_complex, 2>(X.data(), s), inv);
			ctx.run(f, src, dst).wait();
			ctx.run(b, iane::host::make_const_view<iane::fft_complex, 2>(X.data(), s), inv).wait();
		} else {
			const iane::math::vec3i s{n0, n1, n2};
			auto src = iane::host::make_const_view<iane::fft_complex, 3>(x.data(), s);
			auto dst = iane::host::make_view<iane::fft_complex, 3>(X.data(), s);
			auto inv = iane::host::make_view<iane::fft_complex, 3>(back.data(), s);
			auto f = ctx.prepare(iane::fft_forward, src, dst);
			auto b = ctx.prepare(iane::fft_inverse, iane::host::make_const_view<iane::fft_complex, 3>(X.data(), s), inv);
			ctx.run(f, src, dst).wait();
			ctx.run(b, iane::host::make_const_view<iane::fft_complex, 3>(X.data(), s), inv).wait();
		}
		for (int k2 = 0; k2 < n2; ++k2) {
			for (int k1 = 0; k1 < n1; ++k1) {
				for (int k0 = 0; k0 < n0; ++k0) {
					const auto want = dft(k0, k1, k2);
					const auto& got = X[(k2 * n1 + k1) * n0 + k0];
					err = std::max(err, std::abs(std::complex<double>(got.real, got.imag) - want));
					peak = std::max(peak, std::abs(want));
				}
			}
		}
		for (int i = 0; i < count; ++i) {
			round = std::max(round, std::abs(std::complex<double>(back[i].real, back[i].imag) -
				std::complex<double>(x[i].real, x[i].imag) * static_cast<double>(count)));
		}
		CHECK(err <= 2e-6 * peak);
		CHECK(round <= 4e-6 * peak);
		if (err > 2e-6 * peak || round > 4e-6 * peak) {
			std::printf("host FFT %dx%dx%d: error %.3g, round trip %.3g, peak %.3g\n", n0, n1, n2, err, round, peak);
		}
	};
	for (int n : {1, 2, 3, 4, 5, 7, 8, 12, 17, 31, 61, 64, 67, 97, 128, 210, 256, 331, 1009, 2 * 67, 4 * 101}) {
		check_shape(&n, 1);
	}
	const int shape2[][2] = {{6, 67}, {16, 9}, {1, 5}, {67, 3}};
	for (const auto& s : shape2) {
		check_shape(s, 2);
	}
	const int shape3[][3] = {{5, 4, 3}, {8, 1, 7}, {3, 71, 2}};
	for (const auto& s : shape3) {
		check_shape(s, 3);
	}
}

// main.cu:51-81 on the host with the header's host convolve (the reference's
// CPU semantics) and on the device with the DIRECT method in reference arithmetic.
void rl_device_equals_host_loop()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{23, 19};
	const coord_t ksize{5, 7};
	const float eps = 1e-6f;
	iane::host_image<float, 2> observed(size), psf(ksize), flipped(ksize);
	for (int i = 0; i < size[0] * size[1]; ++i) {
		observed.pointer()[i] = static_cast<float>((i * 37) % 101) + (i % 7 == 0 ? 0.0f : 3.5f);
	}
	double sum = 0.0;
	for (int i = 0; i < ksize[0] * ksize[1]; ++i) {
		psf.pointer()[i] = 1.0f + static_cast<float>((i * 13) % 5);
		sum += psf.pointer()[i];
	}
	for (int i = 0; i < ksize[0] * ksize[1]; ++i) {
		psf.pointer()[i] = static_cast<float>(psf.pointer()[i] / sum);
	}
	const int count = ksize[0] * ksize[1];
	std::reverse_copy(psf.pointer(), psf.pointer() + count, flipped.pointer());
	// host loop
	iane::host_image<float, 2> estimate(size), predicted(size), correction(size), sensitivity(size), ones(size);
	for (int i = 0; i < size[0] * size[1]; ++i) {
		ones.pointer()[i] = 1.0f;
		const float y = observed.pointer()[i];
		estimate.pointer()[i] = y > eps ? y : eps;
	}
	const iane::convolution_options zero{.border = iane::border_mode::zero};
	iane::convolve(ones.const_view(), flipped.const_view(), sensitivity.view(), zero);
	const int iterations = 7;
	for (int it = 0; it < iterations; ++it) {
		iane::convolve(estimate.const_view(), psf.const_view(), predicted.view(), zero);
		for (int i = 0; i < size[0] * size[1]; ++i) {
			const float p = predicted.pointer()[i];
			predicted.pointer()[i] = observed.pointer()[i] / (p > eps ? p : eps);
		}
		iane::convolve(predicted.const_view(), flipped.const_view(), correction.view(), zero);
		for (int i = 0; i < size[0] * size[1]; ++i) {
			const float s = sensitivity.pointer()[i];
			const float f = s > 0.0f ? correction.pointer()[i] / (s > eps ? s : eps) : 0.0f;
			estimate.pointer()[i] = estimate.pointer()[i] * f;
		}
	}
	// device prepared RL
	iane::device_image<float, 2> dobs(size), dpsf(ksize), dest(size);
	iane::copy(observed.const_view(), dobs.view()).wait();
	iane::copy(psf.const_view(), dpsf.view()).wait();
	auto run = iane::prepare_rl(dobs.const_view(), dpsf.const_view(), dest.view(), iane::hip_policy{}, iterations, eps,
		iane::rl_method::direct, iane::convolution_precision::reference);
	run().wait();
	iane::host_image<float, 2> result(size);
	iane::copy(dest.const_view(), result.view()).wait();
	CHECK(std::memcmp(result.pointer(), estimate.pointer(), sizeof(float) * size[0] * size[1]) == 0);
}

// Texture images: the cases of test_convolution_cuda.cu:10-163 as data.
static_assert(iane::concepts::DeviceView<iane::texture_image_view<float>>);
static_assert(!iane::concepts::MemoryView<iane::texture_image_view<float>>);
static_assert(!iane::concepts::HostView<iane::texture_image_view<float>>);

// One upload: the texels come from storage.subview(corner, size) of an int
// image whose value at s is bias + sum_a weight[a] * s[a] (host storage with
// the given byte pitch, 0 = packed, or a device copy of packed storage); the
// texture is read back through texture.subview(outer).subview(inner).
template<int N>
struct texture_upload_case {
	using ext = typename iane::math::dim_traits<N>::extents_t;
	const char* name;
	ext storage, pitch, corner, size;
	bool from_device;
	int bias;
	ext weight;
	ext outer_corner, outer_size, inner_corner, inner_size;
};

template<int N, class F>
void for_each_coord(const typename iane::math::dim_traits<N>::extents_t& aSize, F&& aVisit)
{
	if (iane::math::detail::element_count<N>(aSize) == 0) {
		return;
	}
	typename iane::math::dim_traits<N>::coord_t c{};
	do {
		aVisit(c);
	} while (iane::math::detail::advance<N>(c, aSize));
}

template<int N>
void run_texture_upload(const texture_upload_case<N>& aCase)
{
	using iane::math::detail::axis;
	using coord_t = typename iane::math::dim_traits<N>::coord_t;
	const auto value_at = [&](const coord_t& aStorage) {
		int v = aCase.bias;
		for (int a = 0; a < N; ++a) {
			v += axis(aCase.weight, a) * axis(aStorage, a);
		}
		return v;
	};
	// host storage: packed or with the case's pitch
	std::size_t span = sizeof(int);
	bool packed = true;
	for (int a = 0; a < N; ++a) {
		packed = packed && axis(aCase.pitch, a) == 0;
		span += static_cast<std::size_t>(axis(aCase.pitch, a)) * (axis(aCase.storage, a) - 1);
	}
	const auto layout = iane::resource::detail::packed_layout<int, N>(aCase.storage);
	std::vector<int> bytes((packed ? layout.first * sizeof(int) : span) / sizeof(int) + 1, -7);
	const iane::host_image_view<int, N> storage(bytes.data(), aCase.storage, packed ? layout.second : aCase.pitch);
	for_each_coord<N>(aCase.storage, [&](const coord_t& c) { storage[c] = value_at(c); });

	iane::texture_image<int, N> texture(aCase.size);
	iane::device_image<int, N> linear(aCase.storage);
	if (aCase.from_device) {
		iane::copy(iane::const_host_image_view<int, N>(storage), linear.view()).wait();
		iane::copy(linear.const_view().subview(aCase.corner, aCase.size), texture).wait();
	} else {
		iane::copy(iane::const_host_image_view<int, N>(storage).subview(aCase.corner, aCase.size), texture).wait();
	}
	iane::device_image<int, N> window(aCase.inner_size);
	iane::host_image<int, N> result(aCase.inner_size);
	iane::copy(texture.const_view().subview(aCase.outer_corner, aCase.outer_size).subview(aCase.inner_corner,
		aCase.inner_size), window.view());
	iane::copy(window.const_view(), result.view()).wait();
	for_each_coord<N>(aCase.inner_size, [&](const coord_t& w) {
		coord_t s{};
		for (int a = 0; a < N; ++a) {
			axis(s, a) = axis(aCase.corner, a) + axis(aCase.outer_corner, a) + axis(aCase.inner_corner, a) + axis(w, a);
		}
		++g_checks;
		if (result.view()[w] != value_at(s)) {
			++g_failures;
			std::fprintf(stderr, "texture upload %s: wrong texel\n", aCase.name);
		}
	});
}

void texture_uploads()
{
	using v2 = iane::math::vec2i;
	using v3 = iane::math::vec3i;
	// test_convolution_cuda.cu:14-51 (values i + 101 on a 4x3 image, a 2x2 crop at (1, 1);
	// once from the host view, once from a device copy shifted by 100)
	const texture_upload_case<2> planar[] = {
		{"2-D host crop", {4, 3}, {0, 0}, {1, 1}, {2, 2}, false, 101, {1, 4}, {0, 0}, {2, 2}, {0, 0}, {2, 2}},
		{"2-D device crop", {4, 3}, {0, 0}, {1, 1}, {2, 2}, true, 201, {1, 4}, {0, 0}, {2, 2}, {0, 0}, {2, 2}},
	};
	// test_convolution_cuda.cu:110-163 (100 z + 10 y + x; host rows padded to 16 bytes and
	// slices to 44, which is not a whole number of rows; a 5x4x4 device crop at (1, 1, 1)
	// shifted by 1000; read back through the 2x2x2 window at (1, 0, 1))
	const texture_upload_case<3> volumes[] = {
		{"3-D padded host", {3, 2, 3}, {4, 16, 44}, {0, 0, 0}, {3, 2, 3}, false, 0, {1, 10, 100}, {0, 0, 0},
			{3, 2, 3}, {1, 0, 1}, {2, 2, 2}},
		{"3-D device crop", {5, 4, 4}, {0, 0, 0}, {1, 1, 1}, {3, 2, 3}, true, 1000 - 111, {1, 10, 100}, {0, 0, 0},
			{3, 2, 3}, {1, 0, 1}, {2, 2, 2}},
	};
	// test_convolution_cuda.cu:98-108 (100 + i, a 4-sample crop at 1, read through
	// subview(1, 3).subview(1, 2))
	const texture_upload_case<1> lines[] = {
		{"1-D nested crop", 6, 0, 1, 4, false, 100, 1, 1, 3, 1, 2},
	};
	for (const auto& c : planar) {
		run_texture_upload(c);
	}
	for (const auto& c : volumes) {
		run_texture_upload(c);
	}
	for (const auto& c : lines) {
		run_texture_upload(c);
	}

	// rejected uploads and extents
	iane::host_image<int, 2> wide(v2{4, 3});
	iane::texture_image<int> small(v2{2, 2});
	check_throws<std::invalid_argument>([&] { iane::copy(wide.const_view(), small); }, "upload extent mismatch");
	int padded[33]{};
	const iane::host_image_view<int, 3> gapped_rows(padded, v3{2, 2, 2}, v3{8, 16, 44});
	iane::texture_image<int, 3> cube2(v3{2, 2, 2});
	check_throws<std::invalid_argument>([&] { iane::copy(gapped_rows, cube2); }, "non-contiguous rows");
	check_throws<std::invalid_argument>([] { iane::texture_image<float> bad(v2{-1, 2}); }, "negative extent");

	// an empty texture allocates nothing, uploads nothing and survives a move
	iane::texture_image<float> empty(v2{0, 3});
	auto moved = std::move(empty);
	iane::host_image<float, 2> nothing(v2{0, 3});
	iane::copy(nothing.const_view(), moved).wait();
	CHECK(moved.const_view().dimensions() == (v2{0, 3}));
	CHECK(moved.array() == nullptr);
}

// test_convolution_cuda.cu:53-86 with a texture source: the 5x4 image i - 6, the
// 3x3 kernel (i - 3) / 8, the 3x3 crop at (1, 1); equal to the host convolve for
// both borders.
void texture_convolution_matches_cpu()
{
	using v2 = iane::math::vec2i;
	iane::host_image<float, 2> source(v2{5, 4}), kernel(v2{3, 3}), expected(v2{3, 3}), got(v2{3, 3});
	std::iota(source.pointer(), source.pointer() + 20, -6.0f);
	for (int i = 0; i < 9; ++i) {
		kernel.pointer()[i] = static_cast<float>(i - 3) / 8.0f;
	}
	iane::texture_image<float> texture(source.dimensions());
	iane::device_image<float, 2> dkernel(kernel.dimensions()), dout(got.dimensions());
	iane::copy(source.const_view(), texture).wait();
	iane::copy(kernel.const_view(), dkernel.view()).wait();
	for (const auto border : {iane::border_mode::clamp, iane::border_mode::zero}) {
		iane::convolve(source.const_view().subview(v2{1, 1}, v2{3, 3}), kernel.const_view(), expected.view(),
			{.border = border});
		iane::convolve(texture.const_view().subview(v2{1, 1}, v2{3, 3}), dkernel.const_view(), dout.view(),
			{.border = border});
		iane::copy(dout.const_view(), got.view()).wait();
		CHECK(std::memcmp(got.pointer(), expected.pointer(), 9 * sizeof(float)) == 0);
	}
}

// Scan / reduce (test/unit/prepared/sequence_cases.hpp:31-117): a 263 x 5 u8
// image x % 7 + y; row and column scans composed into 2-D prefix sums, flat
// exclusive and inclusive scans seeded with 11, whole and per-row reductions,
// exclusive column scans. The host runs the reference's u64 functor; the device
// (whose C ABI carries operator codes and has no u64) uses std::plus<int32_t>.
struct sequence_plus {
	std::uint64_t operator()(std::uint64_t aLeft, std::uint64_t aRight) const { return aLeft + aRight; }
};

template<bool tDevice>
void sequence_case()
{
	using coord_t = iane::math::vec2i;
	constexpr int width = 263;
	constexpr int height = 5;
	const coord_t size{width, height};
	using acc_t = std::conditional_t<tDevice, std::int32_t, std::uint64_t>;
	using op_t = std::conditional_t<tDevice, std::plus<std::int32_t>, sequence_plus>;
	iane::host_image<std::uint8_t, 2> input(size);
	for (int y = 0; y < height; ++y) {
		for (int x = 0; x < width; ++x) {
			input.view()[coord_t{x, y}] = static_cast<std::uint8_t>(x % 7 + y);
		}
	}
	iane::host_image<acc_t, 2> output(size), prefixes(size);
	iane::host_image<acc_t, 1> total(1), totals(height);
	constexpr acc_t seed = 11;
	const op_t plus{};
	const auto check_results = [&](bool aInclusive) {
		acc_t prefix = seed;
		for (int y = 0; y < height; ++y) {
			acc_t row = seed;
			for (int x = 0; x < width; ++x) {
				const auto pixel = static_cast<acc_t>(x % 7 + y);
				CHECK((prefixes.view()[coord_t{x, y}] == prefix + (aInclusive ? pixel : 0)));
				prefix += pixel;
				row += pixel;
			}
			CHECK(totals.view()[y] == row);
		}
		CHECK(total.view()[0] == prefix);
	};
	const auto check_columns = [&] {
		for (int x = 0; x < width; ++x) {
			acc_t prefix = seed;
			for (int y = 0; y < height; ++y) {
				CHECK((prefixes.view()[coord_t{x, y}] == prefix));
				prefix += static_cast<acc_t>(x % 7 + y);
			}
		}
	};
	const auto check_rectangles = [&] {
		for (int y = 0; y < height; ++y) {
			acc_t rectangle = 0;
			for (int x = 0; x < width; ++x) {
				for (int row = 0; row <= y; ++row) {
					rectangle += static_cast<acc_t>(x % 7 + row);
				}
				CHECK((output.const_view()[coord_t{x, y}] == rectangle));
			}
		}
	};
	if constexpr (tDevice) {
		iane::device_image<std::uint8_t, 2> source(size);
		iane::device_image<acc_t, 2> intermediate(coord_t{width + 2, height}), destination(size);
		iane::device_image<acc_t, 1> scalar(1), row_totals(height);
		iane::copy(input.const_view(), source.view()).wait();
		auto rows = intermediate.view().subview(coord_t{1, 0}, size);
		auto exec = iane::make_execution_context(iane::hip_policy{});
		auto src = source.const_view();
		auto dst = destination.view();
		auto row_scan = exec.prepare(iane::inclusive_scan_along<0>(plus, acc_t{}), src, rows);
		auto column_scan = exec.prepare(iane::inclusive_scan_along<1>(plus, acc_t{}), iane::const_device_image_view<acc_t, 2>(rows), dst);
		auto flat_scan = exec.prepare(iane::exclusive_scan(plus, seed), src, dst);
		auto full_reduce = exec.prepare(iane::reduce(plus, seed), src, scalar.view());
		auto row_reduce = exec.prepare(iane::reduce_along<0>(plus, seed), src, row_totals.view());
		auto exclusive_columns = exec.prepare(iane::exclusive_scan_along<1>(plus, seed), src, dst);
		auto inclusive_flat = exec.prepare(iane::inclusive_scan(plus, seed), src, dst);
		exec.run(full_reduce, src, scalar.view());
		exec.run(row_reduce, src, row_totals.view());
		iane::copy(scalar.const_view(), total.view()).wait();
		iane::copy(row_totals.const_view(), totals.view()).wait();
		for (bool inclusive : {false, true}) {
			exec.run(inclusive ? inclusive_flat : flat_scan, src, dst);
			iane::copy(destination.const_view(), prefixes.view()).wait();
			check_results(inclusive);
		}
		exec.run(exclusive_columns, src, dst);
		iane::copy(destination.const_view(), prefixes.view()).wait();
		check_columns();
		exec.run(row_scan, src, rows);
		exec.run(column_scan, iane::const_device_image_view<acc_t, 2>(rows), dst).wait();
		iane::copy(destination.const_view(), output.view()).wait();
		check_rectangles();
	} else {
		iane::host_image<acc_t, 2> intermediate(coord_t{width + 2, height});
		auto rows = intermediate.view().subview(coord_t{1, 0}, size);
		auto exec = iane::make_execution_context(iane::cpu_policy{.threads = 8, .grain_size = 31});
		auto src = input.const_view();
		auto dst = output.view();
		auto row_scan = exec.prepare(iane::inclusive_scan_along<0>(plus, acc_t{}), src, rows);
		auto column_scan = exec.prepare(iane::inclusive_scan_along<1>(plus, acc_t{}), rows, dst);
		auto flat_scan = exec.prepare(iane::exclusive_scan(plus, seed), src, prefixes.view());
		auto full_reduce = exec.prepare(iane::reduce(plus, seed), src, total.view());
		auto row_reduce = exec.prepare(iane::reduce_along<0>(plus, seed), src, totals.view());
		auto exclusive_columns = exec.prepare(iane::exclusive_scan_along<1>(plus, seed), src, prefixes.view());
		auto inclusive_flat = exec.prepare(iane::inclusive_scan(plus, seed), src, prefixes.view());
		exec.run(full_reduce, src, total.view());
		exec.run(row_reduce, src, totals.view());
		exec.run(flat_scan, src, prefixes.view());
		check_results(false);
		exec.run(inclusive_flat, src, prefixes.view());
		check_results(true);
		exec.run(exclusive_columns, src, prefixes.view());
		check_columns();
		exec.run(row_scan, src, rows);
		exec.run(column_scan, rows, dst);
		check_rectangles();
	}
	check_throws<std::invalid_argument>([&] {
		auto exec = iane::make_execution_context(iane::cpu_policy{});
		iane::host_image<acc_t, 1> wrong(height + 1);
		(void)exec.prepare(iane::reduce_along<0>(plus, seed), input.const_view(), wrong.view());
	}, "reduce_along extent mismatch");
}

// ---------------------------------------------------- expressions ------
// copy_cases.hpp:38-84 (procedural copy over a crop, view algebra),
// test_copy.cpp (laziness, zip extent check), transform_cases.hpp:37-70
// (value / position / locator transforms), test_for_each.cpp:12-56 (empty
// domain, the cpu_policy partition over a strided crop, errors reaching the caller).
void host_expressions()
{
	using v2 = iane::math::vec2i;
	{
		iane::host_image<int, 2> input(v2{3, 2}), output(v2{2, 2});
		std::iota(input.pointer(), input.pointer() + 6, 1);
		const auto source = input.const_view().subview(v2{1, 0}, v2{2, 2});
		const auto mapped = iane::map_view(source, [](int a) { return 2 * a; });
		const auto column = iane::generate_view(source, [](const v2& p) { return p[0]; });
		const auto bias = iane::zip_view(column, iane::constant_view(source, 1), std::plus<int>{});
		static_assert(!iane::concepts::MemoryView<decltype(bias)>);
		iane::copy(iane::zip_view(mapped, bias, std::plus<int>{}), output.view()).wait();
		const int expected[] = {5, 8, 11, 14};
		CHECK(std::equal(expected, expected + 4, output.pointer()));
	}
	{
		iane::host_image<int, 1> input(3), output(3);
		input.pointer()[0] = 2;
		input.pointer()[1] = 3;
		input.pointer()[2] = 5;
		const auto s = input.const_view();
		iane::copy((s * 2 + 1) + (20 / s) + -(10 - s), output.view());
		const int expected[] = {7, 6, 10};
		CHECK(std::equal(expected, expected + 3, output.pointer()));
		// lazy: the map sees later writes; zip checks extents
		const auto doubled = iane::map_view(s, [](int a) { return 2 * a; });
		input.pointer()[1] = 9;
		CHECK(doubled[1] == 18);
		iane::host_image<int, 1> shorter(2);
		check_throws<std::invalid_argument>([&] {
			(void)iane::zip_view(doubled, iane::constant_view(shorter.const_view(), 1), std::plus<int>{});
		}, "zip extents");
		// the math functions of the algebra follow std:: on the pixel type
		iane::host_image<float, 1> f(3), g(3);
		f.pointer()[0] = 1.0f;
		f.pointer()[1] = 4.0f;
		f.pointer()[2] = 9.0f;
		iane::copy(iane::sqrt(f.const_view()) + iane::log(f.const_view()), g.view());
		CHECK(g.pointer()[1] == std::sqrt(4.0f) + std::log(4.0f));
	}
	{
		// transform_cases.hpp:37-70 on a 3 x 3 x 3 pixel pattern
		using v3 = iane::math::vec3i;
		iane::host_image<pixel, 3> input(v3{3, 3, 3}), output(v3{3, 3, 3});
		for (int i = 0; i < 27; ++i) {
			input.pointer()[i] = {i + 1, -i - 1};
		}
		iane::transform(input.const_view(), output.view(), [](pixel a) { return a + pixel{1, -1}; });
		iane::transform(output.const_view(), output.view(), [](pixel a) { return a + pixel{1, -1}; });
		for (int i = 0; i < 27; ++i) {
			CHECK(output.pointer()[i].first == i + 3 && output.pointer()[i].second == -i - 3);
		}
		iane::transform_position(input.const_view(), output.view(), [](pixel a, const v3& p) { return a + pixel{p[2], 0}; });
		for (int i = 0; i < 27; ++i) {
			CHECK(output.pointer()[i].first == i + 1 + i / 9);
		}
		iane::transform_locator(input.const_view(), output.view(), [](auto aLocator) {
			return aLocator[v3{0, 0, 1}];
		}, iane::border_mode::clamp);
		for (int i = 0; i < 27; ++i) {
			const int next = i / 9 == 2 ? i + 1 : i + 10;
			CHECK(output.pointer()[i].first == next);
		}
		iane::transform_locator(input.const_view(), output.view(), [](auto aLocator) {
			return aLocator[v3{0, 0, 1}];
		}, iane::border_mode::zero);
		CHECK(output.pointer()[26].first == 0 && output.pointer()[0].first == 10);
		// image_locator on a crop: the crop's extent is the boundary
		const auto crop = input.const_view().subview(v3{1, 1, 1}, v3{2, 2, 2});
		const iane::image_locator<decltype(crop)> at_corner(crop, v3{0, 0, 0}, iane::border_mode::clamp);
		CHECK(at_corner[v3{-1, 0, 0}].first == 14 && at_corner.position() == (v3{0, 0, 0}));
		const iane::image_locator<decltype(crop)> zeroed(crop, v3{0, 0, 0}, iane::border_mode::zero);
		CHECK(zeroed[v3{-1, 0, 0}].first == 0 && zeroed[v3{1, 1, 1}].first == 27);
	}
	{
		// test_for_each.cpp:12-19, 21-41, 43-56
		iane::host_image<int, 4> empty(iane::math::vec4i{2, 2, 0, 3});
		int visits = 0;
		iane::for_each(empty.view(), [&](int&) { ++visits; });
		CHECK(visits == 0);
		iane::host_image<int, 2> image(v2{7, 5});
		iane::for_each(image.view(), [](int& a) { a = 0; });
		const auto crop = image.view().subview(v2{1, 1}, v2{5, 3});
		std::thread::id workers[15];
		iane::for_each_position(crop, [&](int& a, const v2& p) {
			++a;
			workers[p[0] + 5 * p[1]] = std::this_thread::get_id();
		}, iane::cpu_policy{.threads = 3, .grain_size = 4}).wait();
		for (int y = 0; y < 5; ++y) {
			for (int x = 0; x < 7; ++x) {
				CHECK(image.const_view()[v2{x, y}] == (x >= 1 && x <= 5 && y >= 1 && y <= 3 ? 1 : 0));
			}
		}
		CHECK(std::set<std::thread::id>(std::begin(workers), std::end(workers)).size() == 3u);
		iane::host_image<int, 1> line(17), copy_out(17);
		check_throws<std::invalid_argument>([&] {
			iane::copy(line.const_view(), copy_out.view(), iane::cpu_policy{.threads = 0});
		}, "threads 0");
		check_throws<std::invalid_argument>([&] {
			iane::for_each(line.view(), [](int&) {}, iane::cpu_policy{.grain_size = 0});
		}, "grain 0");
		check_throws<std::logic_error>([&] {
			iane::for_each_position(line.view(), [](int&, int p) {
				if (p == 16) {
					throw std::logic_error("worker failure");
				}
			}, iane::cpu_policy{.threads = 3, .grain_size = 4});
		}, "worker exception");
		std::iota(line.pointer(), line.pointer() + 17, 1);
		int located = 0;
		iane::for_each_locator(line.const_view(), [&](auto aLocator) {
			located += aLocator[1] == (aLocator.position() == 16 ? 0 : aLocator.position() + 2);
		}, iane::border_mode::zero);
		CHECK(located == 17);
	}
	{
		// slice<tI>() and image reset()
		iane::host_image<int, 3> volume(iane::math::vec3i{4, 3, 2});
		std::iota(volume.pointer(), volume.pointer() + 24, 0);
		const auto plane = volume.const_view().slice<2>(1);
		CHECK(plane.dimensions() == (v2{4, 3}) && plane[v2{2, 1}] == 12 + 4 + 2);
		const auto columns = volume.view().slice<0>(3);
		CHECK(columns.dimensions() == (v2{3, 2}) && columns[v2{1, 1}] == 12 + 4 + 3);
		const auto row = plane.slice<1>(2);
		CHECK(row.dimensions() == 4 && row[1] == 12 + 8 + 1);
		volume.reset();
		CHECK(volume.pointer() == nullptr && volume.dimensions() == (iane::math::vec3i{0, 0, 0}));
	}
}

// The pointwise functions of the RL loop (examples/rl_deconvolution/main.cu:11-49)
// as iane::pointwise expressions: valid on host and device views alike.
struct rl_functions {
	float eps;
	auto initial() const { using namespace iane::pointwise; return select(_1 > eps, _1, eps); }
	auto ratio() const { using namespace iane::pointwise; return _1 / select(_2 > eps, _2, eps); }
	auto edge() const { using namespace iane::pointwise; return select(_2 > 0.0f, _1 / select(_2 > eps, _2, eps), 0.0f); }
	auto multiply() const { using namespace iane::pointwise; return _1 * _2; }
};

template<class TImage, int N>
auto make_image(typename iane::math::dim_traits<N>::extents_t aSize)
{
	return TImage(aSize);
}

// prepare_rl of examples/rl_deconvolution/main.cu:51-81, stated with the same
// calls: constant_view sensitivity, transform for the initial estimate, and
// copy(zip_view(...)) updates between the two convolutions per iteration.
template<class TImage, class TSrc, class TPsf, class TDst>
void rl_as_the_reference_writes_it(TSrc aSource, TPsf aPsf, TPsf aAdjoint, TDst aDestination,
	iane::view_policy<TDst> aPolicy, int aIterations, float aEpsilon)
{
	const auto size = aDestination.dimensions();
	TImage estimate(size), predicted(size), correction(size), sensitivity(size);
	const rl_functions f{aEpsilon};
	const iane::convolution_options options{.border = iane::border_mode::zero};
	iane::convolve(iane::constant_view(sensitivity.view(), 1.0f), aAdjoint, sensitivity.view(), options, aPolicy).wait();
	iane::transform(aSource, estimate.view(), f.initial(), aPolicy);
	for (int it = 0; it < aIterations; ++it) {
		iane::convolve(estimate.const_view(), aPsf, predicted.view(), options, aPolicy);
		iane::copy(iane::zip_view(aSource, predicted.const_view(), f.ratio()), predicted.view(), aPolicy);
		iane::convolve(predicted.const_view(), aAdjoint, correction.view(), options, aPolicy);
		const auto factor = iane::zip_view(correction.const_view(), sensitivity.const_view(), f.edge());
		iane::copy(iane::zip_view(estimate.const_view(), factor, f.multiply()), estimate.view(), aPolicy);
	}
	iane::copy(estimate.const_view(), aDestination, aPolicy).wait();
}

std::vector<float> read_floats(const std::string& aPath, std::size_t aCount)
{
	std::vector<float> data(aCount);
	FILE* f = std::fopen(aPath.c_str(), "rb");
	const bool ok = f && std::fread(data.data(), sizeof(float), aCount, f) == aCount;
	if (f) {
		std::fclose(f);
	}
	if (!ok) {
		throw std::runtime_error("cannot read " + aPath);
	}
	return data;
}

// The golden RL fixture in aDir (observed, psf, expected as raw float32 in
// numpy order; meta.txt: rank, extents axis 0 first, psf extents, iterations,
// epsilon) run through rl_as_the_reference_writes_it on host and device views.
template<int N>
void rl_golden_case(const std::string& aDir, const int* aExtent, const int* aPsfExtent, int aIterations, float aEps,
	bool aDevice)
{
	using ext = typename iane::math::dim_traits<N>::extents_t;
	ext size{}, ksize{};
	std::size_t count = 1, kcount = 1;
	for (int a = 0; a < N; ++a) {
		iane::math::detail::axis(size, a) = aExtent[a];
		iane::math::detail::axis(ksize, a) = aPsfExtent[a];
		count *= aExtent[a];
		kcount *= aPsfExtent[a];
	}
	const auto observed = read_floats(aDir + "/observed.f32", count);
	const auto psf = read_floats(aDir + "/psf.f32", kcount);
	const auto expected = read_floats(aDir + "/expected.f32", count);
	std::vector<float> flipped(psf.rbegin(), psf.rend());
	iane::host_image<float, N> result(size);
	if (!aDevice) {
		const auto obs = iane::host::make_const_view<float, N>(observed.data(), size);
		const auto h = iane::host::make_const_view<float, N>(psf.data(), ksize);
		const auto ht = iane::host::make_const_view<float, N>(flipped.data(), ksize);
		rl_as_the_reference_writes_it<iane::host_image<float, N>>(obs, h, ht, result.view(), iane::cpu_policy{},
			aIterations, aEps);
	} else {
		iane::device_image<float, N> dobs(size), dpsf(ksize), dflip(ksize), dout(size);
		iane::copy(iane::host::make_const_view<float, N>(observed.data(), size), dobs.view()).wait();
		iane::copy(iane::host::make_const_view<float, N>(psf.data(), ksize), dpsf.view()).wait();
		iane::copy(iane::host::make_const_view<float, N>(flipped.data(), ksize), dflip.view()).wait();
		rl_as_the_reference_writes_it<iane::device_image<float, N>>(dobs.const_view(), dpsf.const_view(),
			dflip.const_view(), dout.view(), iane::hip_policy{}, aIterations, aEps);
		iane::copy(dout.const_view(), result.view()).wait();
	}
	CHECK(std::memcmp(result.pointer(), expected.data(), count * sizeof(float)) == 0);
}

int rl_golden(const std::string& aDir, bool aDevice)
{
	FILE* f = std::fopen((aDir + "/meta.txt").c_str(), "r");
	if (!f) {
		std::fprintf(stderr, "no meta.txt in %s\n", aDir.c_str());
		return 2;
	}
	int rank = 0, extent[4] = {}, kext[4] = {}, iterations = 0;
	float eps = 0.0f;
	int ok = std::fscanf(f, "%d", &rank) == 1;
	for (int a = 0; ok && a < rank; ++a) {
		ok = std::fscanf(f, "%d", &extent[a]) == 1;
	}
	for (int a = 0; ok && a < rank; ++a) {
		ok = std::fscanf(f, "%d", &kext[a]) == 1;
	}
	ok = ok && std::fscanf(f, "%d %f", &iterations, &eps) == 2;
	std::fclose(f);
	if (!ok) {
		return 2;
	}
	if (rank == 2) {
		rl_golden_case<2>(aDir, extent, kext, iterations, eps, aDevice);
	} else {
		rl_golden_case<3>(aDir, extent, kext, iterations, eps, aDevice);
	}
	return 0;
}

// The device side of the expression API: the pointwise programs against the
// same functions evaluated on the host (bit-exact for + - * / select and the
// comparisons), the procedural copy and view algebra of copy_cases.hpp on
// device views, device for_each on the stream it was given
// (test_for_each_cuda.cu:19-36), and errors.
// texture_image_view::sample on device views, as device expressions.
// test_convolution_cuda.cu:170-191: point samples of a cropped 1-D texture at
// fractional coordinates (clamped to the crop's edge texel centres), an empty
// view giving 0; then a 2-D crop sampled on a coordinate grid, a NaN coordinate
// giving 0, against the same rule evaluated on the host.
void texture_sample()
{
	using namespace iane::pointwise;
	{
		iane::host_image<int, 1> source(4);
		const int values[]{99, 10, 20, 99};
		std::copy_n(values, 4, source.pointer());
		iane::texture_image<int, 1> texture(4);
		iane::copy(source.const_view(), texture).wait();
		const auto crop = texture.const_view().subview(1, 3).subview(0, 2);
		const auto empty = crop.subview(0, 0);
		const float coordinates[]{-1.0f, 0.0f, 0.5f, 0.75f, 1.0f, 1.5f, 2.0f, 0.0f};
		iane::host_image<float, 1> coords(8);
		std::copy_n(coordinates, 8, coords.pointer());
		iane::device_image<float, 1> dcoords(8);
		iane::copy(coords.const_view(), dcoords.view()).wait();
		iane::device_image<int, 1> device_output(8);
		const auto index = iane::coordinate_view(device_output.view(), 0);
		const auto from_crop = iane::sample(crop, dcoords.const_view());
		const auto from_empty = iane::sample(empty, dcoords.const_view());
		const auto picked = iane::device_expression<1>::apply(8, select(_1 == 7.0f, _2, _3),
			{&index, &from_empty, &from_crop});
		iane::copy(picked, device_output.view()).wait();
		iane::host_image<int, 1> output(8);
		iane::copy(device_output.const_view(), output.view()).wait();
		const int expected[]{10, 10, 10, 10, 20, 20, 20, 0};
		CHECK(std::equal(expected, expected + 8, output.pointer()));
	}
	{
		using v2 = iane::math::vec2i;
		iane::host_image<float, 2> source(v2{5, 4});
		for (int y = 0; y < 4; ++y) {
			for (int x = 0; x < 5; ++x) {
				source.view()[v2{x, y}] = static_cast<float>(10 * y + x);
			}
		}
		iane::texture_image<float, 2> texture(v2{5, 4});
		iane::copy(source.const_view(), texture).wait();
		const auto crop = texture.const_view().subview(v2{1, 1}, v2{3, 2});
		const v2 grid{7, 5};
		iane::device_image<float, 2> out(grid);
		const auto gx = iane::coordinate_view(out.view(), 0);
		const auto gy = iane::coordinate_view(out.view(), 1);
		const auto cx = iane::device_expression<2>::apply(grid, select(_1 == 6.0f, value(NAN), _1 * 0.7f - 0.4f), {&gx});
		const auto cy = iane::device_expression<2>::apply(grid, _1 * 0.9f + 0.2f, {&gy});
		iane::copy(iane::sample(crop, cx, cy), out.view()).wait();
		iane::host_image<float, 2> got(grid);
		iane::copy(out.const_view(), got.view()).wait();
		const auto point = [](float c, int extent) {
			const float last = static_cast<float>(extent) - 0.5f;
			return static_cast<int>(std::floor(c < 0.5f ? 0.5f : (c > last ? last : c)));
		};
		bool same = true;
		for (int y = 0; y < grid[1]; ++y) {
			for (int x = 0; x < grid[0]; ++x) {
				const float sx = static_cast<float>(x) * 0.7f - 0.4f;
				const float sy = static_cast<float>(y) * 0.9f + 0.2f;
				const float expect = x == 6 ? 0.0f : source.const_view()[v2{1 + point(sx, 3), 1 + point(sy, 2)}];
				same = same && got.const_view()[v2{x, y}] == expect;
			}
		}
		CHECK(same);
	}
}

void device_expressions()
{
	using v2 = iane::math::vec2i;
	using namespace iane::pointwise;
	const v2 size{37, 23};
	iane::host_image<float, 2> a(size), b(size), got(size);
	iane::host_image<std::uint16_t, 2> c(size);
	for (int i = 0; i < size[0] * size[1]; ++i) {
		a.pointer()[i] = static_cast<float>((i * 7919) % 1000) / 37.0f - 9.0f;
		b.pointer()[i] = static_cast<float>((i * 104729) % 997) / 13.0f;
		c.pointer()[i] = static_cast<std::uint16_t>((i * 31) % 65536);
	}
	iane::device_image<float, 2> da(size), db(size), dout(size);
	iane::device_image<std::uint16_t, 2> dc(size);
	iane::copy(a.const_view(), da.view()).wait();
	iane::copy(b.const_view(), db.view()).wait();
	iane::copy(c.const_view(), dc.view()).wait();
	const auto f = select(_1 > _2, (_1 - _2) / (_2 + 1.5f), min(_1 * _3, max(_2, 0.25f))) + sqrt(abs(_1)) * 0.5f;
	const auto ea = iane::device_expression<2>::load(da.const_view());
	const auto eb = iane::device_expression<2>::load(db.const_view());
	const auto ec = iane::device_expression<2>::load(dc.const_view());
	const auto dev = iane::device_expression<2>::apply(size, f, {&ea, &eb, &ec});
	iane::copy(dev, dout.view()).wait();
	iane::copy(dout.const_view(), got.view()).wait();
	bool same = true;
	for (int i = 0; i < size[0] * size[1]; ++i) {
		same = same && got.pointer()[i] == f(a.pointer()[i], b.pointer()[i], c.pointer()[i]);
	}
	CHECK(same);
	// strided (every other column) destination and a reversed source through zip_view + algebra
	iane::device_image<float, 2> wide(v2{2 * size[0], size[1]});
	const iane::device_image_view<float, 2> every_other(wide.pointer(), size, v2{8, 8 * size[0]});
	iane::copy((da.const_view() + db.const_view()) * 2.0f - 1.0f, every_other).wait();
	iane::host_image<float, 2> wide_host(v2{2 * size[0], size[1]});
	iane::copy(wide.const_view(), wide_host.view()).wait();
	same = true;
	for (int y = 0; y < size[1]; ++y) {
		for (int x = 0; x < size[0]; ++x) {
			const float expect = (a.const_view()[v2{x, y}] + b.const_view()[v2{x, y}]) * 2.0f - 1.0f;
			same = same && wide_host.const_view()[v2{2 * x, y}] == expect;
		}
	}
	CHECK(same);
	// copy_cases.hpp:38-84 on device views (generate_view aside: no device compiler here)
	{
		iane::host_image<int, 2> input(v2{3, 2}), output(v2{2, 2});
		std::iota(input.pointer(), input.pointer() + 6, 1);
		iane::device_image<int, 2> din(input.dimensions()), dres(output.dimensions());
		iane::copy(input.const_view(), din.view()).wait();
		const auto source = din.const_view().subview(v2{1, 0}, v2{2, 2});
		const auto mapped = iane::map_view(source, _1 * 2);
		iane::copy(iane::zip_view(mapped, iane::constant_view(source, 1), _1 + _2), dres.view()).wait();
		iane::copy(dres.const_view(), output.view()).wait();
		const int expected[] = {5, 7, 11, 13};
		CHECK(std::equal(expected, expected + 4, output.pointer()));
		iane::host_image<int, 1> s(3), r(3);
		s.pointer()[0] = 2;
		s.pointer()[1] = 3;
		s.pointer()[2] = 5;
		iane::device_image<int, 1> ds(3), dr(3);
		iane::copy(s.const_view(), ds.view()).wait();
		const auto v = ds.const_view();
		iane::copy((v * 2 + 1) + (20 / v) + -(10 - v), dr.view()).wait();
		iane::copy(dr.const_view(), r.view()).wait();
		const int algebra[] = {7, 6, 10};
		CHECK(std::equal(algebra, algebra + 3, r.pointer()));
	}
	// device for_each on the caller's stream
	{
		iane::host_image<int, 1> output(17);
		iane::run_on(iane::run_options{.backend = iane::backend::hip}, output.view(), [](auto aDst, auto aExecution) {
			if constexpr (iane::concepts::DeviceView<decltype(aDst)>) {
				const iane::hip_policy policy{.stream = aExecution.stream};
				const auto result = iane::for_each(aDst, value(42.0f), policy);
				CHECK(result.stream == aExecution.stream);
				return result;
			} else {
				return aExecution;
			}
		});
		CHECK(std::all_of(output.pointer(), output.pointer() + 17, [](int x) { return x == 42; }));
	}
	check_throws<std::invalid_argument>([&] {
		iane::device_image<float, 2> other(v2{3, 3});
		(void)iane::zip_view(da.const_view(), other.const_view(), _1 + _2);
	}, "device zip extents");
	check_throws<std::invalid_argument>([&] {
		iane::device_image<float, 2> k(v2{3, 3});
		iane::convolve(da.const_view() * 2.0f, k.const_view(), dout.view()).wait();
	}, "expression convolve source");
}

} // namespace

int main(int argc, char** argv)
{
	if (argc > 2 && std::strcmp(argv[1], "rl-golden") == 0) {
		const bool on_device = argc > 3 && std::strcmp(argv[3], "device") == 0;
		if (on_device) {
			int n = 0;
			if (iane_hip_device_count(&n) != IANE_OK || n == 0) {
				std::puts("no HIP device: skipped");
				return 77;
			}
		}
		const int status = rl_golden(argv[2], on_device);
		std::printf("rl-golden %s: %d checks, %d failures\n", on_device ? "device" : "host", g_checks, g_failures);
		return status ? status : (g_failures ? 1 : 0);
	}
	const bool device = argc > 1 && std::strcmp(argv[1], "device") == 0;
	if (!device) {
		host_expressions();
		pattern_host<2>();
		pattern_host<3>();
		pattern_host<4>();
		scalar_coordinates();
		crop_flip(false);
		policy_validation();
		sequence_case<false>();
		host_fft_known_answers();
		host_real_fft_known_answers({4, 3});
		host_real_fft_known_answers({5, 3});
		host_real_fft_known_answers({6, 4});
		host_fft_against_dft();
	} else {
		int n = 0;
		if (iane_hip_device_count(&n) != IANE_OK || n == 0) {
			std::puts("no HIP device: skipped");
			return 77;
		}
		pattern_device<2>();
		pattern_device<3>();
		pattern_device<4>();
		crop_flip(true);
		device_equals_host();
		fft_known_answers();
		real_fft_known_answers({4, 3});
		real_fft_known_answers({5, 3});
		real_fft_known_answers({6, 4});
		rl_device_equals_host_loop();
		texture_uploads();
		texture_convolution_matches_cpu();
		device_expressions();
		texture_sample();
		sequence_case<true>();
	}
	std::printf("%s: %d checks, %d failures\n", device ? "device" : "host", g_checks, g_failures);
	return g_failures ? 1 : 0;
}
