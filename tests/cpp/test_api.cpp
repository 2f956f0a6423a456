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
#include <vector>

namespace {

int g_failures = 0;
int g_checks = 0;

#define CHECK(cond)                                                                     \
	do {                                                                                \
		++g_checks;                                                                     \
		if (!(cond)) {                                                                  \
			++g_failures;                                                               \
			std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
		}                                                                               \
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

// Texture images (test_convolution_cuda.cu:10-163): uploads from strided host and
// device views, convolution of a cropped texture view equal to the host
// convolve for both borders, empty textures and moves, nested 1-D crops, and
// 3-D uploads whose slice stride is not a whole number of rows.
static_assert(iane::concepts::DeviceView<iane::texture_image_view<float>>);
static_assert(!iane::concepts::MemoryView<iane::texture_image_view<float>>);
static_assert(!iane::concepts::HostView<iane::texture_image_view<float>>);

void texture_upload_from_strided_views()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{2, 2};
	iane::host_image<int, 2> host(coord_t{4, 3});
	for (int i = 0; i < 12; ++i) {
		host.pointer()[i] = i + 101;
	}
	iane::host_image<int, 2> output(size);
	iane::texture_image<int> texture(size);
	iane::device_image<int, 2> linear(host.dimensions());
	iane::device_image<int, 2> materialized(size);
	void* stream = nullptr;
	iane::detail::check_status(iane_hip_stream_create(&stream));
	const iane::hip_policy policy{stream};
	check_throws<std::invalid_argument>([&] { iane::copy(host.const_view(), texture, policy); }, "upload extent mismatch");
	const auto verify = [&](int aOffset) {
		iane::copy(texture.const_view(), materialized.view(), policy);
		iane::copy(materialized.const_view(), output.view(), policy).wait();
		const int expected[]{106, 107, 110, 111};
		for (int i = 0; i < 4; ++i) {
			CHECK(output.pointer()[i] == expected[i] + aOffset);
		}
	};
	const coord_t origin{1, 1};
	iane::copy(host.const_view().subview(origin, size), texture, policy);
	verify(0);
	for (int i = 0; i < 12; ++i) {
		host.pointer()[i] += 100;
	}
	iane::copy(host.const_view(), linear.view(), policy);
	iane::copy(linear.const_view().subview(origin, size), texture, policy);
	verify(100);
	iane::detail::check_status(iane_hip_stream_destroy(stream));
}

void texture_convolution_matches_cpu()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{5, 4};
	const coord_t crop_size{3, 3};
	const coord_t origin{1, 1};
	iane::host_image<float, 2> source(size), kernel(coord_t{3, 3}), reference(crop_size), output(crop_size);
	for (int i = 0; i < 20; ++i) {
		source.pointer()[i] = static_cast<float>(i - 6);
	}
	for (int i = 0; i < 9; ++i) {
		kernel.pointer()[i] = static_cast<float>(i - 3) / 8.0f;
	}
	iane::texture_image<float> texture(size);
	iane::device_image<float, 2> device_kernel(kernel.dimensions()), device_output(crop_size);
	iane::copy(source.const_view(), texture).wait();
	iane::copy(kernel.const_view(), device_kernel.view()).wait();
	for (const auto border : {iane::border_mode::clamp, iane::border_mode::zero}) {
		iane::convolve(source.const_view().subview(origin, crop_size), kernel.const_view(), reference.view(),
			{.border = border});
		iane::convolve(texture.const_view().subview(origin, crop_size), device_kernel.const_view(), device_output.view(),
			{.border = border});
		iane::copy(device_output.const_view(), output.view()).wait();
		for (int i = 0; i < 9; ++i) {
			CHECK(output.pointer()[i] == reference.pointer()[i]);
		}
	}
}

void empty_texture_and_move()
{
	using coord_t = iane::math::vec2i;
	const coord_t size{0, 3};
	iane::texture_image<float> empty(size);
	auto moved = std::move(empty);
	iane::host_image<float, 2> source(size);
	iane::copy(source.const_view(), moved).wait();
	CHECK(moved.const_view().dimensions()[0] == size[0]);
	CHECK(moved.const_view().dimensions()[1] == size[1]);
	CHECK(moved.array() == nullptr);
	check_throws<std::invalid_argument>([] { iane::texture_image<float> bad(coord_t{-1, 2}); }, "negative extent");
}

void texture_1d_nested_crop()
{
	iane::host_image<int, 1> source(6);
	for (int i = 0; i < 6; ++i) {
		source.view()[i] = 100 + i;
	}
	iane::texture_image<int, 1> texture(4);
	iane::device_image<int, 1> materialized(2);
	iane::host_image<int, 1> output(2);
	iane::copy(source.const_view().subview(1, 4), texture);
	iane::copy(texture.const_view().subview(1, 3).subview(1, 2), materialized.view());
	iane::copy(materialized.const_view(), output.view()).wait();
	CHECK(output.view()[0] == 103);
	CHECK(output.view()[1] == 104);
}

void texture_3d_upload_from_strided_views()
{
	using coord_t = iane::math::vec3i;
	const coord_t size{3, 2, 3};
	const coord_t crop_size{2, 2, 2};
	const coord_t origin{1, 0, 1};
	iane::texture_image<int, 3> texture(size);
	iane::device_image<int, 3> materialized(crop_size);
	iane::host_image<int, 3> output(crop_size);
	const auto verify = [&](int aOffset) {
		iane::copy(texture.const_view().subview(origin, crop_size), materialized.view());
		iane::copy(materialized.const_view(), output.view()).wait();
		for (int z = 0; z < 2; ++z) {
			for (int y = 0; y < 2; ++y) {
				for (int x = 0; x < 2; ++x) {
					CHECK((output.view()[coord_t{x, y, z}] == 100 * (z + 1) + 10 * y + x + 1 + aOffset));
				}
			}
		}
	};
	// a slice stride (44 bytes) that is not a whole number of row pitches (16 bytes)
	int padded[33]{};
	const iane::host_image_view<int, 3> host(padded, size, coord_t{4, 16, 44});
	for (int z = 0; z < 3; ++z) {
		for (int y = 0; y < 2; ++y) {
			for (int x = 0; x < 3; ++x) {
				host[coord_t{x, y, z}] = 100 * z + 10 * y + x;
			}
		}
	}
	iane::copy(host, texture);
	verify(0);
	// a cropped device view as the source
	const coord_t storage_size{5, 4, 4};
	const coord_t storage_origin{1, 1, 1};
	iane::host_image<int, 3> packed(storage_size);
	for (int i = 0; i < 5 * 4 * 4; ++i) {
		packed.pointer()[i] = -1;
	}
	auto crop = packed.view().subview(storage_origin, size);
	for (int z = 0; z < 3; ++z) {
		for (int y = 0; y < 2; ++y) {
			for (int x = 0; x < 3; ++x) {
				crop[coord_t{x, y, z}] = host[coord_t{x, y, z}] + 1000;
			}
		}
	}
	iane::device_image<int, 3> linear(storage_size);
	iane::copy(packed.const_view(), linear.view());
	iane::copy(linear.const_view().subview(storage_origin, size), texture);
	verify(1000);
	// rows must be contiguous
	const iane::host_image_view<int, 3> strided_rows(padded, coord_t{2, 2, 2}, coord_t{8, 16, 44});
	iane::texture_image<int, 3> small(coord_t{2, 2, 2});
	check_throws<std::invalid_argument>([&] { iane::copy(strided_rows, small); }, "non-contiguous rows");
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

} // namespace

int main(int argc, char** argv)
{
	const bool device = argc > 1 && std::strcmp(argv[1], "device") == 0;
	if (!device) {
		pattern_host<2>();
		pattern_host<3>();
		pattern_host<4>();
		scalar_coordinates();
		crop_flip(false);
		policy_validation();
		sequence_case<false>();
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
		texture_upload_from_strided_views();
		texture_convolution_matches_cpu();
		empty_texture_and_move();
		texture_1d_nested_crop();
		texture_3d_upload_from_strided_views();
		sequence_case<true>();
	}
	std::printf("%s: %d checks, %d failures\n", device ? "device" : "host", g_checks, g_failures);
	return g_failures ? 1 : 0;
}
