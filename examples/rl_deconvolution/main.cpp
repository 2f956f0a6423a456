// rl_deconvolution: the reference's examples/rl_deconvolution/main.cu on this backend.
//
//   rl_deconvolution INPUT OUTPUT --psf psf.txt [--iterations 20] [--epsilon 1e-6]
//                    [--method direct|fft] [--backend cpu|hip] [--device N]
//                    [--warmup N] [--runs N] [--config settings.json]
//
// Each iteration convolves the estimate with the PSF (zero border), divides the
// observed image by it (floored at epsilon), convolves the ratio with the
// flipped PSF and multiplies the estimate by that correction normalised by the
// flipped PSF's response to an all-ones image (main.cu:11-81). The CPU backend
// runs this loop on host views with FP64-accumulated convolutions; the GPU
// backend runs iane::prepare_rl, by default with the direct method in the same
// (reference) arithmetic, so both produce the same bits; --method fft uses the
// spectral method (FP32, for large PSFs). The grey input is read as float
// intensities (an exact conversion), and the final estimate is rounded and
// clipped to 8 bits after the timed runs.
#include "../common/image_io.hpp"
#include "../common/kernel.hpp"
#include "../common/program.hpp"

#include <algorithm>
#include <memory>
#include <tuple>

namespace {

// The reference loop on host views (main.cu:51-81).
struct host_rl_state {
	iane::host_image<float, 2> flipped, sensitivity, predicted, correction;
};

template<class TSrc, class TPsf, class TDst>
auto prepare_host_rl(TSrc aSource, TPsf aPsf, TDst aDestination, int aIterations, float aEpsilon)
{
	const auto size = aDestination.dimensions();
	const auto ksize = aPsf.dimensions();
	auto state = std::make_shared<host_rl_state>(host_rl_state{iane::host_image<float, 2>(ksize),
		iane::host_image<float, 2>(size), iane::host_image<float, 2>(size), iane::host_image<float, 2>(size)});
	// PSF and source views are packed host images here: flip by reversing storage order
	const std::size_t kcount = static_cast<std::size_t>(ksize[0]) * ksize[1];
	std::reverse_copy(aPsf.pointer(), aPsf.pointer() + kcount, state->flipped.pointer());
	constexpr iane::convolution_options zero{.border = iane::border_mode::zero};
	iane::host_image<float, 2> ones(size);
	const std::size_t count = static_cast<std::size_t>(size[0]) * size[1];
	std::fill_n(ones.pointer(), count, 1.0f);
	iane::convolve(ones.const_view(), state->flipped.const_view(), state->sensitivity.view(), zero);
	return [state, aSource, aPsf, aDestination, aIterations, aEpsilon, count, zero]() {
		float* estimate = aDestination.pointer();
		const float* observed = aSource.pointer();
		for (std::size_t i = 0; i < count; ++i) {
			estimate[i] = observed[i] > aEpsilon ? observed[i] : aEpsilon;
		}
		for (int iteration = 0; iteration < aIterations; ++iteration) {
			iane::convolve(aDestination, aPsf, state->predicted.view(), zero);
			float* predicted = state->predicted.pointer();
			for (std::size_t i = 0; i < count; ++i) {
				predicted[i] = observed[i] / (predicted[i] > aEpsilon ? predicted[i] : aEpsilon);
			}
			iane::convolve(state->predicted.const_view(), state->flipped.const_view(), state->correction.view(), zero);
			const float* correction = state->correction.pointer();
			const float* sensitivity = state->sensitivity.pointer();
			for (std::size_t i = 0; i < count; ++i) {
				const float s = sensitivity[i];
				const float factor = s > 0.0f ? correction[i] / (s > aEpsilon ? s : aEpsilon) : 0.0f;
				estimate[i] = estimate[i] * factor;
			}
		}
		return iane::cpu_execution{};
	};
}

} // namespace

int main(int aArgc, char** aArgv)
{
	return cli::run(aArgc, aArgv, "rl_deconvolution",
		{{"psf", "Text file containing the point spread function", std::nullopt, true},
			{"iterations", "Number of iterations per run", "20"},
			{"epsilon", "Positive floor for divisions", "1e-06"},
			{"method", "direct or fft (GPU backend)", "direct"}},
		[](const std::string& aInputPath, const std::string& aOutputPath, const iane::run_options& aExecution,
			const cli::arguments& aArgs) {
			const int iterations = aArgs.integer("iterations");
			const float epsilon = aArgs.real("epsilon");
			if (iterations <= 0 || !std::isfinite(epsilon) || epsilon <= 0.0f) {
				throw std::invalid_argument("Iterations and finite epsilon must be positive");
			}
			const auto method_name = aArgs.text("method");
			if (method_name != "direct" && method_name != "fft") {
				throw std::invalid_argument("Method must be direct or fft");
			}
			const auto method = method_name == "direct" ? iane::rl_method::direct : iane::rl_method::fft;
			const auto precision = method == iane::rl_method::direct ? iane::convolution_precision::reference :
				iane::convolution_precision::fp32;
			const auto input = io::read_gray8(aInputPath);
			const auto psf = io::read_psf(aArgs.text("psf"));
			const auto size = input.dimensions();
			const std::size_t count = static_cast<std::size_t>(size[0]) * size[1];
			iane::host_image<float, 2> observed(size), estimate(size);
			std::transform(input.pointer(), input.pointer() + count, observed.pointer(),
				[](std::uint8_t v) { return static_cast<float>(v); });
			iane::run_prepared_on(aExecution, std::tuple{observed.const_view(), psf.const_view()}, estimate.view(),
				[&](auto aSource, auto aPsf, auto aDestination, auto& aExec) {
					if constexpr (iane::concepts::DeviceView<decltype(aDestination)>) {
						return iane::prepare_rl(aSource, aPsf, aDestination, aExec, iterations, epsilon, method, precision);
					} else {
						return prepare_host_rl(aSource, aPsf, aDestination, iterations, epsilon);
					}
				});
			io::gray_image output(size);
			std::transform(estimate.pointer(), estimate.pointer() + count, output.pointer(), io::to_gray8);
			io::write_gray8(aOutputPath, output);
		});
}
