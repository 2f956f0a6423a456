// convolution: the reference's examples/convolution/main.cu on this backend.
//
//   convolution INPUT OUTPUT --kernel kernel.txt [--roi X Y WIDTH HEIGHT]
//               [--border clamp|zero] [--backend cpu|hip] [--device N]
//               [--warmup N] [--runs N] [--config settings.json]
//
// A grey image convolves as a 2-D u8 view with the float kernel. A colour image
// convolves as one 3-D u8 view over its interleaved pixels (axis 0 = channel,
// 1 = x, 2 = y) with the kernel given extent 1 along the channel axis: each
// channel accumulates in double precision on its own and is rounded to a byte
// once (ties up, saturating), which is what the reference's rgb_sum does
// (main.cu:17-56). The crop is the boundary of the operation; the output has the
// crop's size. Timing lines come from iane::run_on.
#include "../common/image_io.hpp"
#include "../common/kernel.hpp"
#include "../common/program.hpp"

#include <tuple>
#include <type_traits>

int main(int aArgc, char** aArgv)
{
	return cli::run(aArgc, aArgv, "convolution",
		{{"kernel", "Text file containing kernel coefficients", std::nullopt, true},
			{"roi", "Crop as X Y WIDTH HEIGHT", std::nullopt, false, true},
			{"border", "clamp or zero", "clamp"}},
		[](const std::string& aInputPath, const std::string& aOutputPath, const iane::run_options& aExecution,
			const cli::arguments& aArgs) {
			const auto border = aArgs.text("border");
			if (border != "clamp" && border != "zero") {
				throw std::invalid_argument("Border must be clamp or zero");
			}
			const auto input = io::read_image8(aInputPath);
			const auto kernel = io::read_kernel(aArgs.text("kernel"));
			const iane::convolution_options options{
				.border = border == "clamp" ? iane::border_mode::clamp : iane::border_mode::zero,
			};
			std::visit([&](const auto& aImage) {
				using image_t = std::decay_t<decltype(aImage)>;
				const auto size = aImage.dimensions();
				int x0 = 0, y0 = 0, width = size[0], height = size[1];
				if (aArgs.has("roi")) {
					const auto roi = aArgs.integers("roi");
					if (roi.size() != 4 || roi[0] < 0 || roi[1] < 0 || roi[2] <= 0 || roi[3] <= 0 ||
						roi[2] > size[0] - roi[0] || roi[3] > size[1] - roi[1]) {
						throw std::invalid_argument("ROI must be a nonempty rectangle inside the input");
					}
					x0 = roi[0];
					y0 = roi[1];
					width = roi[2];
					height = roi[3];
				}
				const auto convolve = [&](auto aSrc, auto aCoefficients, auto aDst, auto& aExec) {
					return iane::convolve(aSrc, aCoefficients, aDst, options, aExec);
				};
				if constexpr (std::is_same_v<image_t, io::gray_image>) {
					using coord_t = iane::math::vec2i;
					const auto source = aImage.const_view().subview(coord_t{x0, y0}, coord_t{width, height});
					io::gray_image output(coord_t{width, height});
					iane::run_on(aExecution, std::tuple{source, kernel.const_view()}, output.view(), convolve);
					io::write_gray8(aOutputPath, output);
				} else {
					using coord_t = iane::math::vec3i;
					const auto* pixels = reinterpret_cast<const std::uint8_t*>(aImage.pointer());
					const iane::const_host_image_view<std::uint8_t, 3> channels(pixels, coord_t{3, size[0], size[1]},
						coord_t{1, 3, 3 * size[0]});
					const auto source = channels.subview(coord_t{0, x0, y0}, coord_t{3, width, height});
					const auto ksize = kernel.dimensions();
					iane::host_image<float, 3> kernel3(coord_t{1, ksize[0], ksize[1]});
					std::copy_n(kernel.pointer(), static_cast<std::size_t>(ksize[0]) * ksize[1], kernel3.pointer());
					io::rgb_image output(iane::math::vec2i{width, height});
					const iane::host_image_view<std::uint8_t, 3> destination(
						reinterpret_cast<std::uint8_t*>(output.pointer()), coord_t{3, width, height},
						coord_t{1, 3, 3 * width});
					iane::run_on(aExecution, std::tuple{source, kernel3.const_view()}, destination, convolve);
					io::write_rgb8(aOutputPath, output);
				}
			}, input);
		});
}
