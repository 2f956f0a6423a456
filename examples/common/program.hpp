// The shared shell of the example programs (the reference's cli::run,
// examples/common/cli.hpp:28-91): common options, validation, run options for
// iane::run_on, and "NAME: message" on stderr with exit status 1 on error.
//
// Backends: "cpu" (host path of the headers) and "hip" (libiane_hip.so on the
// GPU); "cuda" is accepted as another name for the GPU backend so scripts
// written for the reference keep working.
#pragma once

#include "options.hpp"

#include <iane/iane.hpp>

namespace cli {

template<bool tHasInput = true, class TOperation>
int run(int aArgc, char** aArgv, const std::string& aName, std::vector<option> aOperationOptions, TOperation aOperation)
{
	try {
		std::vector<option> options = {
			{"help", "Show help", std::nullopt, false, false, true},
			{"config", "Read options from a JSON file"},
			{"output", "Output image path (.png, .bmp, .tga, .pgm or .ppm)", std::nullopt, true},
			{"backend", "cpu or hip (cuda is accepted for hip)", "cpu"},
			{"device", "GPU device index", "0"},
			{"warmup", "Number of warm-up runs", "2"},
			{"runs", "Number of measured runs", "10"},
		};
		if constexpr (tHasInput) {
			options.push_back({"input", "Input image path (format detected from contents)", std::nullopt, true});
		}
		options.insert(options.end(), aOperationOptions.begin(), aOperationOptions.end());
		arguments args(options);
		if (!args.parse(aArgc, aArgv, aName + (tHasInput ? " INPUT OUTPUT [options]" : " OUTPUT [options]"),
				tHasInput ? 2 : 1)) {
			return 0;
		}
		const auto backend = args.text("backend");
		const int device = args.integer("device");
		const int warmup = args.integer("warmup");
		const int runs = args.integer("runs");
		if (backend != "cpu" && backend != "hip" && backend != "cuda") {
			throw std::invalid_argument("Backend must be cpu or hip");
		}
		if (device < 0 || warmup < 0 || runs <= 0) {
			throw std::invalid_argument("Device and warmup must be nonnegative and runs must be positive");
		}
		const iane::run_options execution{
			.backend = backend == "cpu" ? iane::backend::cpu : iane::backend::hip,
			.device = device,
			.warmup = warmup,
			.runs = runs,
		};
		if constexpr (tHasInput) {
			aOperation(args.text("input"), args.text("output"), execution, args);
		} else {
			aOperation(args.text("output"), execution, args);
		}
		return 0;
	} catch (const std::exception& error) {
		std::cerr << aName << ": " << error.what() << '\n';
		return 1;
	}
}

} // namespace cli
