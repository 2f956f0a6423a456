// Richardson–Lucy and Wiener deconvolution for device images.
//
// prepare_rl mirrors examples/rl_deconvolution/main.cu:51-81 of the reference:
// it prepares the sensitivity (H^T 1, main.cu:62-63) and returns a callable
// that runs `iterations` RL steps from max(y, eps) (main.cu:68) and returns a
// completion handle. The work images live in the prepared state (main.cu:57-60).
// prepare_wiener mirrors examples/wiener/main.cu:44-83 on the FFT drop-ins.
#pragma once

#include <iane/prepared.hpp>

#include <memory>

namespace iane {

enum class rl_method { direct = IANE_RL_DIRECT, fft = IANE_RL_FFT };

/// A prepared Richardson–Lucy problem for one image shape and PSF (rank 2 or 3, f32).
template<int N>
class rl_plan {
	static_assert(N == 2 || N == 3, "RL plans support rank 2 and 3 images");

public:
	using extents_t = typename math::dim_traits<N>::extents_t;

	template<class TPsf>
	rl_plan(extents_t aSize, TPsf aPsf, rl_method aMethod = rl_method::fft,
		convolution_precision aPrecision = convolution_precision::fp32, float aEpsilon = 1e-6f, hip_policy aPolicy = {})
		: mSize(aSize), mPolicy(aPolicy)
	{
		static_assert(concepts::DeviceView<TPsf> && TPsf::dimension == N);
		int32_t extent[3] = {1, 1, 1};
		for (int i = 0; i < N; ++i) {
			extent[i] = aSize[i];
		}
		const auto psf = detail::describe(aPsf);
		iane_rl_plan* raw = nullptr;
		detail::check_status(iane_hip_rl_plan_create(&raw, N, extent, &psf, static_cast<int>(aMethod),
			static_cast<int>(aPrecision), aEpsilon, aPolicy.stream));
		mPlan.reset(raw);
		std::size_t bytes = 0;
		detail::check_status(iane_hip_rl_plan_workspace_bytes(raw, &bytes));
		mScratch.resize(static_cast<int>((bytes + 7) / 8));
	}

	/// Run `iterations` steps; `init` starts from max(y, eps), otherwise from the estimate.
	template<class TObs, class TEst>
	hip_execution run(TObs aObserved, TEst aEstimate, int aIterations, bool aInit = true, double* aLlHistory = nullptr)
	{
		static_assert(std::same_as<typename TObs::value_type, float> && std::same_as<typename TEst::value_type, float>);
		detail::check_packed_fft_view(aObserved);
		detail::check_packed_fft_view(aEstimate);
		if (!(aObserved.dimensions() == mSize) || !(aEstimate.dimensions() == mSize)) {
			throw std::invalid_argument("Image dimensions must match the prepared shape");
		}
		detail::check_status(iane_hip_rl_run(mPlan.get(), aObserved.pointer(), aEstimate.pointer(), aIterations,
			aInit ? 1 : 0, mScratch.pointer(), aLlHistory, mPolicy.stream));
		return hip_execution{mPolicy.stream};
	}

private:
	struct plan_deleter {
		void operator()(iane_rl_plan* aPlan) const noexcept { iane_hip_rl_plan_destroy(aPlan); }
	};
	extents_t mSize;
	hip_policy mPolicy;
	std::unique_ptr<iane_rl_plan, plan_deleter> mPlan;
	device_image<double, 1> mScratch; // the plan's per-run workspace
};

/**
 * prepare_rl (main.cu:51-81) for device views: returns a callable that runs
 * the whole deconvolution (initial estimate + `iterations` steps) into the
 * destination and returns its completion handle.
 */
template<class TSrc, class TPsf, class TDst>
auto prepare_rl(TSrc aSource, TPsf aPsf, TDst aDestination, hip_policy aPolicy, int aIterations, float aEpsilon,
	rl_method aMethod = rl_method::fft, convolution_precision aPrecision = convolution_precision::fp32)
{
	if (aIterations <= 0 || !std::isfinite(aEpsilon) || aEpsilon <= 0.0f) {
		throw std::invalid_argument("Iterations and finite epsilon must be positive");
	}
	auto plan = std::make_shared<rl_plan<TDst::dimension>>(aDestination.dimensions(), aPsf, aMethod, aPrecision,
		aEpsilon, aPolicy);
	return [plan, aSource, aDestination, aIterations]() { return plan->run(aSource, aDestination, aIterations, true); };
}

/// wiener/main.cu:21-32 on device spectra (in place)
template<class TSpectrum, class TResponse>
hip_execution wiener_apply(TSpectrum aSpectrum, TResponse aResponse, float aNoise, hip_policy aPolicy = {})
{
	const auto count = math::detail::element_count<TSpectrum::dimension>(aSpectrum.dimensions());
	detail::check_status(iane_hip_wiener_apply(aSpectrum.pointer(), aResponse.pointer(),
		static_cast<int64_t>(count), aNoise, aPolicy.stream));
	return hip_execution{aPolicy.stream};
}

} // namespace iane
