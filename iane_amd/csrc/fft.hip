// FFT plans behind iane_hip_fft_*: the drop-in for the reference's prepared
// fft_forward / fft_inverse / rfft / irfft (include/iane/algorithms/fft/fft.hpp:196-231,
// rfft.hpp:107-161), generalised from 2-D to ranks 1..3.
//
// Conventions (fft.hpp:228-231, docs/prepared_algorithms.md:26-29): forward uses
// the negative exponent, the inverse is unnormalised, R2C writes the packed
// half spectrum with extent {n0/2+1, n1, ...} along axis 0 (layout.hpp:15-21),
// C2R leaves its input untouched (rfft.hpp:72-83: the reference copies it into
// scratch; here the first inverse pass reads it and writes the workspace).
//
// A plan is a list of passes, one per axis; each pass is one launch of
// fft_pass_kernel (lines of the axis held in LDS, all radix stages in one kernel).
#include "fft.hpp"

#include <atomic>
#include <complex>
#include <map>
#include <tuple>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace iane_hip {

bool factor_radices(int aM, fft_radix_plan& aPlan)
{
	aPlan.m = aM;
	aPlan.stages = 0;
	int rest = aM;
	auto push = [&](int r) {
		if (aPlan.stages >= kFftMaxStages) {
			return false;
		}
		aPlan.radix[aPlan.stages++] = r;
		rest /= r;
		return true;
	};
	// odd radices first: the first stage writes its outputs R elements apart,
	// which an odd R spreads over distinct LDS banks
	for (int r : {3, 5, 7, 11, 13}) {
		while (rest % r == 0) {
			if (!push(r)) return false;
		}
	}
	for (int r = 17; r <= 64; r += 2) {
		while (rest % r == 0) {
			if (!push(r)) return false;
		}
	}
	while (rest % 8 == 0) {
		if (!push(8)) return false;
	}
	while (rest % 4 == 0) {
		if (!push(4)) return false;
	}
	while (rest % 2 == 0) {
		if (!push(2)) return false;
	}
	if (rest != 1) {
		return false;
	}
	int ns = 1;
	for (int s = 0; s < aPlan.stages; ++s) {
		const int r = aPlan.radix[s];
		fft_stage& st = aPlan.stage[s];
		st.nb = aM / r;
		st.ns = ns;
		st.step = aM / (ns * r);
		st.c_nb = fast_div(st.nb).c;
		st.c_ns = fast_div(ns).c;
		ns *= r;
	}
	return true;
}

namespace {
std::mutex g_twiddle_mutex;
std::map<std::pair<int, int>, float2*> g_twiddles;
} // namespace

const float2* twiddle_table(int aM)
{
	int device = 0;
	check_hip(hipGetDevice(&device), "Query HIP device");
	std::lock_guard lock(g_twiddle_mutex);
	auto& slot = g_twiddles[{device, aM}];
	if (!slot) {
		std::vector<float2> host(aM);
		for (int k = 0; k < aM; ++k) {
			const double angle = -2.0 * 3.14159265358979323846 * static_cast<double>(k) / aM;
			host[k] = make_float2(static_cast<float>(std::cos(angle)), static_cast<float>(std::sin(angle)));
		}
		float2* table = nullptr;
		check_hip(hipMalloc(&table, sizeof(float2) * aM), "Allocate FFT twiddles");
		check_hip(hipMemcpy(table, host.data(), sizeof(float2) * aM, hipMemcpyHostToDevice), "Upload FFT twiddles");
		slot = table;
	}
	return slot;
}

int fixed_plan_id(const fft_radix_plan& aPlan)
{
	if (tune_get("fft.fixed", 1) == 0) {
		return 0;
	}
	const auto is = [&](std::initializer_list<int> aRadices) {
		if (static_cast<int>(aRadices.size()) != aPlan.stages) {
			return false;
		}
		int s = 0;
		for (int r : aRadices) {
			if (aPlan.radix[s++] != r) {
				return false;
			}
		}
		return true;
	};
	static const std::initializer_list<int> table[kFixedPlans] = {
		{5, 7, 4}, {5, 7, 8}, {5, 8, 4}, {8, 8, 2}, {8, 8, 4}, {8, 8, 8}, {8, 8, 8, 2}};
	for (int id = 1; id <= kFixedPlans; ++id) {
		if (is(table[id - 1])) {
			return id;
		}
	}
	return 0;
}

int next_fft_length(int aN)
{
	for (int n = aN + (aN & 1); ; n += 2) {
		int r = n;
		for (int p : {2, 3, 5, 7}) {
			while (r % p == 0) {
				r /= p;
			}
		}
		if (r == 1) {
			return n;
		}
	}
}

namespace {

template<int MODE, bool GENERAL, int FIXED = 0>
__global__ __launch_bounds__(kFftThreads) void fft_pass_kernel(fft_pass_args a)
{
	extern __shared__ float2 lds[];
	fft_pass_body<MODE, GENERAL, FIXED>(a, xcd_remap(blockIdx.x, gridDim.x), lds);
}

template<int MODE>
void launch_staged(const fft_pass_args& a, unsigned aBlocks, size_t aLds, hipStream_t aStream)
{
	with_fixed_plan(a.fixed, [&](auto aId) {
		fft_pass_kernel<MODE, false, decltype(aId)::value><<<aBlocks, kFftThreads, aLds, aStream>>>(a);
	});
}

} // namespace

// LDS bytes one workgroup may use on the current device (cached per device).
size_t max_lds_per_block()
{
	int device = 0;
	check_hip(hipGetDevice(&device), "hipGetDevice");
	static std::atomic<int> cache[64];
	if (device >= 0 && device < 64) {
		const int cached = cache[device].load(std::memory_order_relaxed);
		if (cached > 0) {
			return static_cast<size_t>(cached);
		}
	}
	int bytes = 0;
	check_hip(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "hipDeviceGetAttribute");
	bytes = std::max(bytes, 1);
	if (device >= 0 && device < 64) {
		cache[device].store(bytes, std::memory_order_relaxed);
	}
	return static_cast<size_t>(bytes);
}

namespace {

// Forward DFT in double of a length whose prime factors are small (mixed-radix
// decimation in time; naive DFT for a prime length). Host only: the Bluestein
// filter spectrum.
void host_dft(std::vector<std::complex<double>>& aX)
{
	const size_t n = aX.size();
	if (n <= 1) {
		return;
	}
	size_t p = 2;
	while (n % p != 0) {
		++p;
	}
	const double pi = 3.14159265358979323846;
	std::vector<std::complex<double>> out(n);
	if (p == n) {
		for (size_t k = 0; k < n; ++k) {
			std::complex<double> sum = 0.0;
			for (size_t j = 0; j < n; ++j) {
				sum += aX[j] * std::polar(1.0, -2.0 * pi * static_cast<double>((j * k) % n) / static_cast<double>(n));
			}
			out[k] = sum;
		}
		aX.swap(out);
		return;
	}
	const size_t q = n / p;
	std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(q));
	for (size_t r = 0; r < p; ++r) {
		for (size_t j = 0; j < q; ++j) {
			sub[r][j] = aX[r + p * j];
		}
		host_dft(sub[r]);
	}
	for (size_t k = 0; k < n; ++k) {
		std::complex<double> sum = 0.0;
		for (size_t r = 0; r < p; ++r) {
			sum += sub[r][k % q] * std::polar(1.0, -2.0 * pi * static_cast<double>((r * k) % n) / static_cast<double>(n));
		}
		out[k] = sum;
	}
	aX.swap(out);
}

struct bluestein_tables {
	float2* chirp = nullptr;
	float2* spectrum = nullptr;
};

std::mutex g_bluestein_mutex;
std::map<std::tuple<int, int, int>, bluestein_tables> g_bluestein;

// Chirp exp(-i pi j^2 / n) (j < n; the angle reduced exactly with j^2 mod 2n)
// and the spectrum of the length-aBig filter b[j] = b[aBig - j] = conj(chirp[j])
// (j < n, zero between), divided by aBig. Cached per (device, n, aBig).
bluestein_tables bluestein(int aN, int aBig)
{
	int device = 0;
	check_hip(hipGetDevice(&device), "Query HIP device");
	std::lock_guard lock(g_bluestein_mutex);
	auto& slot = g_bluestein[{device, aN, aBig}];
	if (!slot.chirp) {
		const double pi = 3.14159265358979323846;
		std::vector<std::complex<double>> chirp(aN);
		for (long long j = 0; j < aN; ++j) {
			const long long r = (j * j) % (2LL * aN);
			chirp[j] = std::polar(1.0, -pi * static_cast<double>(r) / static_cast<double>(aN));
		}
		std::vector<std::complex<double>> filter(aBig, 0.0);
		for (int j = 0; j < aN; ++j) {
			filter[j] = std::conj(chirp[j]);
			if (j > 0) {
				filter[aBig - j] = std::conj(chirp[j]);
			}
		}
		host_dft(filter);
		std::vector<float2> c(aN), f(aBig);
		for (int j = 0; j < aN; ++j) {
			c[j] = make_float2(static_cast<float>(chirp[j].real()), static_cast<float>(chirp[j].imag()));
		}
		for (int k = 0; k < aBig; ++k) {
			const std::complex<double> v = filter[k] / static_cast<double>(aBig);
			f[k] = make_float2(static_cast<float>(v.real()), static_cast<float>(v.imag()));
		}
		bluestein_tables t;
		check_hip(hipMalloc(&t.chirp, sizeof(float2) * aN), "Allocate Bluestein chirp");
		check_hip(hipMalloc(&t.spectrum, sizeof(float2) * aBig), "Allocate Bluestein spectrum");
		check_hip(hipMemcpy(t.chirp, c.data(), sizeof(float2) * aN, hipMemcpyHostToDevice), "Upload Bluestein chirp");
		check_hip(hipMemcpy(t.spectrum, f.data(), sizeof(float2) * aBig, hipMemcpyHostToDevice),
			"Upload Bluestein spectrum");
		slot = t;
	}
	return slot;
}

// Complex elements of LDS one workgroup may use (1 KiB kept for the kernels'
// static LDS: batched descriptors, reduction slots).
long long lds_elems_limit()
{
	return (static_cast<long long>(max_lds_per_block()) - 1024) / static_cast<long long>(sizeof(float2));
}

} // namespace

bool fft_line_fits(int aM)
{
	fft_radix_plan plan;
	return factor_radices(aM, plan) && 2LL * (aM + 1) <= lds_elems_limit();
}

namespace {

// 2: the line's own stages fit in LDS; 1: Bluestein's form fits; 0: neither.
int line_feasibility(int aM)
{
	if (fft_line_fits(aM)) {
		return 2;
	}
	return aM > 1 && fft_line_fits(next_fft_length(2 * aM - 1)) ? 1 : 0;
}

} // namespace

// Fill the radix plan, tables, LDS pitch and bundle of a pass for a transform
// of logical length n along lines of `lines` x `outer`. A complex line length m
// whose radices are supported runs its own stages; otherwise Bluestein's form
// with the smallest transform length M >= 2m - 1 of radices 2..7. Both hold a
// line (M + 1 points for Bluestein) twice in LDS: lines longer than that are
// refused here (the plan splits such axes into shorter passes first).
void setup_pass(fft_pass_args& a, int aMode, int aN, int aLines, int aOuter)
{
	a.mode = aMode;
	a.n = aN;
	const bool half = (aMode == PASS_R2C || aMode == PASS_C2R) && (aN % 2 == 0);
	const int m = half ? aN / 2 : aN;
	a.line_m = m;
	a.chirp = nullptr;
	a.bspec = nullptr;
	const long long limit = lds_elems_limit();
	if (!factor_radices(m, a.plan) || 2LL * (m + 1) > limit) {
		if (aMode == PASS_C2C_MULT) {
			fail(IANE_ERR_UNSUPPORTED, "FFT length " + std::to_string(aN) + " needs radices 2..64 in a spectrum pass");
		}
		const int big = next_fft_length(2 * m - 1);
		if (!factor_radices(big, a.plan) || 2LL * (big + 1) > limit) {
			fail(IANE_ERR_UNSUPPORTED, "FFT line length " + std::to_string(aN) + " exceeds the LDS line limit");
		}
		const bluestein_tables t = bluestein(m, big);
		a.chirp = t.chirp;
		a.bspec = t.spectrum;
	}
	const int stages = a.plan.m;
	a.pitch = stages + 1;
	a.tw_m = twiddle_table(stages);
	a.tw_n = half ? twiddle_table(aN) : nullptr;
	a.lines = aLines;
	a.outer = aOuter;
	// the twiddle tables follow the two line buffers in LDS when one line of
	// each buffer and the tables fit; longer lines and Bluestein's form run the
	// general pass kernel, which reads them from memory
	const int tables = pass_table_elems(stages, m, half);
	a.tables_in_lds = a.chirp == nullptr && 2LL * a.pitch + tables <= limit ? 1 : 0;
	a.fixed = a.tables_in_lds ? fixed_plan_id(a.plan) : 0;
	const int staged = a.tables_in_lds ? tables : 0;
	// Lines per workgroup: as many as fit an LDS budget of "fft.lds_elems" complex
	// elements per buffer (autotuned; default kFftLdsElems). A smaller budget means
	// less LDS per workgroup and more resident workgroups per CU. Every line is
	// transformed the same way whatever the bundle, so results do not depend on it.
	const int budget = std::max(a.pitch, std::min(tune_get("fft.lds_elems", kFftLdsElems), kFftLdsElems));
	// the twiddle tables come out of the same budget of 2 * budget elements, so
	// the workgroups resident per CU do not change with them
	a.bundle = std::max(1, std::min(aLines, (2 * budget - staged) / (2 * a.pitch)));
	// fewer lines if the workgroup would exceed the device's LDS
	while (a.bundle > 1 && 2LL * a.bundle * a.pitch + staged > limit) {
		--a.bundle;
	}
	a.lds_elems = a.bundle * a.pitch;
	a.groups_per_outer = ceil_div(aLines, a.bundle);
}

void launch_pass(const fft_pass_args& a, hipStream_t aStream)
{
	const long long blocks = static_cast<long long>(a.groups_per_outer) * a.outer;
	if (blocks == 0) {
		return;
	}
	if (blocks > 0x7fffffffLL) {
		fail(IANE_ERR_UNSUPPORTED, "FFT grid too large");
	}
	const unsigned nb = static_cast<unsigned>(blocks);
	const size_t lds = pass_lds_bytes(a);
	if (!a.tables_in_lds) {
		switch (a.mode) {
		case PASS_C2C: fft_pass_kernel<PASS_C2C, true><<<nb, kFftThreads, lds, aStream>>>(a); break;
		case PASS_R2C: fft_pass_kernel<PASS_R2C, true><<<nb, kFftThreads, lds, aStream>>>(a); break;
		case PASS_C2R: fft_pass_kernel<PASS_C2R, true><<<nb, kFftThreads, lds, aStream>>>(a); break;
		default: fail(IANE_ERR_UNSUPPORTED, "FFT spectrum pass line too long for LDS");
		}
	} else {
		switch (a.mode) {
		case PASS_C2C: launch_staged<PASS_C2C>(a, nb, lds, aStream); break;
		case PASS_R2C: launch_staged<PASS_R2C>(a, nb, lds, aStream); break;
		case PASS_C2R: launch_staged<PASS_C2R>(a, nb, lds, aStream); break;
		default: launch_staged<PASS_C2C_MULT>(a, nb, lds, aStream); break;
		}
	}
	check_hip(hipGetLastError(), "FFT pass launch");
}

// Geometry of the lines along `axis` of a packed array with extents e[0..rank-1].
void axis_lines(const int* e, int rank, int axis, line_geom& g, int& lines, int& outer)
{
	long long inner = 1;
	for (int i = 0; i < axis; ++i) {
		inner *= e[i];
	}
	long long above = 1;
	for (int i = axis + 1; i < rank; ++i) {
		above *= e[i];
	}
	if (axis == 0) {
		g = {1, e[0], 0};
		lines = static_cast<int>(above);
		outer = 1;
	} else {
		g = {inner, 1, inner * e[axis]};
		lines = static_cast<int>(inner);
		outer = static_cast<int>(above);
	}
}

} // namespace iane_hip

using namespace iane_hip;

enum { BUF_IN = 0, BUF_OUT = 1, BUF_WS = 2, BUF_WS2 = 3 };

struct iane_fft_plan {
	int rank = 0;
	int kind = 0;
	int extent[3] = {1, 1, 1};      // logical (real for R2C/C2R) extents
	size_t in_bytes = 0;
	size_t out_bytes = 0;
	size_t workspace = 0;
	size_t ws2_offset = 0;                    // second workspace region (four-step passes of a C2R plan)
	std::vector<fft_pass_args> passes;
	std::vector<std::pair<int, int>> buffers; // (in, out) per pass
};

namespace {

// The complex passes along `axis` of an array with extents e[] (packed spectrum
// layout) from buffer aSrc to buffer aDst. A line the LDS holds (directly or in
// Bluestein's form) is one pass. A longer line n = n1 n2 (both factors such
// lines, preferring radix factors and balanced sizes) is split in four steps:
// with j = j2 + n2 j1 and k = k1 + n1 k2,
//   pass A: length-n1 transforms over j1 (points n2 apart) in the layout of
//           the source, each result times W_n^(j2 k1), into aScratch;
//   pass B: length-n2 transforms over j2, stored to X[k1 + n1 k2] in aDst.
// Returns false when no factorisation fits.
bool add_c2c_axis(iane_fft_plan& aPlan, const int* e, int aRank, int aAxis, float aSign, int aSrc, int aDst,
	int aScratch)
{
	line_geom g;
	int lines, outer;
	axis_lines(e, aRank, aAxis, g, lines, outer);
	const int n = e[aAxis];
	if (line_feasibility(n) > 0) {
		fft_pass_args a{};
		setup_pass(a, PASS_C2C, n, lines, outer);
		a.sign = aSign;
		a.load_n = a.store_n = n;
		a.gin = a.gout = g;
		aPlan.passes.push_back(a);
		aPlan.buffers.push_back({aSrc, aDst});
		return true;
	}
	int best_n1 = 0;
	int best_score = -1;
	for (int d = 2; static_cast<long long>(d) * d <= n; ++d) {
		if (n % d != 0) {
			continue;
		}
		for (const int n1 : {d, n / d}) {
			const int f1 = line_feasibility(n1);
			const int f2 = line_feasibility(n / n1);
			if (f1 == 0 || f2 == 0) {
				continue;
			}
			const int score = (f1 + f2) * (1 << 24) + std::min(n1, n / n1);
			if (score > best_score) {
				best_score = score;
				best_n1 = n1;
			}
		}
	}
	if (best_n1 == 0) {
		return false;
	}
	const int n1 = best_n1;
	const int n2 = n / n1;
	long long inner = 1;
	for (int i = 0; i < aAxis; ++i) {
		inner *= e[i];
	}
	long long above = 1;
	for (int i = aAxis + 1; i < aRank; ++i) {
		above *= e[i];
	}
	const long long stride = inner * n;  // between the lines' outer groups
	if (inner * n2 > 0x7fffffffLL || above * n1 > 0x7fffffffLL) {
		return false;
	}
	fft_pass_args pa{};
	setup_pass(pa, PASS_C2C, n1, static_cast<int>(inner * n2), static_cast<int>(above));
	pa.sign = aSign;
	pa.load_n = pa.store_n = n1;
	pa.gin = pa.gout = {inner * n2, 1, stride, 0, 0};
	pa.post_n = n;
	pa.post_div = static_cast<int>(inner);
	pa.tw_post = twiddle_table(n);
	pa.tables_in_lds = 0;  // the general pass kernel (the twiddle step is there)
	pa.fixed = 0;
	fft_pass_args pb{};
	setup_pass(pb, PASS_C2C, n2, static_cast<int>(inner), static_cast<int>(above * n1));
	pb.sign = aSign;
	pb.load_n = pb.store_n = n2;
	pb.gin = {inner, 1, inner * n2, 0, 0};
	pb.gout = {inner * n1, 1, inner, stride, n1};
	pb.tables_in_lds = 0;
	pb.fixed = 0;
	aPlan.passes.push_back(pa);
	aPlan.buffers.push_back({aSrc, aScratch});
	aPlan.passes.push_back(pb);
	aPlan.buffers.push_back({aScratch, aDst});
	return true;
}

void refuse_long_line(int aN)
{
	fail(IANE_ERR_UNSUPPORTED, "FFT line length " + std::to_string(aN) + " exceeds the LDS line limit");
}

} // namespace

extern "C" {

int iane_hip_fft_plan_create(iane_fft_plan** aPlan, int aRank, const int32_t* aExtent, int aKind)
{
	return guarded([&] {
		if (!aPlan || !aExtent) {
			invalid("FFT plan output or extent is null");
		}
		*aPlan = nullptr;
		if (aRank < 1 || aRank > 3) {
			fail(IANE_ERR_UNSUPPORTED, "FFT rank must be 1..3");
		}
		if (aKind < IANE_FFT_C2C_FORWARD || aKind > IANE_FFT_C2R) {
			invalid("unknown FFT kind");
		}
		auto plan = std::make_unique<iane_fft_plan>();
		plan->rank = aRank;
		plan->kind = aKind;
		int spec[3] = {1, 1, 1};
		long long count = 1;
		for (int i = 0; i < aRank; ++i) {
			if (aExtent[i] <= 0) {
				invalid(aKind >= IANE_FFT_R2C ? "Real FFT requires positive dimensions" : "FFT requires positive dimensions");
			}
			plan->extent[i] = aExtent[i];
			spec[i] = aExtent[i];
			count *= aExtent[i];
		}
		const bool real = aKind == IANE_FFT_R2C || aKind == IANE_FFT_C2R;
		if (real) {
			spec[0] = aExtent[0] / 2 + 1;
		}
		long long spec_count = 1;
		for (int i = 0; i < aRank; ++i) {
			spec_count *= spec[i];
		}
		const size_t real_bytes = static_cast<size_t>(count) * sizeof(float);
		const size_t spec_bytes = static_cast<size_t>(spec_count) * sizeof(float2);
		bool four_step = false;
		const auto axis_pass = [&](int aAxis, float aSign, int aSrc, int aDst, int aScratch) {
			const size_t before = plan->passes.size();
			if (!add_c2c_axis(*plan, spec, aRank, aAxis, aSign, aSrc, aDst, aScratch)) {
				refuse_long_line(spec[aAxis]);
			}
			four_step = four_step || plan->passes.size() - before > 1;
		};
		if (aKind == IANE_FFT_C2C_FORWARD || aKind == IANE_FFT_C2C_INVERSE) {
			const float sign = aKind == IANE_FFT_C2C_FORWARD ? -1.0f : 1.0f;
			for (int axis = 0; axis < aRank; ++axis) {
				axis_pass(axis, sign, axis == 0 ? BUF_IN : BUF_OUT, BUF_OUT, BUF_WS);
			}
			plan->in_bytes = plan->out_bytes = spec_bytes;
			plan->workspace = four_step ? spec_bytes : 0;
		} else if (aKind == IANE_FFT_R2C) {
			fft_pass_args a{};
			line_geom g;
			int lines, outer;
			axis_lines(plan->extent, aRank, 0, g, lines, outer);
			if (line_feasibility(plan->extent[0] % 2 == 0 ? plan->extent[0] / 2 : plan->extent[0]) == 0) {
				refuse_long_line(plan->extent[0]);  // real lines are not split (yet)
			}
			setup_pass(a, PASS_R2C, plan->extent[0], lines, outer);
			a.load_n = plan->extent[0];
			a.store_n = spec[0];
			a.gin = g;
			a.gout = {1, spec[0], 0};
			plan->passes.push_back(a);
			plan->buffers.push_back({BUF_IN, BUF_OUT});
			for (int axis = 1; axis < aRank; ++axis) {
				axis_pass(axis, -1.0f, BUF_OUT, BUF_OUT, BUF_WS);
			}
			plan->in_bytes = real_bytes;
			plan->out_bytes = spec_bytes;
			plan->workspace = four_step ? spec_bytes : 0;
		} else { // C2R
			int src = BUF_IN;
			for (int axis = aRank - 1; axis >= 1; --axis) {
				axis_pass(axis, 1.0f, src, BUF_WS, BUF_WS2);
				src = BUF_WS;
			}
			fft_pass_args a{};
			line_geom g;
			int lines, outer;
			axis_lines(plan->extent, aRank, 0, g, lines, outer);
			if (line_feasibility(plan->extent[0] % 2 == 0 ? plan->extent[0] / 2 : plan->extent[0]) == 0) {
				refuse_long_line(plan->extent[0]);  // real lines are not split (yet)
			}
			setup_pass(a, PASS_C2R, plan->extent[0], lines, outer);
			a.load_n = spec[0];
			a.store_n = plan->extent[0];
			a.gin = {1, spec[0], 0};
			a.gout = g;
			plan->passes.push_back(a);
			plan->buffers.push_back({src, BUF_OUT});
			plan->workspace = aRank > 1 ? spec_bytes : 0;
			if (four_step) {
				plan->ws2_offset = (spec_bytes + 255) / 256 * 256;
				plan->workspace = plan->ws2_offset + spec_bytes;
			}
			plan->in_bytes = spec_bytes;
			plan->out_bytes = real_bytes;
		}
		*aPlan = plan.release();
	});
}

int iane_hip_fft_plan_workspace_bytes(const iane_fft_plan* aPlan, size_t* aBytes)
{
	return guarded([&] {
		if (!aPlan || !aBytes) {
			invalid("FFT plan or output is null");
		}
		*aBytes = aPlan->workspace;
	});
}

// Validation of fft.hpp:20-66 / rfft.hpp:29-47: non-null, complex-aligned,
// disjoint buffers; the workspace must not overlap either.
int iane_hip_fft_exec(iane_fft_plan* aPlan, const void* aIn, void* aOut, void* aWorkspace, void* aStream)
{
	return guarded([&] {
		if (!aPlan) {
			invalid("FFT plan is null");
		}
		auto addr = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
		if (!aIn || !aOut || addr(aIn) % alignof(float2) || addr(aOut) % alignof(float2)) {
			invalid("FFT requires aligned input and output buffers");
		}
		auto overlap = [&](const void* a, size_t na, const void* b, size_t nb) {
			return addr(a) <= addr(b) ? addr(b) - addr(a) < na : addr(a) - addr(b) < nb;
		};
		if (overlap(aIn, aPlan->in_bytes, aOut, aPlan->out_bytes)) {
			invalid("FFT requires disjoint input and output buffers");
		}
		if (aPlan->workspace) {
			if (!aWorkspace || addr(aWorkspace) % alignof(float2)) {
				fail(IANE_ERR_LENGTH, "Prepared operation exceeded its scratch reservation");
			}
			if (overlap(aWorkspace, aPlan->workspace, aIn, aPlan->in_bytes) ||
				overlap(aWorkspace, aPlan->workspace, aOut, aPlan->out_bytes)) {
				invalid("FFT workspace overlaps an input or output buffer");
			}
		}
		auto stream = static_cast<hipStream_t>(aStream);
		void* buffers[4] = {const_cast<void*>(aIn), aOut, aWorkspace,
			aWorkspace ? static_cast<char*>(aWorkspace) + aPlan->ws2_offset : nullptr};
		for (size_t i = 0; i < aPlan->passes.size(); ++i) {
			fft_pass_args a = aPlan->passes[i];
			a.in = buffers[aPlan->buffers[i].first];
			a.out = buffers[aPlan->buffers[i].second];
			launch_pass(a, stream);
		}
	});
}

int iane_hip_fft_plan_destroy(iane_fft_plan* aPlan)
{
	return guarded([&] { delete aPlan; });
}

} // extern "C"
