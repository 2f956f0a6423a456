// Richardson–Lucy and Wiener deconvolution behind iane_hip_rl_* / iane_hip_wiener_apply.
//
// The loop is the reference's (examples/rl_deconvolution/main.cu:51-81):
//   x0 = max(y, eps);  per iteration:
//   p = H x (zero border);  r = y / max(p, eps);  c = H^T r (flipped PSF);
//   x <- x * (s > 0 ? c / max(s, eps) : 0),  s = H^T 1.
// Pointwise steps use the same float operations as the reference functors
// (main.cu:11-49), so the DIRECT method in REFERENCE precision is bit-identical
// to the reference loop.
//
// The FFT method works on a zero-padded domain n_a >= N_a + K_a/2 (smooth
// lengths), which makes the circular product equal to the zero-border linear
// convolution on the image. One convolution is five line passes:
//   R2C along x (rows of the image) -> C2C along y -> [C2C along z, multiply by
//   the PSF spectrum, inverse along z] in one kernel -> inverse along y -> C2R
//   along x. Loads read only the valid (unpadded) extent and stores write only
//   what a later pass reads. The C2R of one convolution and the R2C of the next
//   happen in the same row kernel with the RL pointwise step between them
//   (ratio + log-likelihood, or the multiplicative update), so neither the
//   prediction nor the ratio image ever touches HBM.
#include "fft.hpp"

#include <memory>
#include <mutex>
#include <vector>

namespace iane_hip {

void setup_pass(fft_pass_args& a, int aMode, int aN, int aLines, int aOuter);
void launch_pass(const fft_pass_args& a, hipStream_t aStream);

namespace {

constexpr int kPointThreads = 256;
constexpr int kPartialBlocks = 1024;

// ---------------------------------------------------------------- pointwise --
// main.cu:12-19
__global__ void rl_initial_kernel(const float* y, float* x, long long n, float eps)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float v = y[i];
		x[i] = v > eps ? v : eps;
	}
}

// Log-likelihood sums stay in FP32: a compensated pair (hi, lo) where lo
// carries the rounding error of every addition into hi (Knuth's TwoSum), so
// the sum of millions of terms keeps the resolution of the terms themselves
// (each term y log max(p, eps) - p is formed in FP32). Same size as a double:
// the partial buffers hold one pair per block.
struct ll_sum {
	float hi;
	float lo;
};

__device__ __forceinline__ ll_sum ll_add(ll_sum aSum, float aTerm)
{
	const float s = aSum.hi + aTerm;
	const float b = s - aSum.hi;
	const float err = (aSum.hi - (s - b)) + (aTerm - b);
	return ll_sum{s, aSum.lo + err};
}

__device__ __forceinline__ ll_sum ll_merge(ll_sum aA, ll_sum aB)
{
	ll_sum r = ll_add(aA, aB.hi);
	r.lo += aB.lo;
	return r;
}

__device__ __forceinline__ float ll_term(float aObs, float aPred, float aFloor)
{
	return aObs * logf(aFloor) - aPred;
}

// Deterministic block sum (fixed shuffle tree, waves in order); the result is
// valid in thread 0.
__device__ __forceinline__ ll_sum block_sum(ll_sum v, ll_sum* shared)
{
	for (int off = 32; off > 0; off >>= 1) {
		const ll_sum other{__shfl_down(v.hi, off, 64), __shfl_down(v.lo, off, 64)};
		v = ll_merge(v, other);
	}
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	if (lane == 0) {
		shared[wave] = v;
	}
	__syncthreads();
	ll_sum total{0.0f, 0.0f};
	if (threadIdx.x == 0) {
		const int waves = static_cast<int>((blockDim.x + 63) / 64);
		for (int w = 0; w < waves; ++w) {
			total = ll_merge(total, shared[w]);
		}
	}
	return total;
}

// main.cu:22-29 plus the Poisson log-likelihood partial sum(y log max(p, eps) - p)
__global__ void rl_ratio_kernel(const float* y, const float* p, float* r, long long n, float eps, ll_sum* partials)
{
	__shared__ ll_sum shared[kPointThreads / 64];
	ll_sum ll{0.0f, 0.0f};
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float obs = y[i];
		const float pred = p[i];
		const float floor = pred > eps ? pred : eps;
		r[i] = obs / floor;
		if (partials) {
			ll = ll_add(ll, ll_term(obs, pred, floor));
		}
	}
	if (partials) {
		const ll_sum total = block_sum(ll, shared);
		if (threadIdx.x == 0) {
			partials[blockIdx.x] = total;
		}
	}
}

// main.cu:32-49, :74-77
__device__ __forceinline__ float edge_factor(float c, float s, float eps)
{
	return s > 0.0f ? c / (s > eps ? s : eps) : 0.0f;
}

__global__ void rl_update_kernel(float* x, const float* c, const float* s, long long n, float eps)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		x[i] = x[i] * edge_factor(c[i], s[i], eps);
	}
}

// Deterministic final sum of the per-block partials (fixed order); the pair is
// widened to the double the interface returns.
__global__ void ll_finalize_kernel(const ll_sum* partials, int n, double* out)
{
	__shared__ ll_sum shared[kPointThreads / 64];
	ll_sum v{0.0f, 0.0f};
	for (int i = threadIdx.x; i < n; i += blockDim.x) {
		v = ll_merge(v, partials[i]);
	}
	const ll_sum total = block_sum(v, shared);
	if (threadIdx.x == 0) {
		*out = static_cast<double>(total.hi) + static_cast<double>(total.lo);
	}
}

// wiener/main.cu:21-32: X * conj(H) / (|H|^2 + noise), factor in double
__global__ void wiener_kernel(float2* spectrum, const float2* response, long long n, float noise)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float2 o = spectrum[i];
		const double real = response[i].x;
		const double imag = response[i].y;
		const double denominator = real * real + imag * imag + noise;
		spectrum[i] = make_float2(static_cast<float>((o.x * real + o.y * imag) / denominator),
			static_cast<float>((o.y * real - o.x * imag) / denominator));
	}
}

// wiener/main.cu:13-19 (real_sample) and :35-41 (encode_sample's real part x scale)
__global__ void real_to_complex_kernel(const float* in, float2* out, long long n)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		out[i] = make_float2(in[i], 0.0f);
	}
}

__global__ void complex_real_scale_kernel(const float2* in, float* out, long long n, float scale)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		out[i] = in[i].x * scale;
	}
}

unsigned point_blocks(long long n)
{
	return static_cast<unsigned>(std::min<long long>(kPartialBlocks, std::max<long long>(1, (n + 255) / 256)));
}

// ------------------------------------------------------------- FFT row step --
enum { EPI_RATIO = 0, EPI_UPDATE = 1 };

struct rl_row_args {
	fft_pass_args f;        // spectrum rows (in and out), m = n0/2, lines = ny rows, outer = nz planes
	float* x;               // estimate (EPI_UPDATE)
	const float* y;         // observed (EPI_RATIO)
	const float* s;         // sensitivity (EPI_UPDATE)
	long long real_line;    // real image row stride (= nx)
	long long real_outer;   // real image plane stride (= nx*ny)
	int nx;
	float eps;
	int forward_next;       // run the R2C of the new row (0 on the last update)
	ll_sum* partials;       // EPI_RATIO: log-likelihood partials per block (may be null)
};

// C2R of `count` spectrum rows, the RL pointwise step on the real row, R2C of
// the result back into the same spectrum rows.
// One workgroup's rows (group `aBlock` of `a`); the log-likelihood partial of
// the group goes to a.partials[aSlot].
template<int EPI, int FIXED = 0>
__device__ __forceinline__ void rl_row_body(const rl_row_args& a, unsigned aBlock, unsigned aSlot, float2* lds,
	ll_sum* shared)
{
	const fft_pass_args& f = a.f;
	float2* A = lds;
	float2* B = lds + f.lds_elems;
	const twiddle_fetch fetched = fetch_twiddles(f);
	const float2* tw_m = lds + 2 * f.lds_elems; // staged tables (RL plans check tables_in_lds)
	const float2* tw_n = tw_m + f.plan.m;       // W_n^0 .. W_n^m (the rows are half-length real transforms)
	int o, l0, count;
	decode_group(f, aBlock, o, l0, count);
	const int n = f.n;
	const int m = n / 2;
	load_complex_lines(static_cast<const float2*>(f.in), f.gin, o, l0, count, f.load_n, m + 1, A, f.pitch);
	store_twiddles(fetched, f, lds + 2 * f.lds_elems);
	__syncthreads();
	c2r_pre(A, B, count, n, f.pitch, tw_n);
	__syncthreads();
	float2* r = run_plan<FIXED>(f.plan, B, A, count, f.pitch, tw_m, 1.0f);
	// pointwise step on real samples 2j, 2j+1 of each row
	ll_sum ll{0.0f, 0.0f};
	const long long base = o * a.real_outer + l0 * a.real_line;
	const bool want_ll = EPI == EPI_RATIO && a.partials != nullptr; // uniform
	const fast_div by_m(m);
	// the image samples of a row pair are loaded for a whole batch first
	// (samples past nx read sample nx - 1 and are not used)
	using samples_t = std::conditional_t<EPI == EPI_RATIO, float2, float4>;
	batched_loop(count * m,
		[&](int t) {
			const int w = by_m.div(t);
			const int j = t - w * m;
			const long long row = base + w * a.real_line;
			const long long i0 = row + min(2 * j, a.nx - 1);
			const long long i1 = row + min(2 * j + 1, a.nx - 1);
			if constexpr (EPI == EPI_RATIO) {
				return make_float2(a.y[i0], a.y[i1]);
			} else {
				return make_float4(a.x[i0], a.x[i1], a.s[i0], a.s[i1]);
			}
		},
		[&](int t, samples_t aSamples) {
			const int w = by_m.div(t);
			const int j = t - w * m;
			const float2 v = r[w * f.pitch + j];
			float out[2] = {0.0f, 0.0f};
			const float in[2] = {v.x, v.y};
#pragma unroll
			for (int h = 0; h < 2; ++h) {
				const int i = 2 * j + h;
				if (i < a.nx) {
					const long long idx = base + w * a.real_line + i;
					if constexpr (EPI == EPI_RATIO) {
						const float obs = h == 0 ? aSamples.x : aSamples.y;
						const float pred = in[h];
						const float floor = pred > a.eps ? pred : a.eps;
						out[h] = obs / floor;
						if (want_ll) {
							ll = ll_add(ll, ll_term(obs, pred, floor));
						}
					} else {
						const float x = h == 0 ? aSamples.x : aSamples.y;
						const float sens = h == 0 ? aSamples.z : aSamples.w;
						const float xn = x * edge_factor(in[h], sens, a.eps);
						a.x[idx] = xn;
						out[h] = xn;
					}
				}
			}
			r[w * f.pitch + j] = make_float2(out[0], out[1]);
		});
	if constexpr (EPI == EPI_RATIO) {
		if (a.partials) {
			const ll_sum total = block_sum(ll, shared);
			if (threadIdx.x == 0) {
				a.partials[aSlot] = total;
			}
		}
	}
	if (!a.forward_next) {
		return;
	}
	__syncthreads();
	float2* z = run_plan<FIXED>(f.plan, r, r == A ? B : A, count, f.pitch, tw_m, -1.0f);
	float2* xb = z == A ? B : A;
	r2c_post(z, xb, count, n, f.pitch, tw_n);
	__syncthreads();
	store_complex_lines(static_cast<float2*>(f.out), f.gout, o, l0, count, f.store_n, xb, f.pitch);
}

template<int EPI, int FIXED>
__global__ __launch_bounds__(kFftThreads) void rl_row_kernel(rl_row_args a)
{
	extern __shared__ float2 lds[];
	__shared__ ll_sum shared[kFftThreads / 64];
	// the partial of line group g goes to slot g (the batched kernel does the same)
	const unsigned group = xcd_remap(blockIdx.x, gridDim.x);
	rl_row_body<EPI, FIXED>(a, group, group, lds, shared);
}

// ---------------------------------------------------- batched (many volumes) --
// One launch covers the line groups of every volume: block b belongs to the
// volume v with offsets[v] <= b < offsets[v + 1] (offsets are prefix sums).
__device__ __forceinline__ int volume_of(const unsigned* offsets, int count, unsigned b)
{
	int lo = 0, hi = count;
	while (hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if (offsets[mid] <= b) {
			lo = mid;
		} else {
			hi = mid;
		}
	}
	return lo;
}

// The workgroup copies its volume's descriptor into LDS once: the pass body
// reads descriptor fields in its inner loops, and LDS reads of a uniform
// address are far cheaper there than repeated global loads.
template<class T>
__device__ __forceinline__ void copy_descriptor(T& aDst, const T& aSrc)
{
	static_assert(sizeof(T) % 4 == 0 && sizeof(T) <= 4 * kFftThreads, "descriptor copied one word per thread");
	if (threadIdx.x < sizeof(T) / 4) {
		reinterpret_cast<int*>(&aDst)[threadIdx.x] = reinterpret_cast<const int*>(&aSrc)[threadIdx.x];
	}
	__syncthreads();
}
// FIXED: the compile-time plan every volume of the launch shares (0: run-time stages).
template<int MODE, int FIXED = 0>
__global__ __launch_bounds__(kFftThreads) void fft_pass_batched_kernel(
	const fft_pass_args* __restrict__ args, const unsigned* __restrict__ offsets, int count)
{
	extern __shared__ float2 lds[];
	__shared__ fft_pass_args desc;
	const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
	const int v = __builtin_amdgcn_readfirstlane(volume_of(offsets, count, b));
	copy_descriptor(desc, args[v]);
	fft_pass_body<MODE, false, FIXED>(desc, b - offsets[v], lds);
}

template<int EPI, int FIXED = 0>
__global__ __launch_bounds__(kFftThreads) void rl_row_batched_kernel(
	const rl_row_args* __restrict__ args, const unsigned* __restrict__ offsets, int count)
{
	extern __shared__ float2 lds[];
	__shared__ ll_sum shared[kFftThreads / 64];
	__shared__ rl_row_args desc;
	const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
	const int v = __builtin_amdgcn_readfirstlane(volume_of(offsets, count, b));
	copy_descriptor(desc, args[v]);
	rl_row_body<EPI, FIXED>(desc, b - offsets[v], b, lds, shared);
}

// log-likelihood of volume v (partials offsets[v] .. offsets[v+1]-1) into out[v * stride]
__global__ void ll_finalize_batched_kernel(const ll_sum* partials, const unsigned* offsets, double* out, int stride)
{
	__shared__ ll_sum shared[kPointThreads / 64];
	const int v = blockIdx.x;
	ll_sum s{0.0f, 0.0f};
	for (unsigned i = offsets[v] + threadIdx.x; i < offsets[v + 1]; i += blockDim.x) {
		s = ll_merge(s, partials[i]);
	}
	const ll_sum total = block_sum(s, shared);
	if (threadIdx.x == 0) {
		out[static_cast<long long>(v) * stride] = static_cast<double>(total.hi) + static_cast<double>(total.lo);
	}
}

__global__ void rl_initial_batched_kernel(const float* const* y, float* const* x, const long long* counts, float eps)
{
	const int v = blockIdx.y;
	const long long n = counts[v];
	const float* yv = y[v];
	float* xv = x[v];
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const float o = yv[i];
		xv[i] = o > eps ? o : eps;
	}
}

// place the PSF with its centre at the origin of the padded real domain (wrap-around)
__global__ void psf_wrap_kernel(view4 psf, float* out, int n0, int n1, int n2, float scale)
{
	const long long count = static_cast<long long>(psf.extent[0]) * psf.extent[1] * psf.extent[2];
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < count;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		const int q0 = static_cast<int>(i % psf.extent[0]);
		const int q1 = static_cast<int>((i / psf.extent[0]) % psf.extent[1]);
		const int q2 = static_cast<int>(i / (static_cast<long long>(psf.extent[0]) * psf.extent[1]));
		const float v = load_as_float(psf.ptr + q0 * psf.stride[0] + q1 * psf.stride[1] + q2 * psf.stride[2], psf.type);
		const int d0 = ((q0 - psf.extent[0] / 2) % n0 + n0) % n0;
		const int d1 = ((q1 - psf.extent[1] / 2) % n1 + n1) % n1;
		const int d2 = ((q2 - psf.extent[2] / 2) % n2 + n2) % n2;
		out[(static_cast<long long>(d2) * n1 + d1) * n0 + d0] = v * scale;
	}
}

// reverse a packed buffer (flips every axis, main.cu:106-107) and convert to float
__global__ void flip_kernel(view4 psf, float* out)
{
	const long long count = static_cast<long long>(psf.extent[0]) * psf.extent[1] * psf.extent[2] * psf.extent[3];
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < count;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		long long rest = i;
		long long off = 0;
		for (int a = 0; a < 4; ++a) {
			off += (rest % psf.extent[a]) * psf.stride[a];
			rest /= psf.extent[a];
		}
		out[count - 1 - i] = load_as_float(psf.ptr + off, psf.type);
	}
}

__global__ void fill_kernel(float* out, long long n, float v)
{
	for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
		 i += static_cast<long long>(gridDim.x) * blockDim.x) {
		out[i] = v;
	}
}

} // namespace

// ----------------------------------------------------------------- plan ------
struct rl_plan_impl {
	int rank = 0;
	int method = 0;
	int precision = 0;
	float eps = 1e-6f;
	int n[3] = {1, 1, 1};       // image extents (IANE order)
	int k[3] = {1, 1, 1};       // PSF extents
	long long count = 0;
	// persistent device data
	float* sensitivity = nullptr;
	float* psf = nullptr;       // packed float PSF (direct)
	float* flipped = nullptr;   // packed flipped PSF (direct)
	float2* spectrum_h = nullptr; // PSF spectrum scaled by 1/(n0 n1 n2) (fft)
	// split-precision direct method: kernel parts (scale + Toeplitz table) of
	// the PSF and of its flip, built once; the source part is run scratch
	char* split_psf = nullptr;
	char* split_flipped = nullptr;
	size_t split_scratch = 0;     // bytes of the per-call source part (0: no split path)
	size_t split_offset = 0;      // where the source part starts in the run workspace
	// fft geometry
	int pad[3] = {1, 1, 1};
	int h0 = 1;
	size_t spectrum_bytes = 0;
	std::vector<fft_pass_args> forward_mid;  // middle passes with H
	std::vector<fft_pass_args> adjoint_mid;  // middle passes with conj(H)
	fft_pass_args r2c{};                     // R2C of image rows into the spectrum
	fft_pass_args c2r{};                     // C2R of spectrum rows into an image
	rl_row_args row_tpl{};                    // row-kernel template
	int partial_capacity = kPartialBlocks;   // log-likelihood partial sums (one per block)
	size_t workspace = 0;

	~rl_plan_impl()
	{
		// destructors cannot report: a failed free is ignored
		(void)hipFree(sensitivity);
		(void)hipFree(psf);
		(void)hipFree(flipped);
		(void)hipFree(spectrum_h);
		(void)hipFree(split_psf);
		(void)hipFree(split_flipped);
	}

	view4 image_view(float* p) const
	{
		view4 v{};
		v.ptr = reinterpret_cast<char*>(p);
		v.type = IANE_F32;
		v.rank = rank;
		long long stride = 4;
		for (int i = 0; i < 4; ++i) {
			v.extent[i] = i < rank ? n[i] : 1;
			v.stride[i] = i < rank ? stride : 0;
			stride *= v.extent[i];
		}
		return v;
	}

	view4 kernel_view(float* p) const
	{
		view4 v{};
		v.ptr = reinterpret_cast<char*>(p);
		v.type = IANE_F32;
		v.rank = rank;
		long long stride = 4;
		for (int i = 0; i < 4; ++i) {
			v.extent[i] = i < rank ? k[i] : 1;
			v.stride[i] = i < rank ? stride : 0;
			stride *= v.extent[i];
		}
		return v;
	}

	// out = in (*) kernel, zero border; `scratch` holds split_scratch bytes when
	// the split path runs (kernel is psf or flipped).
	void convolve(const float* in, const float* kernel, float* out, void* scratch, hipStream_t stream) const
	{
		const view4 src = image_view(const_cast<float*>(in));
		const view4 ker = kernel_view(const_cast<float*>(kernel));
		const view4 dst = image_view(out);
		if (split_scratch) {
			split_convolve_prepared(src, ker, dst, IANE_BORDER_ZERO, precision,
				kernel == psf ? split_psf : split_flipped, scratch, stream);
			return;
		}
		convolve_direct(src, ker, dst, IANE_BORDER_ZERO,
			precision == IANE_PREC_REFERENCE ? IANE_PREC_REFERENCE : IANE_PREC_FP32, stream, false);
	}

	// Spectrum geometry helpers (packed h0 x n1 x n2 complex array).
	void build_fft(hipStream_t stream);
	void run_mid(const std::vector<fft_pass_args>& passes, float2* spec, hipStream_t stream) const
	{
		for (fft_pass_args a : passes) {
			a.in = spec;
			a.out = spec;
			launch_pass(a, stream);
		}
	}
};

void rl_plan_impl::build_fft(hipStream_t stream)
{
	for (int i = 0; i < 3; ++i) {
		pad[i] = i < rank ? next_fft_length(n[i] + k[i] / 2) : 1;
	}
	h0 = pad[0] / 2 + 1;
	const long long spec_count = static_cast<long long>(h0) * pad[1] * pad[2];
	spectrum_bytes = static_cast<size_t>(spec_count) * sizeof(float2);
	const long long row = h0;               // spectrum row stride
	const long long plane = row * pad[1];   // spectrum plane stride
	const int nx = n[0], ny = n[1], nz = rank == 3 ? n[2] : 1;
	// R2C of image rows (y < ny, z < nz), zero beyond nx
	setup_pass(r2c, PASS_R2C, pad[0], ny, nz);
	r2c.load_n = nx;
	r2c.store_n = h0;
	r2c.gin = {1, nx, static_cast<long long>(nx) * ny};
	r2c.gout = {1, row, plane};
	// C2R of spectrum rows, only x < nx stored
	setup_pass(c2r, PASS_C2R, pad[0], ny, nz);
	c2r.load_n = h0;
	c2r.store_n = nx;
	c2r.gin = {1, row, plane};
	c2r.gout = {1, nx, static_cast<long long>(nx) * ny};
	// middle passes
	const float scale = 1.0f;
	auto mid = [&](int conj) {
		std::vector<fft_pass_args> passes;
		if (rank == 3) {
			fft_pass_args yf{};
			setup_pass(yf, PASS_C2C, pad[1], h0, nz);
			yf.sign = -1.0f;
			yf.load_n = ny;
			yf.store_n = pad[1];
			yf.gin = yf.gout = {row, 1, plane};
			fft_pass_args zm{};
			setup_pass(zm, PASS_C2C_MULT, pad[2], static_cast<int>(row * pad[1]), 1);
			zm.load_n = nz;
			zm.store_n = nz;
			zm.gin = zm.gout = zm.gmult = {plane, 1, 0};
			zm.mult = spectrum_h;
			zm.mult_conj = conj;
			zm.mult_scale = scale;
			fft_pass_args yi{};
			setup_pass(yi, PASS_C2C, pad[1], h0, nz);
			yi.sign = 1.0f;
			yi.load_n = pad[1];
			yi.store_n = ny;
			yi.gin = yi.gout = {row, 1, plane};
			passes = {yf, zm, yi};
		} else {
			fft_pass_args ym{};
			setup_pass(ym, PASS_C2C_MULT, pad[1], h0, 1);
			ym.load_n = ny;
			ym.store_n = ny;
			ym.gin = ym.gout = ym.gmult = {row, 1, 0};
			ym.mult = spectrum_h;
			ym.mult_conj = conj;
			ym.mult_scale = scale;
			passes = {ym};
		}
		return passes;
	};
	// PSF spectrum: wrap the PSF into the padded domain, full R2C + forward passes
	const long long real_count = static_cast<long long>(pad[0]) * pad[1] * pad[2];
	float* wrapped = nullptr;
	check_hip(hipMallocAsync(reinterpret_cast<void**>(&wrapped), real_count * sizeof(float), stream), "Allocate PSF buffer");
	check_hip(hipMemsetAsync(wrapped, 0, real_count * sizeof(float), stream), "Clear PSF buffer");
	check_hip(hipMalloc(&spectrum_h, spectrum_bytes), "Allocate PSF spectrum");
	{
		view4 kp = kernel_view(psf);
		psf_wrap_kernel<<<point_blocks(static_cast<long long>(k[0]) * k[1] * k[2]), kPointThreads, 0, stream>>>(
			kp, wrapped, pad[0], pad[1], pad[2], static_cast<float>(1.0 / static_cast<double>(real_count)));
		check_hip(hipGetLastError(), "PSF wrap launch");
		fft_pass_args full{};
		setup_pass(full, PASS_R2C, pad[0], pad[1], pad[2]);
		full.load_n = pad[0];
		full.store_n = h0;
		full.gin = {1, pad[0], static_cast<long long>(pad[0]) * pad[1]};
		full.gout = {1, row, plane};
		full.in = wrapped;
		full.out = spectrum_h;
		launch_pass(full, stream);
		for (int axis = 1; axis < rank; ++axis) {
			fft_pass_args c{};
			if (axis == 1) {
				setup_pass(c, PASS_C2C, pad[1], h0, pad[2]);
				c.gin = c.gout = {row, 1, plane};
			} else {
				setup_pass(c, PASS_C2C, pad[2], static_cast<int>(row * pad[1]), 1);
				c.gin = c.gout = {plane, 1, 0};
			}
			c.sign = -1.0f;
			c.load_n = c.store_n = pad[axis];
			c.in = c.out = spectrum_h;
			launch_pass(c, stream);
		}
	}
	check_hip(hipFreeAsync(wrapped, stream), "Free PSF buffer");
	forward_mid = mid(0);
	adjoint_mid = mid(1);
	// the row kernels and the batched passes read the twiddles from their LDS copies
	bool staged = r2c.tables_in_lds && c2r.tables_in_lds;
	for (const fft_pass_args& a : forward_mid) {
		staged = staged && a.tables_in_lds;
	}
	if (!staged) {
		fail(IANE_ERR_UNSUPPORTED, "RL padded FFT line exceeds the LDS line limit");
	}
	// row-kernel template: the spectrum rows of the image region
	row_tpl.f = c2r;
	row_tpl.f.gout = {1, row_tpl.f.gin.line, row_tpl.f.gin.outer};
	row_tpl.f.store_n = h0;
	row_tpl.real_line = nx;
	row_tpl.real_outer = static_cast<long long>(nx) * ny;
	row_tpl.nx = nx;
	row_tpl.eps = eps;
	partial_capacity = static_cast<int>(std::max<long long>(kPartialBlocks,
		static_cast<long long>(row_tpl.f.groups_per_outer) * row_tpl.f.outer));
	workspace = ((spectrum_bytes + 255) / 256) * 256 + static_cast<size_t>(partial_capacity) * sizeof(ll_sum);
}

} // namespace iane_hip

using namespace iane_hip;

struct iane_rl_plan {
	std::unique_ptr<rl_plan_impl> impl;
};

// Validate and prepare one RL problem (shared by single plans and batches).
std::unique_ptr<rl_plan_impl> make_rl_plan(int aRank, const int32_t* aExtent, const view4& psf, int aMethod,
	int aPrecision, float aEpsilon, hipStream_t stream)
{
	if (aRank != 2 && aRank != 3) {
		fail(IANE_ERR_UNSUPPORTED, "RL plans support rank 2 and 3 images");
	}
	if (aMethod != IANE_RL_DIRECT && aMethod != IANE_RL_FFT) {
		invalid("unknown RL method");
	}
	if (!(aEpsilon > 0.0f) || !std::isfinite(aEpsilon)) {
		invalid("Iterations and finite epsilon must be positive");
	}
	if (psf.rank != aRank) {
		invalid("Convolution requires a kernel and images of matching rank");
	}
	auto impl = std::make_unique<rl_plan_impl>();
	impl->rank = aRank;
	impl->method = aMethod;
	impl->precision = aPrecision;
	impl->eps = aEpsilon;
	impl->count = 1;
	for (int i = 0; i < aRank; ++i) {
		if (aExtent[i] <= 0) {
			invalid("RL requires positive image dimensions");
		}
		if (psf.extent[i] <= 0 || psf.extent[i] % 2 == 0) {
			invalid("Convolution requires positive odd kernel dimensions");
		}
		impl->n[i] = aExtent[i];
		impl->k[i] = psf.extent[i];
		impl->count *= aExtent[i];
	}
	const long long kcount = static_cast<long long>(impl->k[0]) * impl->k[1] * impl->k[2];
	check_hip(hipMalloc(&impl->psf, kcount * sizeof(float)), "Allocate PSF");
	check_hip(hipMalloc(&impl->flipped, kcount * sizeof(float)), "Allocate flipped PSF");
	check_hip(hipMalloc(&impl->sensitivity, impl->count * sizeof(float)), "Allocate sensitivity");
	{
		// packed float copy of the PSF, then its flip (reverse of the packed buffer)
		view4 packed = impl->kernel_view(impl->flipped);
		flip_kernel<<<point_blocks(kcount), kPointThreads, 0, stream>>>(psf, impl->flipped);
		flip_kernel<<<point_blocks(kcount), kPointThreads, 0, stream>>>(packed, impl->psf);
		check_hip(hipGetLastError(), "PSF copy launch");
	}
	if (aMethod == IANE_RL_DIRECT) {
		const view4 image = impl->image_view(nullptr);
		const view4 kernel = impl->kernel_view(impl->psf);
		const bool split = (aPrecision == IANE_PREC_SPLIT_BF16 || aPrecision == IANE_PREC_SPLIT_FP16) &&
			split_applicable(image, kernel, image);
		if (split) {
			// the PSF's and its flip's scale and Toeplitz table, once per plan
			const size_t kernel_bytes = split_kernel_bytes(kernel, aPrecision);
			check_hip(hipMalloc(&impl->split_psf, kernel_bytes), "Allocate split PSF table");
			check_hip(hipMalloc(&impl->split_flipped, kernel_bytes), "Allocate split PSF table");
			split_prepare_kernel(kernel, aPrecision, impl->split_psf, stream);
			split_prepare_kernel(impl->kernel_view(impl->flipped), aPrecision, impl->split_flipped, stream);
			impl->split_scratch = split_source_bytes(image, kernel, aPrecision);
		}
		// run workspace: predicted, correction, log-likelihood partials, split source part
		const size_t images = (2 * impl->count * sizeof(float) + 255) / 256 * 256;
		const size_t partials = (static_cast<size_t>(kPartialBlocks + 64) * sizeof(ll_sum) + 255) / 256 * 256;
		impl->split_offset = images + partials;
		impl->workspace = impl->split_offset + impl->split_scratch;
		// s = H^T 1 (main.cu:62-63): the flipped PSF over an image of ones, zero border
		float* ones = nullptr;
		char* scratch = nullptr;
		check_hip(hipMallocAsync(reinterpret_cast<void**>(&ones), impl->count * sizeof(float), stream), "Allocate ones");
		if (impl->split_scratch) {
			check_hip(hipMallocAsync(reinterpret_cast<void**>(&scratch), impl->split_scratch, stream),
				"Allocate split scratch");
		}
		fill_kernel<<<point_blocks(impl->count), kPointThreads, 0, stream>>>(ones, impl->count, 1.0f);
		impl->convolve(ones, impl->flipped, impl->sensitivity, scratch, stream);
		check_hip(hipFreeAsync(ones, stream), "Free ones");
		if (scratch) {
			check_hip(hipFreeAsync(scratch, stream), "Free split scratch");
		}
	} else {
		impl->build_fft(stream);
		// s = H^T 1 through the same pruned passes
		float* ones = nullptr;
		float2* spec = nullptr;
		check_hip(hipMallocAsync(reinterpret_cast<void**>(&ones), impl->count * sizeof(float), stream), "Allocate ones");
		check_hip(hipMallocAsync(reinterpret_cast<void**>(&spec), impl->spectrum_bytes, stream), "Allocate spectrum");
		fill_kernel<<<point_blocks(impl->count), kPointThreads, 0, stream>>>(ones, impl->count, 1.0f);
		fft_pass_args a = impl->r2c;
		a.in = ones;
		a.out = spec;
		launch_pass(a, stream);
		impl->run_mid(impl->adjoint_mid, spec, stream);
		fft_pass_args c = impl->c2r;
		c.in = spec;
		c.out = impl->sensitivity;
		launch_pass(c, stream);
		check_hip(hipFreeAsync(ones, stream), "Free ones");
		check_hip(hipFreeAsync(spec, stream), "Free spectrum");
	}
	return impl;
}

namespace {

unsigned row_blocks(const fft_pass_args& a)
{
	return static_cast<unsigned>(static_cast<long long>(a.groups_per_outer) * a.outer);
}

template<int EPI>
void launch_row(const rl_row_args& a, hipStream_t s)
{
	const unsigned blocks = row_blocks(a.f);
	const size_t lds = pass_lds_bytes(a.f);
	with_fixed_plan(a.f.fixed, [&](auto aId) {
		rl_row_kernel<EPI, decltype(aId)::value><<<blocks, kFftThreads, lds, s>>>(a);
	});
	check_hip(hipGetLastError(), "RL row launch");
}

} // namespace

extern "C" {

int iane_hip_rl_initial(const float* aObserved, float* aEstimate, int64_t aCount, float aEps, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aObserved || !aEstimate))) {
			invalid("RL buffers are null");
		}
		if (aCount) {
			rl_initial_kernel<<<point_blocks(aCount), kPointThreads, 0, static_cast<hipStream_t>(aStream)>>>(
				aObserved, aEstimate, aCount, aEps);
			check_hip(hipGetLastError(), "RL initial launch");
		}
	});
}

int iane_hip_rl_ratio_workspace_bytes(int64_t aCount, size_t* aBytes)
{
	return guarded([&] {
		if (!aBytes || aCount < 0) {
			invalid("RL ratio workspace query needs a non-negative count and an output");
		}
		// one compensated partial sum per block of the ratio kernel
		*aBytes = aCount ? static_cast<size_t>(point_blocks(aCount)) * sizeof(ll_sum) : 0;
	});
}

int iane_hip_rl_ratio(const float* aObserved, const float* aPredicted, float* aRatio, int64_t aCount, float aEps,
	double* aLlAccum, void* aWorkspace, size_t aWorkspaceBytes, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aObserved || !aPredicted || !aRatio))) {
			invalid("RL buffers are null");
		}
		if (aCount == 0) {
			return;
		}
		auto stream = static_cast<hipStream_t>(aStream);
		const unsigned blocks = point_blocks(aCount);
		ll_sum* partials = nullptr;
		if (aLlAccum) {
			check_workspace(aWorkspace, aWorkspaceBytes, blocks * sizeof(ll_sum));
			partials = static_cast<ll_sum*>(aWorkspace);
		}
		rl_ratio_kernel<<<blocks, kPointThreads, 0, stream>>>(aObserved, aPredicted, aRatio, aCount, aEps, partials);
		check_hip(hipGetLastError(), "RL ratio launch");
		if (aLlAccum) {
			ll_finalize_kernel<<<1, kPointThreads, 0, stream>>>(partials, blocks, aLlAccum);
			check_hip(hipGetLastError(), "RL log-likelihood launch");
		}
	});
}

} // extern "C"

// ------------------------------------------------ allocation-free LL ratio --
// iane_hip_rl_ratio_noalloc keeps the per-block log-likelihood partial sums in
// this file-scope device array instead of a caller workspace. The ratio
// kernel's grid is capped at kPartialBlocks blocks (point_blocks), so
// kPartialBlocks compensated sums (8 KiB) serve every element count. Each GPU
// holds its own copy for the life of the process; calls that overlap on one
// GPU share it and race.
namespace iane_hip {
namespace {

constexpr int kStaticPartials = kPartialBlocks;
static_assert(sizeof(ll_sum) * kStaticPartials <= 8192, "the static scratch stays small");
static __device__ __align__(256) ll_sum scratch[kStaticPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

// The address of `scratch` on the current device, looked up once per device
// (host code must never pass the symbol itself to a kernel).
ll_sum* device_scratch()
{
	int device = 0;
	check_hip(hipGetDevice(&device), "Query current device");
	if (device < 0 || device >= kMaxDevices) {
		fail(IANE_ERR_UNSUPPORTED, "device index beyond the static scratch table");
	}
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_address[device]) {
		void* address = nullptr;
		check_hip(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "Locate static log-likelihood scratch");
		g_scratch_address[device] = address;
	}
	return static_cast<ll_sum*>(g_scratch_address[device]);
}

} // namespace
} // namespace iane_hip

extern "C" {

int iane_hip_rl_ratio_noalloc(const float* aObserved, const float* aPredicted, float* aRatio, int64_t aCount,
	float aEps, double* aLlAccum, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aObserved || !aPredicted || !aRatio))) {
			invalid("RL buffers are null");
		}
		if (aCount == 0) {
			return;
		}
		const unsigned blocks = point_blocks(aCount);
		if (blocks > static_cast<unsigned>(kStaticPartials)) {
			fail(IANE_ERR_LENGTH, "RL ratio needs more partial sums than the static scratch holds");
		}
		auto stream = static_cast<hipStream_t>(aStream);
		ll_sum* partials = aLlAccum ? device_scratch() : nullptr;
		// every partial the finalize step reads is written by this launch: nothing to clear
		rl_ratio_kernel<<<blocks, kPointThreads, 0, stream>>>(aObserved, aPredicted, aRatio, aCount, aEps, partials);
		check_hip(hipGetLastError(), "RL ratio launch");
		if (aLlAccum) {
			ll_finalize_kernel<<<1, kPointThreads, 0, stream>>>(partials, blocks, aLlAccum);
			check_hip(hipGetLastError(), "RL log-likelihood launch");
		}
	});
}

int iane_hip_rl_update(float* aEstimate, const float* aCorrection, const float* aSensitivity, int64_t aCount,
	float aEps, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aEstimate || !aCorrection || !aSensitivity))) {
			invalid("RL buffers are null");
		}
		if (aCount) {
			rl_update_kernel<<<point_blocks(aCount), kPointThreads, 0, static_cast<hipStream_t>(aStream)>>>(
				aEstimate, aCorrection, aSensitivity, aCount, aEps);
			check_hip(hipGetLastError(), "RL update launch");
		}
	});
}

int iane_hip_wiener_apply(void* aSpectrum, const void* aResponse, int64_t aCount, float aNoise, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aSpectrum || !aResponse))) {
			invalid("Wiener buffers are null");
		}
		if (!(aNoise > 0.0f) || !std::isfinite(aNoise)) {
			invalid("Noise must be finite and positive");
		}
		if (aCount) {
			wiener_kernel<<<point_blocks(aCount), kPointThreads, 0, static_cast<hipStream_t>(aStream)>>>(
				static_cast<float2*>(aSpectrum), static_cast<const float2*>(aResponse), aCount, aNoise);
			check_hip(hipGetLastError(), "Wiener launch");
		}
	});
}

int iane_hip_real_to_complex(const float* aIn, void* aOut, int64_t aCount, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aIn || !aOut))) {
			invalid("buffers are null");
		}
		if (aCount) {
			real_to_complex_kernel<<<point_blocks(aCount), kPointThreads, 0, static_cast<hipStream_t>(aStream)>>>(
				aIn, static_cast<float2*>(aOut), aCount);
			check_hip(hipGetLastError(), "real to complex launch");
		}
	});
}

int iane_hip_complex_real_scale(const void* aIn, float* aOut, int64_t aCount, float aScale, void* aStream)
{
	return guarded([&] {
		if (aCount < 0 || (aCount && (!aIn || !aOut))) {
			invalid("buffers are null");
		}
		if (aCount) {
			complex_real_scale_kernel<<<point_blocks(aCount), kPointThreads, 0, static_cast<hipStream_t>(aStream)>>>(
				static_cast<const float2*>(aIn), aOut, aCount, aScale);
			check_hip(hipGetLastError(), "complex real scale launch");
		}
	});
}

int iane_hip_rl_plan_create(iane_rl_plan** aPlan, int aRank, const int32_t* aExtent, const iane_view_desc* aPsf,
	int aMethod, int aPrecision, float aEpsilon, void* aStream)
{
	return guarded([&] {
		if (!aPlan || !aExtent) {
			invalid("RL plan output or extent is null");
		}
		*aPlan = nullptr;
		auto impl = make_rl_plan(aRank, aExtent, make_view4(aPsf, "psf"), aMethod, aPrecision, aEpsilon,
			static_cast<hipStream_t>(aStream));
		auto plan = std::make_unique<iane_rl_plan>();
		plan->impl = std::move(impl);
		*aPlan = plan.release();
	});
}

int iane_hip_rl_plan_workspace_bytes(const iane_rl_plan* aPlan, size_t* aBytes)
{
	return guarded([&] {
		if (!aPlan || !aBytes) {
			invalid("RL plan or output is null");
		}
		*aBytes = aPlan->impl->workspace;
	});
}

int iane_hip_rl_plan_blur(iane_rl_plan* aPlan, const float* aIn, float* aOut, void* aWorkspace, void* aStream)
{
	return guarded([&] {
		if (!aPlan || !aIn || !aOut || !aWorkspace) {
			invalid("RL blur buffers are null");
		}
		const rl_plan_impl& p = *aPlan->impl;
		auto stream = static_cast<hipStream_t>(aStream);
		if (p.method == IANE_RL_DIRECT) {
			p.convolve(aIn, p.psf, aOut, static_cast<char*>(aWorkspace) + p.split_offset, stream);
			return;
		}
		float2* spec = static_cast<float2*>(aWorkspace);
		fft_pass_args a = p.r2c;
		a.in = aIn;
		a.out = spec;
		launch_pass(a, stream);
		p.run_mid(p.forward_mid, spec, stream);
		fft_pass_args c = p.c2r;
		c.in = spec;
		c.out = aOut;
		launch_pass(c, stream);
	});
}

int iane_hip_rl_run(iane_rl_plan* aPlan, const float* aObserved, float* aEstimate, int aIterations, int aInit,
	void* aWorkspace, double* aLlHistory, void* aStream)
{
	return guarded([&] {
		if (!aPlan) {
			invalid("RL plan is null");
		}
		if (!aObserved || !aEstimate) {
			invalid("RL buffers are null");
		}
		if (aIterations < 0) {
			invalid("Iterations and finite epsilon must be positive");
		}
		const rl_plan_impl& p = *aPlan->impl;
		if (!aWorkspace) {
			fail(IANE_ERR_LENGTH, "Prepared operation exceeded its scratch reservation");
		}
		auto stream = static_cast<hipStream_t>(aStream);
		const long long n = p.count;
		if (aInit) {
			rl_initial_kernel<<<point_blocks(n), kPointThreads, 0, stream>>>(aObserved, aEstimate, n, p.eps);
			check_hip(hipGetLastError(), "RL initial launch");
		}
		if (aIterations == 0) {
			return;
		}
		char* ws = static_cast<char*>(aWorkspace);
		if (p.method == IANE_RL_DIRECT) {
			float* predicted = reinterpret_cast<float*>(ws);
			float* correction = predicted + n;
			ll_sum* partials = reinterpret_cast<ll_sum*>(ws + ((2 * n * sizeof(float) + 255) / 256) * 256);
			char* split = ws + p.split_offset;
			const unsigned blocks = point_blocks(n);
			for (int it = 0; it < aIterations; ++it) {
				p.convolve(aEstimate, p.psf, predicted, split, stream);
				rl_ratio_kernel<<<blocks, kPointThreads, 0, stream>>>(aObserved, predicted, predicted, n, p.eps,
					aLlHistory ? partials : nullptr);
				if (aLlHistory) {
					ll_finalize_kernel<<<1, kPointThreads, 0, stream>>>(partials, blocks, aLlHistory + it);
				}
				p.convolve(predicted, p.flipped, correction, split, stream);
				rl_update_kernel<<<blocks, kPointThreads, 0, stream>>>(aEstimate, correction, p.sensitivity, n, p.eps);
			}
			check_hip(hipGetLastError(), "RL iteration launch");
			return;
		}
		float2* spec = reinterpret_cast<float2*>(ws);
		ll_sum* partials = reinterpret_cast<ll_sum*>(ws + ((p.spectrum_bytes + 255) / 256) * 256);
		// spectrum of the current estimate
		fft_pass_args a = p.r2c;
		a.in = aEstimate;
		a.out = spec;
		launch_pass(a, stream);
		p.run_mid(p.forward_mid, spec, stream);
		rl_row_args ratio = p.row_tpl;
		ratio.f.in = spec;
		ratio.f.out = spec;
		ratio.y = aObserved;
		ratio.forward_next = 1;
		rl_row_args update = ratio;
		update.x = aEstimate;
		update.s = p.sensitivity;
		const unsigned rb = row_blocks(ratio.f);
		for (int it = 0; it < aIterations; ++it) {
			ratio.partials = aLlHistory ? partials : nullptr;
			launch_row<EPI_RATIO>(ratio, stream);
			if (aLlHistory) {
				ll_finalize_kernel<<<1, kPointThreads, 0, stream>>>(partials, rb, aLlHistory + it);
			}
			p.run_mid(p.adjoint_mid, spec, stream);
			update.forward_next = it + 1 < aIterations;
			launch_row<EPI_UPDATE>(update, stream);
			if (it + 1 < aIterations) {
				p.run_mid(p.forward_mid, spec, stream);
			}
		}
	});
}

int iane_hip_rl_plan_destroy(iane_rl_plan* aPlan)
{
	return guarded([&] { delete aPlan; });
}

} // extern "C"

// ------------------------------------------------------------------ batches --
// Many RL problems of uneven sizes, each with its own PSF, advanced together:
// every pass of the iteration is ONE launch that covers the line groups of all
// volumes (per-volume pass descriptors in device memory, prefix-summed block
// offsets), and the whole `iterations`-long sequence is captured once into a
// hipGraph and replayed with a single graph launch. Results per volume are
// identical to the single-volume FFT plan.
namespace {

template<class Args>
struct batched_launch {
	std::vector<Args> host;
	std::vector<unsigned> offsets;
	Args* d_args = nullptr;
	unsigned* d_offsets = nullptr;
	unsigned blocks = 0;
	size_t lds_bytes = 0;  // dynamic LDS of the launch: the largest of the volumes' passes
	int fixed = 0;         // the compile-time plan of every volume's pass, if they share one

	batched_launch() = default;
	batched_launch(const batched_launch&) = delete;
	batched_launch& operator=(const batched_launch&) = delete;
	~batched_launch()
	{
		(void)hipFree(d_args);
		(void)hipFree(d_offsets);
	}

	// offsets from the per-volume group counts; device copies of both tables
	void finalize(const fft_pass_args& (*aPass)(const Args&))
	{
		offsets.assign(host.size() + 1, 0);
		fixed = host.empty() ? 0 : aPass(host[0]).fixed;
		for (size_t v = 0; v < host.size(); ++v) {
			const fft_pass_args& f = aPass(host[v]);
			fixed = f.fixed == fixed ? fixed : 0;
			offsets[v + 1] = offsets[v] + static_cast<unsigned>(static_cast<long long>(f.groups_per_outer) * f.outer);
			lds_bytes = std::max(lds_bytes, pass_lds_bytes(f));
		}
		blocks = offsets.back();
		check_hip(hipMalloc(&d_args, sizeof(Args) * host.size()), "Allocate batch descriptors");
		check_hip(hipMalloc(&d_offsets, sizeof(unsigned) * offsets.size()), "Allocate batch offsets");
		check_hip(hipMemcpy(d_offsets, offsets.data(), sizeof(unsigned) * offsets.size(), hipMemcpyHostToDevice),
			"Upload batch offsets");
	}

	void upload(hipStream_t aStream) const
	{
		check_hip(hipMemcpyAsync(d_args, host.data(), sizeof(Args) * host.size(), hipMemcpyHostToDevice, aStream),
			"Upload batch descriptors");
	}
};

const fft_pass_args& pass_of(const fft_pass_args& a) { return a; }
const fft_pass_args& pass_of_row(const rl_row_args& a) { return a.f; }

template<int MODE>
void launch_batched_fixed(const batched_launch<fft_pass_args>& b, hipStream_t s)
{
	const int n = static_cast<int>(b.host.size());
	with_fixed_plan(b.fixed, [&](auto aId) {
		fft_pass_batched_kernel<MODE, decltype(aId)::value><<<b.blocks, kFftThreads, b.lds_bytes, s>>>(b.d_args,
			b.d_offsets, n);
	});
}

void launch_batched(const batched_launch<fft_pass_args>& b, int aMode, hipStream_t s)
{
	const int n = static_cast<int>(b.host.size());
	switch (aMode) {
	case PASS_C2C: launch_batched_fixed<PASS_C2C>(b, s); break;
	case PASS_R2C: fft_pass_batched_kernel<PASS_R2C><<<b.blocks, kFftThreads, b.lds_bytes, s>>>(b.d_args, b.d_offsets, n); break;
	case PASS_C2R: fft_pass_batched_kernel<PASS_C2R><<<b.blocks, kFftThreads, b.lds_bytes, s>>>(b.d_args, b.d_offsets, n); break;
	default: launch_batched_fixed<PASS_C2C_MULT>(b, s); break;
	}
}

template<int EPI>
void launch_row_batched(const batched_launch<rl_row_args>& b, int aCount, hipStream_t s)
{
	with_fixed_plan(b.fixed, [&](auto aId) {
		rl_row_batched_kernel<EPI, decltype(aId)::value><<<b.blocks, kFftThreads, b.lds_bytes, s>>>(b.d_args,
			b.d_offsets, aCount);
	});
}

} // namespace

struct iane_rl_batch {
	int count = 0;
	int rank = 0;
	float eps = 1e-6f;
	std::vector<std::unique_ptr<rl_plan_impl>> plans;
	batched_launch<fft_pass_args> r2c;
	std::vector<std::unique_ptr<batched_launch<fft_pass_args>>> fwd;
	std::vector<std::unique_ptr<batched_launch<fft_pass_args>>> adj;
	std::vector<int> mid_modes;
	batched_launch<rl_row_args> ratio;
	batched_launch<rl_row_args> update_next;
	batched_launch<rl_row_args> update_last;
	std::vector<size_t> spectrum_offset;
	size_t partial_offset = 0;
	size_t workspace = 0;
	const float** d_observed = nullptr;
	float** d_estimate = nullptr;
	long long* d_counts = nullptr;
	long long max_count = 0;
	// replayable graph of one run configuration
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	int g_iterations = -1;
	int g_init = -1;
	double* g_ll = nullptr;
	void* g_stream = nullptr;
	void* g_workspace = nullptr;

	~iane_rl_batch()
	{
		// destructors cannot report: failed releases are ignored
		if (exec) {
			(void)hipGraphExecDestroy(exec);
		}
		if (graph) {
			(void)hipGraphDestroy(graph);
		}
		(void)hipFree(d_observed);
		(void)hipFree(d_estimate);
		(void)hipFree(d_counts);
	}

	// the launch sequence of `iterations` RL steps (captured into the graph)
	void enqueue(int aIterations, int aInit, double* aLl, hipStream_t s) const
	{
		if (aInit) {
			const unsigned gx = static_cast<unsigned>(std::min<long long>(256, (max_count + 255) / 256));
			rl_initial_batched_kernel<<<dim3(gx, count), kPointThreads, 0, s>>>(d_observed, d_estimate, d_counts, eps);
		}
		launch_batched(r2c, PASS_R2C, s);
		for (size_t p = 0; p < fwd.size(); ++p) {
			launch_batched(*fwd[p], mid_modes[p], s);
		}
		for (int it = 0; it < aIterations; ++it) {
			launch_row_batched<EPI_RATIO>(ratio, count, s);
			if (aLl) {
				ll_finalize_batched_kernel<<<count, kPointThreads, 0, s>>>(ratio.host[0].partials, ratio.d_offsets,
					aLl + it, aIterations);
			}
			for (size_t p = 0; p < adj.size(); ++p) {
				launch_batched(*adj[p], mid_modes[p], s);
			}
			const auto& up = it + 1 < aIterations ? update_next : update_last;
			launch_row_batched<EPI_UPDATE>(up, count, s);
			if (it + 1 < aIterations) {
				for (size_t p = 0; p < fwd.size(); ++p) {
					launch_batched(*fwd[p], mid_modes[p], s);
				}
			}
		}
		check_hip(hipGetLastError(), "RL batch launch");
	}
};

extern "C" {

int iane_hip_rl_batch_create(iane_rl_batch** aBatch, int aCount, int aRank, const int32_t* aExtents,
	const iane_view_desc* aPsfs, float aEpsilon, void* aStream)
{
	return guarded([&] {
		if (!aBatch || !aExtents || !aPsfs) {
			invalid("RL batch output, extents or PSFs are null");
		}
		*aBatch = nullptr;
		if (aCount <= 0) {
			invalid("RL batch needs at least one volume");
		}
		auto stream = static_cast<hipStream_t>(aStream);
		auto batch = std::make_unique<iane_rl_batch>();
		batch->count = aCount;
		batch->rank = aRank;
		batch->eps = aEpsilon;
		size_t offset = 0;
		auto align = [](size_t x) { return (x + 255) / 256 * 256; };
		for (int v = 0; v < aCount; ++v) {
			batch->plans.push_back(make_rl_plan(aRank, aExtents + v * aRank, make_view4(&aPsfs[v], "psf"), IANE_RL_FFT,
				IANE_PREC_FP32, aEpsilon, stream));
			batch->spectrum_offset.push_back(offset);
			offset += align(batch->plans.back()->spectrum_bytes);
			batch->max_count = std::max(batch->max_count, batch->plans.back()->count);
		}
		const size_t mids = batch->plans[0]->forward_mid.size();
		for (size_t p = 0; p < mids; ++p) {
			batch->fwd.push_back(std::make_unique<batched_launch<fft_pass_args>>());
			batch->adj.push_back(std::make_unique<batched_launch<fft_pass_args>>());
			batch->mid_modes.push_back(batch->plans[0]->forward_mid[p].mode);
		}
		for (const auto& plan : batch->plans) {
			batch->r2c.host.push_back(plan->r2c);
			for (size_t p = 0; p < mids; ++p) {
				batch->fwd[p]->host.push_back(plan->forward_mid[p]);
				batch->adj[p]->host.push_back(plan->adjoint_mid[p]);
			}
			rl_row_args row = plan->row_tpl;
			row.forward_next = 1;
			batch->ratio.host.push_back(row);
			batch->update_next.host.push_back(row);
			row.forward_next = 0;
			batch->update_last.host.push_back(row);
		}
		batch->r2c.finalize(pass_of);
		for (size_t p = 0; p < mids; ++p) {
			batch->fwd[p]->finalize(pass_of);
			batch->adj[p]->finalize(pass_of);
		}
		batch->ratio.finalize(pass_of_row);
		batch->update_next.finalize(pass_of_row);
		batch->update_last.finalize(pass_of_row);
		batch->partial_offset = offset;
		batch->workspace = offset + static_cast<size_t>(batch->ratio.blocks) * sizeof(ll_sum);
		check_hip(hipMalloc(reinterpret_cast<void**>(&batch->d_observed), sizeof(float*) * aCount), "Allocate tables");
		check_hip(hipMalloc(reinterpret_cast<void**>(&batch->d_estimate), sizeof(float*) * aCount), "Allocate tables");
		check_hip(hipMalloc(reinterpret_cast<void**>(&batch->d_counts), sizeof(long long) * aCount), "Allocate tables");
		std::vector<long long> counts;
		for (const auto& plan : batch->plans) {
			counts.push_back(plan->count);
		}
		check_hip(hipMemcpy(batch->d_counts, counts.data(), sizeof(long long) * aCount, hipMemcpyHostToDevice),
			"Upload counts");
		check_hip(hipStreamSynchronize(stream), "RL batch prepare");
		*aBatch = batch.release();
	});
}

int iane_hip_rl_batch_workspace_bytes(const iane_rl_batch* aBatch, size_t* aBytes)
{
	return guarded([&] {
		if (!aBatch || !aBytes) {
			invalid("RL batch or output is null");
		}
		*aBytes = aBatch->workspace;
	});
}

int iane_hip_rl_batch_run(iane_rl_batch* aBatch, const float* const* aObserved, float* const* aEstimates,
	int aIterations, int aInit, void* aWorkspace, double* aLlHistory, int aUseGraph, void* aStream)
{
	return guarded([&] {
		if (!aBatch || !aObserved || !aEstimates) {
			invalid("RL batch or buffer tables are null");
		}
		if (!aWorkspace) {
			fail(IANE_ERR_LENGTH, "Prepared operation exceeded its scratch reservation");
		}
		if (aIterations < 0) {
			invalid("Iterations and finite epsilon must be positive");
		}
		iane_rl_batch& b = *aBatch;
		auto stream = static_cast<hipStream_t>(aStream);
		if (aUseGraph && !stream) {
			invalid("graph capture needs a created (non-null) stream");
		}
		char* ws = static_cast<char*>(aWorkspace);
		ll_sum* partials = reinterpret_cast<ll_sum*>(ws + b.partial_offset);
		for (int v = 0; v < b.count; ++v) {
			if (!aObserved[v] || !aEstimates[v]) {
				invalid("RL batch buffers are null");
			}
			float2* spec = reinterpret_cast<float2*>(ws + b.spectrum_offset[v]);
			auto bind = [&](fft_pass_args& f, const void* in, void* out) {
				f.in = in;
				f.out = out;
			};
			bind(b.r2c.host[v], aEstimates[v], spec);
			for (size_t p = 0; p < b.fwd.size(); ++p) {
				bind(b.fwd[p]->host[v], spec, spec);
				bind(b.adj[p]->host[v], spec, spec);
			}
			for (auto* rows : {&b.ratio, &b.update_next, &b.update_last}) {
				rl_row_args& r = rows->host[v];
				bind(r.f, spec, spec);
				r.y = aObserved[v];
				r.x = aEstimates[v];
				r.s = b.plans[v]->sensitivity;
				r.partials = partials;
			}
		}
		// descriptor tables are read by the kernels (and by the captured graph) from device memory
		b.r2c.upload(stream);
		for (size_t p = 0; p < b.fwd.size(); ++p) {
			b.fwd[p]->upload(stream);
			b.adj[p]->upload(stream);
		}
		b.ratio.upload(stream);
		b.update_next.upload(stream);
		b.update_last.upload(stream);
		check_hip(hipMemcpyAsync(b.d_observed, aObserved, sizeof(float*) * b.count, hipMemcpyHostToDevice, stream),
			"Upload tables");
		check_hip(hipMemcpyAsync(b.d_estimate, aEstimates, sizeof(float*) * b.count, hipMemcpyHostToDevice, stream),
			"Upload tables");
		if (aIterations == 0 && !aInit) {
			return;
		}
		if (!aUseGraph) {
			b.enqueue(aIterations, aInit, aLlHistory, stream);
			return;
		}
		const bool reuse = b.exec && b.g_iterations == aIterations && b.g_init == aInit && b.g_ll == aLlHistory &&
			b.g_stream == aStream && b.g_workspace == aWorkspace;
		if (!reuse) {
			if (b.exec) {
				check_hip(hipGraphExecDestroy(b.exec), "Destroy RL graph");
				b.exec = nullptr;
			}
			if (b.graph) {
				check_hip(hipGraphDestroy(b.graph), "Destroy RL graph");
				b.graph = nullptr;
			}
			check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "Begin RL graph capture");
			try {
				b.enqueue(aIterations, aInit, aLlHistory, stream);
			} catch (...) {
				// end the capture before reporting the original error
				hipGraph_t discard = nullptr;
				(void)hipStreamEndCapture(stream, &discard);
				if (discard) {
					(void)hipGraphDestroy(discard);
				}
				throw;
			}
			check_hip(hipStreamEndCapture(stream, &b.graph), "End RL graph capture");
			check_hip(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "Instantiate RL graph");
			b.g_iterations = aIterations;
			b.g_init = aInit;
			b.g_ll = aLlHistory;
			b.g_stream = aStream;
			b.g_workspace = aWorkspace;
		}
		check_hip(hipGraphLaunch(b.exec, stream), "Launch RL graph");
	});
}

int iane_hip_rl_batch_destroy(iane_rl_batch* aBatch)
{
	return guarded([&] { delete aBatch; });
}

} // extern "C"
