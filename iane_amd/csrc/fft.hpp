// Stockham FFT building blocks shared by the FFT plans (fft.hip) and the
// FFT-based Richardson–Lucy pipeline (rl.hip).
//
// Every kernel works on "lines": complete 1-D transforms of length m held in
// LDS by one workgroup. A workgroup processes `bundle` lines at once; lines
// along axis 0 are contiguous rows, lines along axes 1/2 are bundled with
// their axis-0 neighbours so global loads and stores stay coalesced. Radix
// stages (2, 3, 4, 5, 7, 8, 11, 13, generic odd) ping-pong between two LDS
// buffers in the Stockham autosort order, so results land in natural order.
//
// Real transforms along axis 0 use the half-length complex trick for even n
// (z[j] = x[2j] + i x[2j+1], one m = n/2 FFT and an O(n) post/pre-process)
// and a full-length complex FFT for odd n.
//
// A line length m with a prime factor above 64 runs Bluestein's chirp-z form in
// the same LDS buffers: with w[j] = exp(sign i pi j^2 / m),
//   X[k] = w[k] sum_j (x[j] w[j]) conj(w[k - j]),
// a linear convolution evaluated with a length-M transform (M >= 2m - 1, radices
// 2..7): chirp and zero-pad, forward FFT, multiply by the spectrum of the
// wrapped conj(w) (precomputed in double, divided by M), inverse FFT, chirp.
#pragma once

#include "common.hpp"

#include <type_traits>

namespace iane_hip {

constexpr int kFftThreads = 256;
constexpr int kFftLdsElems = 4096;   // complex elements per LDS buffer (32 KiB; two buffers)
constexpr int kFftMaxStages = 16;   // radix >= 2 and lines <= 10240 points: at most 14 stages

// One Stockham stage with its constants worked out on the host: nb = m / R
// butterflies per line, ns = product of the earlier radices, step = m / (ns R)
// (twiddle stride), and the fast_div multipliers of nb and ns.
struct fft_stage {
	int nb;
	int ns;
	int step;
	unsigned c_nb;
	unsigned c_ns;
};

struct fft_radix_plan {
	int m;                       // complex line length
	int stages;
	int radix[kFftMaxStages];
	fft_stage stage[kFftMaxStages];
};

// Division of n < 2^16 by a wave-uniform d in [1, 2^16): q = hi32(n * c) with
// c = ceil(2^32 / d) is exact in that range (c = (2^32 + e) / d with e < d, so
// the error term n e / (d 2^32) stays below 1 / d). One 32-bit division per
// divisor instead of one per element.
struct fast_div {
	unsigned d;
	unsigned c;
	__host__ __device__ __forceinline__ explicit fast_div(int aD)
		: d(static_cast<unsigned>(aD)), c(aD > 1 ? 0xFFFFFFFFu / static_cast<unsigned>(aD) + 1u : 0u) {}
	// a divisor whose multiplier the host already worked out
	__device__ __forceinline__ fast_div(int aD, unsigned aC) : d(static_cast<unsigned>(aD)), c(aC) {}
	__device__ __forceinline__ int div(int aN) const
	{
		return d == 1 ? aN : static_cast<int>(__umulhi(static_cast<unsigned>(aN), c));
	}
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// i * a
__device__ __forceinline__ float2 cmuli(float2 a) { return make_float2(-a.y, a.x); }

// cos(2 pi j / R) and sin(2 pi j / R), j = 0..R-1, for the odd prime radices
// with their own butterfly (values of the mathematical constants, rounded to float).
template<int R>
struct unit_roots;
template<>
struct unit_roots<7> {
	static constexpr float c[7] = {1.000000000e+00f, 6.234898019e-01f, -2.225209340e-01f, -9.009688679e-01f, -9.009688679e-01f, -2.225209340e-01f, 6.234898019e-01f};
	static constexpr float s[7] = {0.000000000e+00f, 7.818314825e-01f, 9.749279122e-01f, 4.338837391e-01f, -4.338837391e-01f, -9.749279122e-01f, -7.818314825e-01f};
};
template<>
struct unit_roots<11> {
	static constexpr float c[11] = {1.000000000e+00f, 8.412535328e-01f, 4.154150130e-01f, -1.423148383e-01f, -6.548607339e-01f, -9.594929736e-01f, -9.594929736e-01f, -6.548607339e-01f, -1.423148383e-01f, 4.154150130e-01f, 8.412535328e-01f};
	static constexpr float s[11] = {0.000000000e+00f, 5.406408175e-01f, 9.096319954e-01f, 9.898214419e-01f, 7.557495744e-01f, 2.817325568e-01f, -2.817325568e-01f, -7.557495744e-01f, -9.898214419e-01f, -9.096319954e-01f, -5.406408175e-01f};
};
template<>
struct unit_roots<13> {
	static constexpr float c[13] = {1.000000000e+00f, 8.854560257e-01f, 5.680647467e-01f, 1.205366803e-01f, -3.546048870e-01f, -7.485107482e-01f, -9.709418174e-01f, -9.709418174e-01f, -7.485107482e-01f, -3.546048870e-01f, 1.205366803e-01f, 5.680647467e-01f, 8.854560257e-01f};
	static constexpr float s[13] = {0.000000000e+00f, 4.647231720e-01f, 8.229838659e-01f, 9.927088741e-01f, 9.350162427e-01f, 6.631226582e-01f, 2.393156643e-01f, -2.393156643e-01f, -6.631226582e-01f, -9.350162427e-01f, -9.927088741e-01f, -8.229838659e-01f, -4.647231720e-01f};
};

// R-point DFT in registers; sign -1 forward, +1 inverse; tw = W_m table, m = line length.
template<int R>
__device__ __forceinline__ void dft(float2 (&v)[R], float sign, const float2* tw, int m)
{
	if constexpr (R == 2) {
		const float2 a = v[0], b = v[1];
		v[0] = cadd(a, b);
		v[1] = csub(a, b);
	} else if constexpr (R == 3) {
		const float c = -0.5f;
		const float s = sign * 0.86602540378443864676f;
		const float2 t = cadd(v[1], v[2]);
		const float2 u = csub(v[1], v[2]);
		const float2 base = make_float2(v[0].x + c * t.x, v[0].y + c * t.y);
		const float2 rot = cmuli(cscale(u, s));
		v[0] = cadd(v[0], t);
		v[1] = cadd(base, rot);
		v[2] = csub(base, rot);
	} else if constexpr (R == 4) {
		const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
		const float2 b0 = cadd(v[1], v[3]);
		float2 b1 = csub(v[1], v[3]);
		b1 = sign < 0 ? make_float2(b1.y, -b1.x) : make_float2(-b1.y, b1.x); // (sign i) * b1
		v[0] = cadd(a0, b0);
		v[2] = csub(a0, b0);
		v[1] = cadd(a1, b1);
		v[3] = csub(a1, b1);
	} else if constexpr (R == 5) {
		const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
		const float s1 = sign * 0.95105651629515357212f, s2 = sign * 0.58778525229247312917f;
		const float2 b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
		const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
		const float2 r1 = make_float2(v[0].x + c1 * b1.x + c2 * b2.x, v[0].y + c1 * b1.y + c2 * b2.y);
		const float2 r2 = make_float2(v[0].x + c2 * b1.x + c1 * b2.x, v[0].y + c2 * b1.y + c1 * b2.y);
		const float2 i1 = cmuli(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
		const float2 i2 = cmuli(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
		v[0] = cadd(v[0], cadd(b1, b2));
		v[1] = cadd(r1, i1);
		v[4] = csub(r1, i1);
		v[2] = cadd(r2, i2);
		v[3] = csub(r2, i2);
	} else if constexpr (R == 8) {
		float2 e[4] = {v[0], v[2], v[4], v[6]};
		float2 o[4] = {v[1], v[3], v[5], v[7]};
		dft<4>(e, sign, tw, m);
		dft<4>(o, sign, tw, m);
		const float r = 0.70710678118654752440f;
		// w8^k = exp(sign * 2 pi i k / 8)
		const float2 w1 = make_float2(r, sign * r);
		const float2 w3 = make_float2(-r, sign * r);
		o[1] = cmul(o[1], w1);
		o[2] = sign < 0 ? make_float2(o[2].y, -o[2].x) : make_float2(-o[2].y, o[2].x);
		o[3] = cmul(o[3], w3);
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			v[k] = cadd(e[k], o[k]);
			v[k + 4] = csub(e[k], o[k]);
		}
	} else {
		// odd prime radix R (7, 11, 13): with a_k = v[k] + v[R-k], b_k = v[k] - v[R-k]
		// (k = 1..h, h = (R-1)/2), X[0] = v[0] + sum a_k and, for q = 1..h,
		// X[q] = C_q + (sign i) S_q, X[R-q] = C_q - (sign i) S_q with
		// C_q = v[0] + sum a_k cos(2 pi qk / R), S_q = sum b_k sin(2 pi qk / R).
		constexpr int h = (R - 1) / 2;
		float2 a[h + 1], b[h + 1];
		float2 sum = v[0];
#pragma unroll
		for (int k = 1; k <= h; ++k) {
			a[k] = cadd(v[k], v[R - k]);
			b[k] = csub(v[k], v[R - k]);
			sum = cadd(sum, a[k]);
		}
		float2 out[R];
		out[0] = sum;
#pragma unroll
		for (int q = 1; q <= h; ++q) {
			float2 cq = v[0];
			float2 sq = make_float2(0.0f, 0.0f);
#pragma unroll
			for (int k = 1; k <= h; ++k) {
				const float c = unit_roots<R>::c[(q * k) % R];
				const float sn = unit_roots<R>::s[(q * k) % R];
				cq = make_float2(cq.x + c * a[k].x, cq.y + c * a[k].y);
				sq = make_float2(sq.x + sn * b[k].x, sq.y + sn * b[k].y);
			}
			const float2 rot = sign < 0 ? make_float2(sq.y, -sq.x) : make_float2(-sq.y, sq.x); // (sign i) * S_q
			out[q] = cadd(cq, rot);
			out[R - q] = csub(cq, rot);
		}
#pragma unroll
		for (int k = 0; k < R; ++k) {
			v[k] = out[k];
		}
		(void)tw;
		(void)m;
	}
}

// One Stockham stage of radix R over `bundle` lines (LDS line pitch `pitch`).
template<int R>
__device__ __forceinline__ void stockham_stage(const float2* src, float2* dst, const fft_stage& st, int m, int bundle,
	int pitch, const float2* tw, float sign)
{
	const int nb = st.nb;
	const int Ns = st.ns;
	const int total = bundle * nb;
	const int step = st.step;
	const fast_div by_nb(nb, st.c_nb), by_ns(Ns, st.c_ns);
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int w = by_nb.div(t);
		const int j = t - w * nb;
		const float2* s = src + w * pitch;
		float2 v[R];
#pragma unroll
		for (int r = 0; r < R; ++r) {
			v[r] = s[j + r * nb];
		}
		const int k = j - by_ns.div(j) * Ns;
		if (Ns > 1) {
#pragma unroll
			for (int r = 1; r < R; ++r) {
				float2 wv = tw[k * r * step];
				if (sign > 0) {
					wv = cconj(wv);
				}
				v[r] = cmul(v[r], wv);
			}
		}
		dft<R>(v, sign, tw, m);
		const int out = (j - k) * R + k;
		float2* d = dst + w * pitch;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			d[out + r * Ns] = v[r];
		}
	}
}

// Generic radix (any odd factor without a specialisation), runtime R <= 64.
__device__ inline void stockham_stage_generic(const float2* src, float2* dst, int m, int R, int Ns, int bundle,
	int pitch, const float2* tw, float sign)
{
	const int nb = m / R;
	const int total = bundle * nb;
	const int step = m / (Ns * R);
	const int dstep = m / R;
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int w = t / nb;
		const int j = t - w * nb;
		const float2* s = src + w * pitch;
		const int k = j % Ns;
		const int out = (j - k) * R + k;
		float2* d = dst + w * pitch;
		for (int q = 0; q < R; ++q) {
			float2 acc = make_float2(0.0f, 0.0f);
			for (int r = 0; r < R; ++r) {
				float2 v = s[j + r * nb];
				if (Ns > 1 && r > 0) {
					float2 wv = tw[k * r * step];
					if (sign > 0) {
						wv = cconj(wv);
					}
					v = cmul(v, wv);
				}
				float2 wd = tw[((r * q) % R) * dstep];
				if (sign > 0) {
					wd = cconj(wd);
				}
				acc = cadd(acc, cmul(v, wd));
			}
			d[out + q * Ns] = acc;
		}
	}
}

// Run every stage of `plan` on the lines in buffer `a` (scratch `b`).
// Returns the buffer holding the result. Ends with a barrier.
// Always inlined: as a called function it needs a scratch stack frame and
// its register count caps the occupancy of every kernel that calls it.
__device__ __forceinline__ float2* run_stages(const fft_radix_plan& plan, float2* a, float2* b, int bundle, int pitch,
	const float2* tw, float sign)
{
	for (int s = 0; s < plan.stages; ++s) {
		const int R = plan.radix[s];
		const fft_stage& st = plan.stage[s];
		switch (R) {
		case 2: stockham_stage<2>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 3: stockham_stage<3>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 4: stockham_stage<4>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 5: stockham_stage<5>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 7: stockham_stage<7>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 8: stockham_stage<8>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 11: stockham_stage<11>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		case 13: stockham_stage<13>(a, b, st, plan.m, bundle, pitch, tw, sign); break;
		default: stockham_stage_generic(a, b, plan.m, R, st.ns, bundle, pitch, tw, sign); break;
		}
		__syncthreads();
		float2* t = a;
		a = b;
		b = t;
	}
	return a;
}

// ---------------------------------------------------------------------------
// Plans fixed at compile time. A line length whose radix sequence is in the
// table below also runs as a chain of stages with the length, the radices and
// the strides as template arguments: the same operations in the same order as
// run_stages (so the same bits), without the per-stage division constants,
// stride products and scalar spills of the run-time loop.
template<int M, int NS, int R>
__device__ __forceinline__ void stockham_stage_fixed(const float2* src, float2* dst, int bundle, int pitch,
	const float2* tw, float sign)
{
	constexpr int nb = M / R;
	constexpr int step = M / (NS * R);
	const unsigned total = static_cast<unsigned>(bundle * nb);
	for (unsigned t = threadIdx.x; t < total; t += blockDim.x) {
		const unsigned w = t / nb;
		const int j = static_cast<int>(t - w * nb);
		const float2* s = src + w * pitch;
		float2 v[R];
#pragma unroll
		for (int r = 0; r < R; ++r) {
			v[r] = s[j + r * nb];
		}
		const int k = static_cast<int>(static_cast<unsigned>(j) % NS);
		if constexpr (NS > 1) {
#pragma unroll
			for (int r = 1; r < R; ++r) {
				float2 wv = tw[k * r * step];
				if (sign > 0) {
					wv = cconj(wv);
				}
				v[r] = cmul(v[r], wv);
			}
		}
		dft<R>(v, sign, tw, M);
		const int out = (j - k) * R + k;
		float2* d = dst + w * pitch;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			d[out + r * NS] = v[r];
		}
	}
}

template<int M, int NS, int R, int... REST>
__device__ __forceinline__ float2* fixed_stages(float2* a, float2* b, int bundle, int pitch, const float2* tw,
	float sign)
{
	stockham_stage_fixed<M, NS, R>(a, b, bundle, pitch, tw, sign);
	__syncthreads();
	if constexpr (sizeof...(REST) == 0) {
		return b;
	} else {
		return fixed_stages<M, NS * R, REST...>(b, a, bundle, pitch, tw, sign);
	}
}

// The table: id -> line length and radix sequence (the order factor_radices
// produces). id 0 is the run-time plan. The lengths are those of the RL padded
// domain of 256 x 256 x 128 volumes with PSFs up to 31 x 31 x 63 (rows 140,
// y 280, z 160) and powers of two from 128 to 1024.
template<int M, int... R>
struct fixed_sequence {
	static constexpr int m = M;
	__device__ static float2* run(float2* a, float2* b, int n, int p, const float2* tw, float s)
	{
		return fixed_stages<M, 1, R...>(a, b, n, p, tw, s);
	}
};
template<int ID>
struct fixed_plan;
template<> struct fixed_plan<1> : fixed_sequence<140, 5, 7, 4> {};
template<> struct fixed_plan<2> : fixed_sequence<280, 5, 7, 8> {};
template<> struct fixed_plan<3> : fixed_sequence<160, 5, 8, 4> {};
template<> struct fixed_plan<4> : fixed_sequence<128, 8, 8, 2> {};
template<> struct fixed_plan<5> : fixed_sequence<256, 8, 8, 4> {};
template<> struct fixed_plan<6> : fixed_sequence<512, 8, 8, 8> {};
template<> struct fixed_plan<7> : fixed_sequence<1024, 8, 8, 8, 2> {};
constexpr int kFixedPlans = 7;

// Call aFn(std::integral_constant<int, id>) for the plan id (0 .. kFixedPlans):
// the launch sites instantiate one kernel per id.
template<class Fn>
inline void with_fixed_plan(int aId, Fn aFn)
{
	switch (aId) {
	case 1: aFn(std::integral_constant<int, 1>{}); break;
	case 2: aFn(std::integral_constant<int, 2>{}); break;
	case 3: aFn(std::integral_constant<int, 3>{}); break;
	case 4: aFn(std::integral_constant<int, 4>{}); break;
	case 5: aFn(std::integral_constant<int, 5>{}); break;
	case 6: aFn(std::integral_constant<int, 6>{}); break;
	case 7: aFn(std::integral_constant<int, 7>{}); break;
	default: aFn(std::integral_constant<int, 0>{}); break;
	}
}

// run_stages, or the fixed plan FIXED (its sequence must be the plan's).
template<int FIXED>
__device__ __forceinline__ float2* run_plan(const fft_radix_plan& plan, float2* a, float2* b, int bundle, int pitch,
	const float2* tw, float sign)
{
	if constexpr (FIXED == 0) {
		return run_stages(plan, a, b, bundle, pitch, tw, sign);
	} else {
		return fixed_plan<FIXED>::run(a, b, bundle, pitch, tw, sign);
	}
}

// Host: the fixed plan id of a radix sequence (0: none).
int fixed_plan_id(const fft_radix_plan& aPlan);

// Host: factor m into supported radices (largest first; any leftover factor
// becomes one generic stage, limited to 64).
bool factor_radices(int aM, fft_radix_plan& aPlan);

// Host: device table W_m^k = exp(-2 pi i k / m), k < m, computed in double;
// cached per (device, m) for the life of the process.
const float2* twiddle_table(int aM);

// Smallest even length >= n whose prime factors are all in {2, 3, 5, 7}.
int next_fft_length(int aN);

// ---------------------------------------------------------------------------
// Line geometry: point k of line (o, l) sits at o*outer + l*line + k*elem
// (element units of the buffer's type).
struct line_geom {
	long long elem;
	long long line;
	long long outer;
	// a second outer level (0: none): outer group o sits at
	// (o % outer_split) * outer + (o / outer_split) * outer_hi
	long long outer_hi;
	int outer_split;
};

__host__ __device__ __forceinline__ long long outer_offset(const line_geom& g, int aOuter)
{
	return g.outer_split == 0 ? aOuter * g.outer :
		(aOuter % g.outer_split) * g.outer + (aOuter / g.outer_split) * g.outer_hi;
}

enum { PASS_C2C = 0, PASS_R2C = 1, PASS_C2R = 2, PASS_C2C_MULT = 3 };

struct fft_pass_args {
	const void* in;
	void* out;
	int mode;
	int n;            // logical length: complex length (C2C) or real length (R2C/C2R)
	float sign;       // C2C direction (-1 forward, +1 inverse)
	int load_n;       // points loaded per line (reals for R2C, bins for C2R), the rest are zero
	int store_n;      // points stored per line (complex for C2C/R2C, reals for C2R)
	line_geom gin;
	line_geom gout;
	int lines;        // lines per outer group
	int outer;        // outer groups
	int bundle;       // lines per workgroup
	int pitch;        // LDS pitch per line (complex elements)
	int lds_elems;    // complex elements per LDS buffer (bundle * pitch); two buffers
	unsigned groups_per_outer;
	fft_radix_plan plan; // the complex transform run in LDS (length m; M for Bluestein)
	int line_m;          // complex length of the line transform
	int fixed;           // compile-time plan id of a.plan (0: run-time stages)
	int post_n;          // four-step first pass (else 0): point k of line l times W_post_n^((l / post_div) k)
	int post_div;
	const float2* tw_post; // W_post_n table
	int tables_in_lds;   // 1: twiddle tables staged in LDS after the line buffers; 0: read from memory
	const float2* chirp; // Bluestein (else null): exp(-i pi j^2 / line_m), j < line_m
	const float2* bspec; // Bluestein: spectrum of the wrapped conj(chirp), length M, divided by M
	const float2* tw_m;
	const float2* tw_n;  // W_n table for the real pre/post processing (even n)
	const float2* mult;  // PASS_C2C_MULT: spectrum multiplier, read with gmult
	line_geom gmult;
	int mult_conj;
	float mult_scale;
};

__device__ __forceinline__ void decode_group(const fft_pass_args& a, unsigned aBlock, int& aOuter, int& aLine0,
	int& aCount)
{
	aOuter = static_cast<int>(aBlock / a.groups_per_outer);
	aLine0 = static_cast<int>(aBlock % a.groups_per_outer) * a.bundle;
	aCount = min(a.bundle, a.lines - aLine0);
}

// Load complex points k < load_n (zero up to `width`) of `count` lines into buf.
// Element loop over t = threadIdx.x + i * blockDim.x < total with kLineBatch
// elements per thread in flight: aFetch(t) (a global or LDS load, from an
// in-range address) runs for the whole batch before aUse(t, value) consumes
// any of it. Nothing branches between a load and its use, so the compiler's
// wait before each use covers only that element's load (a load followed at
// once by its use leaves one load per thread in flight).
constexpr int kLineBatch = 8;

template<class Fetch, class Use>
__device__ __forceinline__ void batched_loop(int total, Fetch aFetch, Use aUse)
{
	const int stride = blockDim.x;
	for (int t0 = threadIdx.x; t0 < total; t0 += kLineBatch * stride) {
		using value_t = decltype(aFetch(0));
		value_t v[kLineBatch];
#pragma unroll
		for (int u = 0; u < kLineBatch; ++u) {
			const int t = t0 + u * stride;
			v[u] = aFetch(t < total ? t : t0);
		}
#pragma unroll
		for (int u = 0; u < kLineBatch; ++u) {
			const int t = t0 + u * stride;
			if (t < total) {
				aUse(t, v[u]);
			}
		}
	}
}

// Points k >= load_n of a line are zero: they read the line's point
// min(k, load_n - 1) (in range) and store 0.
__device__ inline void load_complex_lines(const float2* in, const line_geom& g, int aOuter, int aLine0, int count,
	int load_n, int width, float2* buf, int pitch)
{
	const long long base = outer_offset(g, aOuter) + aLine0 * g.line;
	const int total = count * width;
	const int last = load_n > 0 ? load_n - 1 : 0;
	const float2 zero = make_float2(0.0f, 0.0f);
	// the fetch also returns the LDS slot (negative: a point past load_n, stored as 0)
	struct fetched {
		float2 v;
		int slot;
	};
	if (g.elem == 1) {
		const fast_div by_width(width);
		batched_loop(total,
			[&](int t) {
				const int w = by_width.div(t);
				const int k = t - w * width;
				const int slot = w * pitch + k;
				return fetched{in[base + w * g.line + min(k, last)], k < load_n ? slot : -1 - slot};
			},
			[&](int, fetched f) { buf[f.slot < 0 ? -1 - f.slot : f.slot] = f.slot < 0 ? zero : f.v; });
	} else {
		const fast_div by_count(count);
		batched_loop(total,
			[&](int t) {
				const int k = by_count.div(t);
				const int w = t - k * count;
				const int slot = w * pitch + k;
				return fetched{in[base + w * g.line + min(k, last) * g.elem], k < load_n ? slot : -1 - slot};
			},
			[&](int, fetched f) { buf[f.slot < 0 ? -1 - f.slot : f.slot] = f.slot < 0 ? zero : f.v; });
	}
}

__device__ inline void store_complex_lines(float2* out, const line_geom& g, int aOuter, int aLine0, int count,
	int store_n, const float2* buf, int pitch)
{
	const long long base = outer_offset(g, aOuter) + aLine0 * g.line;
	const int total = count * store_n;
	// the fetch (an LDS read) also returns the destination index
	struct fetched {
		float2 v;
		long long at;
	};
	if (g.elem == 1) {
		const fast_div by_store(store_n);
		batched_loop(total,
			[&](int t) {
				const int w = by_store.div(t);
				const int k = t - w * store_n;
				return fetched{buf[w * pitch + k], base + w * g.line + k};
			},
			[&](int, fetched f) { out[f.at] = f.v; });
	} else {
		const fast_div by_count(count);
		batched_loop(total,
			[&](int t) {
				const int k = by_count.div(t);
				const int w = t - k * count;
				return fetched{buf[w * pitch + k], base + w * g.line + k * g.elem};
			},
			[&](int, fetched f) { out[f.at] = f.v; });
	}
}

// Real line of length n (points >= load_n are zero) packed for the complex FFT:
// even n: z[j] = x[2j] + i x[2j+1] (j < n/2); odd n: z[j] = x[j].
__device__ inline void load_real_lines(const float* in, const line_geom& g, int aOuter, int aLine0, int count,
	int n, int load_n, float2* buf, int pitch)
{
	const long long base = outer_offset(g, aOuter) + aLine0 * g.line;
	const bool even = (n & 1) == 0;
	const int width = even ? n / 2 : n;
	const int total = count * width;
	const int last = load_n > 0 ? load_n - 1 : 0;
	const fast_div by_width(width), by_count(count);
	// (line w, position j) of element t
	const auto place = [&](int t) {
		if (g.elem == 1) {
			const int w = by_width.div(t);
			return make_int2(w, t - w * width);
		}
		const int j = by_count.div(t);
		return make_int2(t - j * count, j);
	};
	// points >= load_n read point load_n - 1 and store 0
	batched_loop(total,
		[&](int t) {
			const int2 wj = place(t);
			const float* line = in + base + wj.x * g.line;
			const int i0 = even ? 2 * wj.y : wj.y;
			return make_float2(line[min(i0, last) * g.elem], even ? line[min(i0 + 1, last) * g.elem] : 0.0f);
		},
		[&](int t, float2 v) {
			const int2 wj = place(t);
			const int i0 = even ? 2 * wj.y : wj.y;
			buf[wj.x * pitch + wj.y] = make_float2(i0 < load_n ? v.x : 0.0f, even && i0 + 1 < load_n ? v.y : 0.0f);
		});
}

// even n, after the m = n/2 forward FFT of the packed line in `z`:
// X[k] = E[k] + W_n^k O[k], k = 0..m, written to `x` (pitch >= m + 1).
__device__ inline void r2c_post(const float2* z, float2* x, int count, int n, int pitch, const float2* tw_n)
{
	const int m = n / 2;
	const int total = count * (m + 1);
	const fast_div by_line(m + 1);
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int w = by_line.div(t);
		const int k = t - w * (m + 1);
		const float2* zl = z + w * pitch;
		const float2 zk = zl[k == m ? 0 : k];
		const float2 zc = cconj(zl[k == 0 ? 0 : m - k]);
		const float2 e = cscale(cadd(zk, zc), 0.5f);
		const float2 d = csub(zk, zc);
		const float2 o = make_float2(0.5f * d.y, -0.5f * d.x); // d / (2i)
		x[w * pitch + k] = cadd(e, cmul(tw_n[k], o));
	}
}

// even n, before the m = n/2 inverse FFT: from bins X[0..m] in `x` build
// Z'[k] = (X[k] + conj X[m-k]) + i (X[k] - conj X[m-k]) W_n^{-k}, k < m, into `z`.
__device__ inline void c2r_pre(const float2* x, float2* z, int count, int n, int pitch, const float2* tw_n)
{
	const int m = n / 2;
	const int total = count * m;
	const fast_div by_line(m);
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int w = by_line.div(t);
		const int k = t - w * m;
		const float2* xl = x + w * pitch;
		const float2 xk = xl[k];
		const float2 xc = cconj(xl[m - k]);
		const float2 s = cadd(xk, xc);
		const float2 d = cmul(csub(xk, xc), cconj(tw_n[k]));
		z[w * pitch + k] = cadd(s, cmuli(d));
	}
}

// odd n, before the inverse FFT: Hermitian extension of bins X[0..n/2] in place.
__device__ inline void c2r_extend_odd(float2* x, int count, int n, int pitch)
{
	const int h = n / 2;
	const int total = count * (n - h - 1);
	const fast_div by_line(n - h - 1);
	for (int t = threadIdx.x; t < total; t += blockDim.x) {
		const int w = by_line.div(t);
		const int k = h + 1 + (t - w * (n - h - 1));
		x[w * pitch + k] = cconj(x[w * pitch + n - k]);
	}
}

__device__ inline void store_real_lines(float* out, const line_geom& g, int aOuter, int aLine0, int count, int n,
	int store_n, const float2* buf, int pitch)
{
	const long long base = outer_offset(g, aOuter) + aLine0 * g.line;
	const bool even = (n & 1) == 0;
	const int total = count * store_n;
	const fast_div by_store(store_n), by_count(count);
	// (line w, point i) of element t
	const auto place = [&](int t) {
		if (g.elem == 1) {
			const int w = by_store.div(t);
			return make_int2(w, t - w * store_n);
		}
		const int i = by_count.div(t);
		return make_int2(t - i * count, i);
	};
	batched_loop(total,
		[&](int t) {
			const int2 wi = place(t);
			const float2 v = buf[wi.x * pitch + (even ? wi.y / 2 : wi.y)];
			return (even && (wi.y & 1)) ? v.y : v.x;
		},
		[&](int t, float v) {
			const int2 wi = place(t);
			out[base + wi.x * g.line + wi.y * g.elem] = v;
		});
}

// Twiddle entries a pass keeps in LDS: W of the transform length (aStages
// entries) and, for a half-length real transform of line length aM, W_n^k for
// k = 0 .. aM (the only ones its pre/post step reads).
__host__ __device__ inline int pass_table_elems(int aStages, int aM, bool aHalf)
{
	return aStages + (aHalf ? aM + 1 : 0);
}

// Dynamic LDS of a launch running pass `a`: two buffers of a.lds_elems, then
// the twiddle tables when they are staged.
inline size_t pass_lds_bytes(const fft_pass_args& a)
{
	const size_t tables = a.tables_in_lds ? pass_table_elems(a.plan.m, a.line_m, a.tw_n != nullptr) : 0;
	return sizeof(float2) * (2 * static_cast<size_t>(a.lds_elems) + tables);
}

// The twiddle tables are copied into LDS after the two line buffers (W_m, then
// W_n^0 .. W_n^m for a half-length real transform); the stages and the real
// pre/post steps read them there (an LDS read instead of a global one per
// butterfly input or bin). fetch_twiddles loads this thread's first two
// entries of each table (all of a table of up to 2 * blockDim.x) before the
// line loads are issued, store_twiddles writes them (and any further entries)
// once the lines are loaded, so the tables cost no extra memory round trip;
// the caller's next barrier, before the first stage, publishes them.
struct twiddle_fetch {
	float2 m[2];
	float2 n[2];
};

__device__ __forceinline__ twiddle_fetch fetch_twiddles(const fft_pass_args& a)
{
	twiddle_fetch f;
	const int m = a.plan.m;
	// unconditional loads (no branch before their use): without a W_n table the
	// second pair re-reads W_m and is never stored
	const float2* tn = a.tw_n ? a.tw_n : a.tw_m;
	const int last_n = a.tw_n ? a.line_m : m - 1;
#pragma unroll
	for (int u = 0; u < 2; ++u) {
		const int i = static_cast<int>(threadIdx.x + u * blockDim.x);
		f.m[u] = a.tw_m[min(i, m - 1)];
		f.n[u] = tn[min(i, last_n)];
	}
	return f;
}

__device__ __forceinline__ void store_table(const float2 (&aFetched)[2], const float2* aTable, int aCount, float2* aLds)
{
#pragma unroll
	for (int u = 0; u < 2; ++u) {
		const int i = static_cast<int>(threadIdx.x + u * blockDim.x);
		if (i < aCount) {
			aLds[i] = aFetched[u];
		}
	}
	for (int i = static_cast<int>(threadIdx.x + 2 * blockDim.x); i < aCount; i += blockDim.x) {
		aLds[i] = aTable[i];
	}
}

// aLds: the table area (W_m at aLds, W_n at aLds + m). Only passes whose
// tables are staged (tables_in_lds) call it.
__device__ __forceinline__ void store_twiddles(const twiddle_fetch& aFetched, const fft_pass_args& a, float2* aLds)
{
	const int m = a.plan.m;
	store_table(aFetched.m, a.tw_m, m, aLds);
	if (a.tw_n) {
		store_table(aFetched.n, a.tw_n, a.line_m + 1, aLds + m);
	}
}

// The complex transform (direction `sign`) of the `count` lines in buffer x
// (scratch y): the radix stages of a.plan, or Bluestein's form when a.chirp is
// set. Returns the buffer holding the result; ends with a barrier.
__device__ __forceinline__ float2* run_line(const fft_pass_args& a, float2* x, float2* y, int count, const float2* tw,
	float sign)
{
	if (a.chirp == nullptr) {
		return sign < 0.0f ? run_stages(a.plan, x, y, count, a.pitch, tw, -1.0f)
		                   : run_stages(a.plan, x, y, count, a.pitch, tw, 1.0f);
	}
	const int n = a.line_m;
	const int big = a.plan.m;
	const bool inverse = sign > 0.0f;
	const fast_div by_big(big), by_n(n);
	// x[j] w[j] for j < n, zero up to M (w = chirp, conjugated for the inverse)
	for (int t = threadIdx.x; t < count * big; t += blockDim.x) {
		const int w = by_big.div(t);
		const int j = t - w * big;
		float2* at = x + w * a.pitch + j;
		if (j < n) {
			const float2 c = a.chirp[j];
			*at = cmul(*at, inverse ? cconj(c) : c);
		} else {
			*at = make_float2(0.0f, 0.0f);
		}
	}
	__syncthreads();
	float2* r = run_stages(a.plan, x, y, count, a.pitch, tw, -1.0f);
	// times the filter spectrum: B for the forward direction; for the inverse the
	// filter is the conjugate sequence, whose spectrum is conj(B[-k])
	for (int t = threadIdx.x; t < count * big; t += blockDim.x) {
		const int w = by_big.div(t);
		const int k = t - w * big;
		const float2 h = inverse ? cconj(a.bspec[k == 0 ? 0 : big - k]) : a.bspec[k];
		float2* at = r + w * a.pitch + k;
		*at = cmul(*at, h);
	}
	__syncthreads();
	float2* q = run_stages(a.plan, r, r == x ? y : x, count, a.pitch, tw, 1.0f);
	for (int t = threadIdx.x; t < count * n; t += blockDim.x) {
		const int w = by_n.div(t);
		const int k = t - w * n;
		const float2 c = a.chirp[k];
		float2* at = q + w * a.pitch + k;
		*at = cmul(*at, inverse ? cconj(c) : c);
	}
	__syncthreads();
	return q;
}

// One workgroup's share of a line pass: the lines of group `aBlock` of `a`.
// `lds` holds two buffers of a.lds_elems complex elements and, unless GENERAL,
// the twiddle tables. GENERAL passes (a.tables_in_lds == 0: lines too long for
// the tables in LDS, or Bluestein's form) read the tables from memory; the
// others read the staged copies, whose LDS address space the compiler sees.
template<int MODE, bool GENERAL = false, int FIXED = 0>
__device__ __forceinline__ void fft_pass_body(const fft_pass_args& a, unsigned aBlock, float2* lds)
{
	float2* A = lds;
	float2* B = lds + a.lds_elems;
	const int m = a.line_m;
	const twiddle_fetch fetched = fetch_twiddles(a);
	float2* const tw_lds = lds + 2 * a.lds_elems;
	const float2* tw_m = GENERAL ? a.tw_m : tw_lds;
	const float2* tw_n = GENERAL ? a.tw_n : tw_lds + a.plan.m; // read only by the half-length real modes
	int o, l0, count;
	decode_group(a, aBlock, o, l0, count);
	const int n = a.n;
	const auto stage_tables = [&] {
		if constexpr (!GENERAL) {
			store_twiddles(fetched, a, tw_lds);
		}
	};
	// the line transform of the bundle: Bluestein-capable only in GENERAL passes;
	// a constant sign in each copy folds the conjugations and (sign i) rotations
	const auto line = [&](float2* x, float2* y, float sign) -> float2* {
		if constexpr (GENERAL) {
			return run_line(a, x, y, count, tw_m, sign);
		} else {
			return sign < 0.0f ? run_plan<FIXED>(a.plan, x, y, count, a.pitch, tw_m, -1.0f)
			                   : run_plan<FIXED>(a.plan, x, y, count, a.pitch, tw_m, 1.0f);
		}
	};
	if constexpr (MODE == PASS_C2C) {
		load_complex_lines(static_cast<const float2*>(a.in), a.gin, o, l0, count, a.load_n, m, A, a.pitch);
		stage_tables();
		__syncthreads();
		float2* r = line(A, B, a.sign);
		if constexpr (GENERAL) {
			if (a.post_n) {
				// four-step: the twiddle between the two line passes
				for (int t = threadIdx.x; t < count * m; t += blockDim.x) {
					const int w = t / m;
					const int k = t - w * m;
					const int j2 = (l0 + w) / a.post_div;
					float2 tw = a.tw_post[static_cast<long long>(j2) * k];
					if (a.sign > 0.0f) {
						tw = cconj(tw);
					}
					r[w * a.pitch + k] = cmul(r[w * a.pitch + k], tw);
				}
				__syncthreads();
			}
		}
		store_complex_lines(static_cast<float2*>(a.out), a.gout, o, l0, count, a.store_n, r, a.pitch);
	} else if constexpr (MODE == PASS_C2C_MULT) {
		load_complex_lines(static_cast<const float2*>(a.in), a.gin, o, l0, count, a.load_n, m, A, a.pitch);
		stage_tables();
		__syncthreads();
		float2* r = run_plan<FIXED>(a.plan, A, B, count, a.pitch, tw_m, -1.0f);
		const long long hbase = o * a.gmult.outer + l0 * a.gmult.line;
		const fast_div by_m(m), by_count(count);
		// (line w, bin k) of element t
		const auto place = [&](int t) {
			if (a.gmult.elem == 1) {
				const int w = by_m.div(t);
				return make_int2(w, t - w * m);
			}
			const int k = by_count.div(t);
			return make_int2(t - k * count, k);
		};
		struct fetched {
			float2 h;
			int slot;
		};
		batched_loop(count * m,
			[&](int t) {
				const int2 wk = place(t);
				return fetched{a.mult[hbase + wk.x * a.gmult.line + wk.y * a.gmult.elem], wk.x * a.pitch + wk.y};
			},
			[&](int, fetched f) {
				const float2 h = a.mult_conj ? cconj(f.h) : f.h;
				r[f.slot] = cscale(cmul(r[f.slot], h), a.mult_scale);
			});
		__syncthreads();
		const float2* q = run_plan<FIXED>(a.plan, r, r == A ? B : A, count, a.pitch, tw_m, 1.0f);
		store_complex_lines(static_cast<float2*>(a.out), a.gout, o, l0, count, a.store_n, q, a.pitch);
	} else if constexpr (MODE == PASS_R2C) {
		load_real_lines(static_cast<const float*>(a.in), a.gin, o, l0, count, n, a.load_n, A, a.pitch);
		stage_tables();
		__syncthreads();
		float2* r = line(A, B, -1.0f);
		if ((n & 1) == 0) {
			float2* x = r == A ? B : A;
			r2c_post(r, x, count, n, a.pitch, tw_n);
			__syncthreads();
			r = x;
		}
		store_complex_lines(static_cast<float2*>(a.out), a.gout, o, l0, count, a.store_n, r, a.pitch);
	} else { // PASS_C2R
		if ((n & 1) == 0) {
			load_complex_lines(static_cast<const float2*>(a.in), a.gin, o, l0, count, a.load_n, m + 1, A, a.pitch);
			stage_tables();
			__syncthreads();
			c2r_pre(A, B, count, n, a.pitch, tw_n);
			__syncthreads();
			const float2* r = line(B, A, 1.0f);
			store_real_lines(static_cast<float*>(a.out), a.gout, o, l0, count, n, a.store_n, r, a.pitch);
		} else {
			load_complex_lines(static_cast<const float2*>(a.in), a.gin, o, l0, count, a.load_n, m, A, a.pitch);
			stage_tables();
			__syncthreads();
			c2r_extend_odd(A, count, n, a.pitch);
			__syncthreads();
			const float2* r = line(A, B, 1.0f);
			store_real_lines(static_cast<float*>(a.out), a.gout, o, l0, count, n, a.store_n, r, a.pitch);
		}
	}
}

} // namespace iane_hip
