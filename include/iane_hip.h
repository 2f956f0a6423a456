/*
 * iane_hip.h — C-ABI of libiane_hip.so, the MI355X (gfx950) backend of the
 * IANE convolution / deconvolution path.
 *
 * Every entry point takes plain pointers, sizes and small descriptor structs.
 * No HIP, torch or C++ types appear in the signatures: a stream is passed as
 * `void*` (a hipStream_t, NULL = the null stream). All work is enqueued on that
 * stream and the call returns without waiting, exactly like the reference's
 * device path returns a completion handle (include/iane/execution/context.hpp:51-59).
 *
 * Status convention: every function returns an int status, 0 on success.
 * The message of the last failure on the calling thread is returned by
 * iane_hip_last_error(). The C++ headers in include/iane map the codes to the
 * exception types the reference throws (std::invalid_argument,
 * std::length_error, std::overflow_error, std::runtime_error).
 *
 * Image layout follows the reference views (include/iane/view/host_image_view.hpp:11-81):
 * a base pointer, `int` extents and `int` BYTE strides per axis, axis 0 fastest.
 */
#ifndef IANE_HIP_H
#define IANE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
	IANE_OK = 0,
	IANE_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
	IANE_ERR_UNSUPPORTED = 2,      /* dtype / rank / size not provided by this backend */
	IANE_ERR_HIP = 3,              /* std::runtime_error("<op>: <hip error>") */
	IANE_ERR_LENGTH = 4,           /* std::length_error (scratch reservation exceeded) */
	IANE_ERR_OVERFLOW = 5          /* std::overflow_error (layout not representable) */
};

/* ---- element types (arithmetic elements of the reference's convolve) ---- */
enum {
	IANE_U8 = 0,
	IANE_U16 = 1,
	IANE_I32 = 2,
	IANE_F32 = 3,
	IANE_F64 = 4,
	IANE_C64 = 5 /* iane::fft_complex {float real, imag}, include/iane/algorithms/fft/complex.hpp:6-9 */
};

/* ---- border rule, include/iane/locator/image_locator.hpp:14 (same order) ---- */
enum { IANE_BORDER_CLAMP = 0, IANE_BORDER_ZERO = 1 };

/* ---- arithmetic of a convolution ---------------------------------------
 * REFERENCE: FP64 fma per tap in kernel storage order, exactly the sum the
 *   reference forms (include/iane/algorithms/convolution.hpp:36-49), then the
 *   reference's output conversion (:51-75). Bit-exact to the reference.
 * FP32: FP32 fma chains (kernel planes from the back, rows and x taps in
 *   storage order); float outputs within ~1e-6 relative.
 * SPLIT_BF16 / SPLIT_FP16: MFMA on operands split into 16-bit parts (BF16:
 *   hi + mid + lo, FP16: hi + lo) with FP32 accumulation (large kernels);
 *   float destinations only; within ~1e-6 relative like FP32.          */
enum {
	IANE_PREC_REFERENCE = 0,
	IANE_PREC_FP32 = 1,
	IANE_PREC_SPLIT_BF16 = 2,
	IANE_PREC_SPLIT_FP16 = 3
};

#define IANE_MAX_RANK 4

/* A strided view: include/iane/view/device_image_view.hpp:11-60 as plain data. */
typedef struct iane_view_desc {
	void* ptr;
	int32_t rank;                        /* 1..IANE_MAX_RANK */
	int32_t dtype;                       /* IANE_U8 .. IANE_C64 */
	int32_t extent[IANE_MAX_RANK];       /* axis 0 fastest */
	int32_t stride_bytes[IANE_MAX_RANK]; /* byte strides, may be negative */
} iane_view_desc;

/* ---- runtime --------------------------------------------------------- */
const char* iane_hip_last_error(void);
int iane_hip_version(void);
int iane_hip_device_count(int* count);
int iane_hip_set_device(int device);
int iane_hip_malloc(void** ptr, size_t bytes);           /* device_image_resource, resource_cuda.hpp:26-41 */
int iane_hip_free(void* ptr);
int iane_hip_host_malloc(void** ptr, size_t bytes);      /* pinned staging for run_on */
int iane_hip_host_free(void* ptr);
int iane_hip_stream_create(void** stream);                /* run_on.hpp:145-153 (non-blocking stream) */
int iane_hip_stream_destroy(void* stream);
int iane_hip_stream_sync(void* stream);                   /* cuda_execution::wait, context.hpp:57-58 */
int iane_hip_device_sync(void);
/* kind: 0 host->host, 1 host->device, 2 device->host, 3 device->device (hipMemcpyKind order) */
int iane_hip_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream);
/* Strided copy between two views of the same element type and extent
 * (copy.hpp:55-108: host<->device transfer of packed or per-plane strided views). */
int iane_hip_copy_view(const iane_view_desc* src, const iane_view_desc* dst, int kind, void* stream);
int iane_hip_memset_async(void* dst, int value, size_t bytes, void* stream);
/* Launch-shape knobs (tile sizes, thread counts, z chunking) chosen by the
 * autotuner (iane_amd/autotune.py); a negative value restores the default.
 * Known keys: "sep3.threads" (256|512), "sep3.chunk", "sep3.min_blocks". */
int iane_hip_tune_set(const char* key, int value);
int iane_hip_tune_get(const char* key, int* value);
/* Events for kernel timing on a given stream. */
int iane_hip_event_create(void** event);
int iane_hip_event_destroy(void* event);
int iane_hip_event_record(void* event, void* stream);
int iane_hip_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---- convolution ----------------------------------------------------
 * Replaces iane::convolve(src, kernel, dst, convolution_options{border},
 * cuda_policy{stream}) for device views, include/iane/algorithms/convolution.hpp:97-117.
 * Kernel extents must be positive and odd (:107-113), ranks equal (:100-101),
 * dst extents equal src extents (include/iane/algorithms/transform.hpp:51-63).
 * Element types: src/kernel/dst each one of U8, U16, I32, F32, F64.     */
/* Workspace: the split-precision MFMA path (SPLIT_BF16 / SPLIT_FP16 on rank <= 3
 * images with float destinations and kernels up to 49 wide along x) needs
 * caller scratch for the source's padded hi/lo planes, the kernel's Toeplitz
 * table and their scales; every other path needs none (0 bytes; workspace may
 * be NULL). The size depends only on the views' shapes and the mode. The
 * workspace must start on a 256-byte boundary; too small a workspace fails
 * with IANE_ERR_LENGTH. The library never allocates on this call (the
 * reference's device work area comes from execution_context scratch,
 * include/iane/execution/prepared.hpp:120-146, algorithms/fft/fft.hpp:136-144). */
int iane_hip_convolve_workspace_bytes(const iane_view_desc* src, const iane_view_desc* kernel,
	const iane_view_desc* dst, int border, int precision, size_t* bytes);
int iane_hip_convolve(const iane_view_desc* src, const iane_view_desc* kernel,
	const iane_view_desc* dst, int border, int precision, void* workspace, size_t workspace_bytes, void* stream);

/* Separable convolution: the kernel is the outer product taps[0] (axis 0)
 * x taps[1] x ... (each a rank-1 F32/F64 view of odd extent). Equal, up to
 * rounding, to iane_hip_convolve with the dense outer-product kernel.
 * Precision REFERENCE (FP64 per-axis passes), FP32, or SPLIT_FP16 (rank-3 F32
 * images, zero border, x extent a multiple of 4: every filter on FP16 matrix
 * instructions with hi + lo operands and FP32 sums, error relative to the
 * largest sample of a 32 x 32 plane region and a K-plane band; other shapes
 * fall back to the FP32 engines). SPLIT_BF16 is refused (UNSUPPORTED).
 * No reference counterpart: the reference always runs the dense kernel. */
int iane_hip_convolve_separable_workspace_bytes(const iane_view_desc* src, const iane_view_desc* taps,
	int n_taps, const iane_view_desc* dst, int border, int precision, size_t* bytes);
/* workspace: as for iane_hip_convolve; only the per-axis path (shapes the
 * single-kernel path does not cover) needs it, for two intermediate images. */
int iane_hip_convolve_separable(const iane_view_desc* src, const iane_view_desc* taps,
	int n_taps, const iane_view_desc* dst, int border, int precision, void* workspace, size_t workspace_bytes,
	void* stream);

/* ---- pointwise expressions: the same-backend copy / transform of lazy views
 * (include/iane/algorithms/copy.hpp:28-47 and transform.hpp:72-92 over the
 * procedural views and view algebra of view/procedural.hpp:93-150 and
 * view/algebra.hpp:52-151) for device images. The caller has no device
 * compiler, so the per-pixel function is a register program: operation k
 * writes register ops[k].dst from registers a, b, c (or loads input a, or
 * sets `value`); registers are read only after an earlier operation wrote
 * them; `result` is stored to dst. Every input has dst's extents; values are
 * FP32 (inputs converted on load; the result assigned to the destination
 * element, integers truncated toward zero and saturated, NaN -> 0).       */
#define IANE_PW_MAX_INPUTS 8
#define IANE_PW_MAX_OPS 32
enum {
	IANE_PW_INPUT = 0,  /* r[dst] = float(inputs[a][p]) */
	IANE_PW_CONST = 1,  /* r[dst] = value */
	IANE_PW_ADD = 2, IANE_PW_SUB = 3, IANE_PW_MUL = 4, IANE_PW_DIV = 5,
	IANE_PW_MIN = 6,    /* std::min(a, b): b < a ? b : a */
	IANE_PW_MAX = 7,    /* std::max(a, b): a < b ? b : a */
	IANE_PW_NEG = 8, IANE_PW_ABS = 9, IANE_PW_SQRT = 10, IANE_PW_EXP = 11, IANE_PW_LOG = 12,
	IANE_PW_SIN = 13, IANE_PW_COS = 14, IANE_PW_LOG1P = 15, IANE_PW_LOG2 = 16, IANE_PW_LOG10 = 17,
	IANE_PW_EXP2 = 18,  /* unary: r[dst] = f(r[a]) */
	IANE_PW_POW = 19,   /* powf(a, b) */
	IANE_PW_LT = 20, IANE_PW_LE = 21, IANE_PW_GT = 22, IANE_PW_GE = 23, IANE_PW_EQ = 24,
	IANE_PW_NE = 25,    /* 1.0f if the comparison holds, else 0.0f */
	IANE_PW_SELECT = 26, /* r[a] != 0 ? r[b] : r[c] */
	IANE_PW_FMA = 27,    /* fmaf(r[a], r[b], r[c]) */
	IANE_PW_COORD = 28,  /* r[dst] = the output element's coordinate along axis a */
	/* r[dst] = point sample of input a (rank 1..3, any extent) at view-local
	 * coordinates r[b] (axis 0), r[c] (axis 1), r[(int)value] (axis 2), measured
	 * from the outer corner: per axis the coordinate is clamped to the edge texel
	 * centres [0.5, extent - 0.5] and the texel floor(coordinate) is read; an
	 * empty axis or a nonfinite coordinate gives 0. The point sampling of
	 * texture_image_view::sample, include/iane/view/texture_image_view.hpp:56-70. */
	IANE_PW_SAMPLE = 29
};
typedef struct iane_pw_op {
	int32_t code;
	int32_t dst, a, b, c;
	float value;
} iane_pw_op;
typedef struct iane_pw_program {
	int32_t n_inputs, n_ops, result;
	iane_view_desc inputs[IANE_PW_MAX_INPUTS];
	iane_pw_op ops[IANE_PW_MAX_OPS];
} iane_pw_program;
int iane_hip_pointwise(const iane_pw_program* program, const iane_view_desc* dst, void* stream);

/* convolve(constant_view(dst, value), kernel, dst, {border}) — the sensitivity
 * step of examples/rl_deconvolution/main.cu:62-63 — without materialising the
 * constant image: every in-range (or clamped) sample is `value`, a zero-border
 * sample outside dst's extent is 0; taps and arithmetic as iane_hip_convolve
 * (precision REFERENCE: the reference's FP64 chain and output conversion;
 * other modes accumulate in FP32). One thread per output: a prepare-time step. */
int iane_hip_convolve_constant(double value, const iane_view_desc* kernel, const iane_view_desc* dst, int border,
	int precision, void* stream);

/* ---- scan / reduce: replace the prepared reduce / reduce_along /
 * inclusive_scan / exclusive_scan / inclusive_scan_along / exclusive_scan_along
 * (include/iane/algorithms/reduce.hpp:7-36, scan.hpp:7-46, built on
 * algorithms/detail/sequence.hpp:255-301) for device views ---------------
 * axis -1: the whole view in storage order (axis 0 fastest); otherwise
 * independent lines along `axis`. The accumulator and destination element
 * type is `init_type` and `init` points to one value of it: it is applied once
 * per line (before the first element) and is the result of an empty
 * reduction. Shapes follow sequence.hpp:262-293: scans keep the source
 * extents; a whole reduction (or one of a rank-1 view) writes a rank-1 view
 * of extent 1; reduce along an axis removes that axis. The operator is one of
 * IANE_OP_*, applied as op(accumulated, next) in the accumulator type
 * (integer sums and products wrap; min and max follow std::min / std::max).
 * Floating-point grouping may vary, as the reference allows. */
enum { IANE_SEQ_INCLUSIVE = 0, IANE_SEQ_EXCLUSIVE = 1, IANE_SEQ_REDUCE = 2 };
enum { IANE_OP_SUM = 0, IANE_OP_PRODUCT = 1, IANE_OP_MIN = 2, IANE_OP_MAX = 3 };
typedef struct iane_sequence_plan iane_sequence_plan;
/* sequence_algorithm::prepare (sequence.hpp:262-300): validates the prototype
 * views' shapes and picks the work partition. */
int iane_hip_sequence_plan_create(iane_sequence_plan** plan, int kind, int axis, int op, int init_type,
	const void* init, const iane_view_desc* src, const iane_view_desc* dst);
/* scratch_bytes() (sequence.hpp:201-211): line totals and carries. */
int iane_hip_sequence_plan_workspace_bytes(const iane_sequence_plan* plan, size_t* bytes);
/* run (sequence.hpp:213-235): views must match the prepared shape. */
int iane_hip_sequence_run(const iane_sequence_plan* plan, const iane_view_desc* src, const iane_view_desc* dst,
	void* workspace, void* stream);
int iane_hip_sequence_plan_destroy(iane_sequence_plan* plan);

/* ---- FFT: replaces the prepared fft_forward / fft_inverse / rfft / irfft
 * (include/iane/algorithms/fft/fft.hpp:196-231, rfft.hpp:107-161) ---------
 * Stockham FFT, FP32, packed buffers, rank 1..3 (the reference is 2-D only,
 * fft.hpp:22). Forward uses the negative exponent, the inverse is
 * unnormalised. extent = the complex extents (C2C) or the REAL extents
 * (R2C/C2R); the spectrum of R2C/C2R is {n0/2+1, n1, n2} (layout.hpp:15-21).
 * C2R leaves its input intact (rfft.hpp:72-83). Lengths: any whose prime
 * factors are <= 64, each line <= 4095 points (the line is held in LDS).
 * Buffers must be complex-aligned and disjoint (fft.hpp:20-66).           */
enum { IANE_FFT_C2C_FORWARD = 0, IANE_FFT_C2C_INVERSE = 1, IANE_FFT_R2C = 2, IANE_FFT_C2R = 3 };
typedef struct iane_fft_plan iane_fft_plan;
int iane_hip_fft_plan_create(iane_fft_plan** plan, int rank, const int32_t* extent, int kind);
/* scratch the caller supplies to exec (execution_context scratch, prepared.hpp:120-146) */
int iane_hip_fft_plan_workspace_bytes(const iane_fft_plan* plan, size_t* bytes);
int iane_hip_fft_exec(iane_fft_plan* plan, const void* in, void* out, void* workspace, void* stream);
int iane_hip_fft_plan_destroy(iane_fft_plan* plan);

/* ---- Richardson–Lucy / Wiener pointwise steps: the functors of
 * examples/rl_deconvolution/main.cu:11-49 and examples/wiener/main.cu:21-32
 * with the same float arithmetic. count = elements; packed buffers.       */
/* estimate = observed > eps ? observed : eps                  (main.cu:12-19) */
int iane_hip_rl_initial(const float* observed, float* estimate, int64_t count, float epsilon, void* stream);
/* ratio = observed / max(predicted, eps) (main.cu:22-29); ratio may alias predicted.
 * ll_out != NULL: *ll_out (device double) = sum(y log max(p, eps) - p), the
 * Poisson log-likelihood without log(y!) (no reference counterpart): FP32
 * terms summed as a compensated FP32 pair, widened to double at the end; the
 * per-block partial sums go to `workspace` (iane_hip_rl_ratio_workspace_bytes,
 * 256-byte aligned; unused and may be NULL when ll_out is NULL).            */
int iane_hip_rl_ratio_workspace_bytes(int64_t count, size_t* bytes);
int iane_hip_rl_ratio(const float* observed, const float* predicted, float* ratio, int64_t count,
	float epsilon, double* ll_out, void* workspace, size_t workspace_bytes, void* stream);
/* iane_hip_rl_ratio without a workspace argument: the log-likelihood partial
 * sums live in a static device array of the library (1024 compensated sums,
 * 8 KiB per GPU, held until the process ends), so the call allocates nothing.
 * No reference counterpart: it is for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Results
 * are the bits of iane_hip_rl_ratio. Every element count fits (the ratio
 * kernel never uses more than 1024 blocks). All calls on one GPU share the
 * array: calls that overlap on one GPU race.                                */
int iane_hip_rl_ratio_noalloc(const float* observed, const float* predicted, float* ratio, int64_t count,
	float epsilon, double* ll_out, void* stream);
/* estimate *= (sens > 0 ? corr / max(sens, eps) : 0)           (main.cu:32-49,74-77) */
int iane_hip_rl_update(float* estimate, const float* correction, const float* sensitivity,
	int64_t count, float epsilon, void* stream);
/* spectrum = spectrum * conj(H) / (|H|^2 + noise), factor in double (wiener/main.cu:21-32);
 * spectrum/response: IANE_C64 buffers of count elements. */
int iane_hip_wiener_apply(void* spectrum, const void* response, int64_t count, float noise, void* stream);
/* out = (in, 0) as IANE_C64 (wiener/main.cu:13-19, real_sample) */
int iane_hip_real_to_complex(const float* in, void* out, int64_t count, void* stream);
/* out = real(in) * scale (wiener/main.cu:35-41, the normalisation of encode_sample) */
int iane_hip_complex_real_scale(const void* in, float* out, int64_t count, float scale, void* stream);

/* ---- Richardson–Lucy prepared algorithm: prepare_rl + its closure,
 * examples/rl_deconvolution/main.cu:51-81, rank 2 or 3, FP32 images, zero border.
 * IANE_RL_DIRECT: direct convolutions with the PSF and its flip (main.cu:106-107)
 *   at `precision`; REFERENCE precision is bit-identical to the reference loop.
 * IANE_RL_FFT: convolutions on a zero-padded domain (n >= N + K/2 per axis,
 *   equal to the zero-border convolution up to FP32 rounding).
 * The plan owns the PSF, its flip / spectrum and the sensitivity H^T 1
 * (main.cu:62-63); the per-run temporaries live in the caller's workspace.  */
enum { IANE_RL_DIRECT = 0, IANE_RL_FFT = 1 };
typedef struct iane_rl_plan iane_rl_plan;
/* extent: image extents (axis 0 fastest); psf: rank-matching view, odd extents */
int iane_hip_rl_plan_create(iane_rl_plan** plan, int rank, const int32_t* extent,
	const iane_view_desc* psf, int method, int precision, float epsilon, void* stream);
int iane_hip_rl_plan_workspace_bytes(const iane_rl_plan* plan, size_t* bytes);
/* init != 0: estimate = max(observed, eps) first (main.cu:68); otherwise the
 * estimate buffer is the warm start. ll_history (device, `iterations` doubles,
 * may be NULL): log-likelihood of each iteration's prediction H x_i.        */
int iane_hip_rl_run(iane_rl_plan* plan, const float* observed, float* estimate, int iterations,
	int init, void* workspace, double* ll_history, void* stream);
int iane_hip_rl_plan_destroy(iane_rl_plan* plan);
/* out = H in (the plan's zero-border convolution with the PSF); workspace as for run */
int iane_hip_rl_plan_blur(iane_rl_plan* plan, const float* in, float* out, void* workspace, void* stream);

/* ---- Batched Richardson–Lucy (no reference counterpart) ----------------
 * `count` volumes of one rank with their own extents (extents[v*rank + a])
 * and PSFs, FFT method, FP32. Each pass of an iteration is one launch over
 * the lines of every volume; with use_graph != 0 the whole run is captured
 * into a hipGraph on first use and replayed (same iterations/init/ll/
 * workspace/stream) with one graph launch; the stream must not be NULL then.
 * Per volume the result equals iane_hip_rl_run with IANE_RL_FFT bit for bit.
 * observed/estimates: host arrays of `count` device pointers (packed f32).
 * ll_history: device doubles, count x iterations (row v = volume v), or NULL. */
typedef struct iane_rl_batch iane_rl_batch;
int iane_hip_rl_batch_create(iane_rl_batch** batch, int count, int rank, const int32_t* extents,
	const iane_view_desc* psfs, float epsilon, void* stream);
int iane_hip_rl_batch_workspace_bytes(const iane_rl_batch* batch, size_t* bytes);
int iane_hip_rl_batch_run(iane_rl_batch* batch, const float* const* observed, float* const* estimates,
	int iterations, int init, void* workspace, double* ll_history, int use_graph, void* stream);
int iane_hip_rl_batch_destroy(iane_rl_batch* batch);

/* ---- Learned RL warm start: kernels of a small 3-D CNN (no reference counterpart) ----
 * Activations channels-last [z][y][x][c] (BF16 unless stated), weights BF16
 * [cout][27][cin], tap = (dz*3 + dy)*3 + dx, zero padding. FP32 accumulation. */
/* out = [relu](bias + 3x3x3 cross-correlation); out f32 or bf16; cout <= 64 */
int iane_hip_cnn_conv3d(const void* in, const void* weights, const float* bias, void* out, int out_bf16, int relu,
	int nx, int ny, int nz, int cin, int cout, void* stream);
/* wt[ci][t][co] = w[co][26 - t][ci]: the input gradient is conv3d with wt */
int iane_hip_cnn_transpose_weights(const void* w, void* wt, int cin, int cout, void* stream);
/* dw[co][t][ci] += sum_p dy[p][co] in[p + d(t) - 1][ci], dbias[co] += sum_p dy[p][co] (FP32, accumulating) */
int iane_hip_cnn_wgrad(const void* in, const float* dy, float* dw, float* dbias, int nx, int ny, int nz, int cin,
	int cout, void* stream);
int iane_hip_cnn_relu_backward(const void* y, float* dy, int64_t n, void* stream);
int iane_hip_cnn_pool2(const void* in, void* out, int nx, int ny, int nz, int c, void* stream);
int iane_hip_cnn_pool2_backward(const float* dy, float* dx, int nx, int ny, int nz, int c, int accumulate, void* stream);
int iane_hip_cnn_up2_add(const void* coarse, const void* skip, void* out, int nx, int ny, int nz, int c, void* stream);
int iane_hip_cnn_up2_backward(const float* dfine, float* dcoarse, int nx, int ny, int nz, int c, void* stream);
int iane_hip_cnn_to_bf16(const float* in, void* out, int64_t n, float scale, void* stream);
int iane_hip_cnn_adam(float* w, const float* g, float* m, float* v, void* w_bf16, int64_t n, float lr, float beta1,
	float beta2, float eps, int step, void* stream);
int iane_hip_cnn_residual_mse(const float* u, const float* r, const float* target, float* pred, float* grad, int64_t n,
	double* partials, int* blocks_used, void* stream);
/* x = max(scale * (u + r), eps) */
int iane_hip_cnn_warm_start(const float* u, const float* r, float* x, int64_t n, float scale, float eps, void* stream);
/* training-data synthesis on the device: blobs (cz, cy, cx, sigma, amp) x count, Poisson
 * sampling (inversion / PTRS over Philox4x32-10), max reduction, scaling */
int iane_hip_synth_blobs(float* out, int nx, int ny, int nz, float background, const float* blobs, int count,
	void* stream);
int iane_hip_synth_poisson(const float* lambda, float* out, int64_t n, unsigned seed, void* stream);
int iane_hip_max(const float* in, int64_t n, float* partial, float* out, void* stream);
int iane_hip_scale(float* x, int64_t n, float scale, void* stream);

/* Test hooks: the untiled one-thread-per-output kernel and the per-axis
 * separable passes, so tests can compare every path against the oracle. */
int iane_hip_convolve_untiled(const iane_view_desc* src, const iane_view_desc* kernel,
	const iane_view_desc* dst, int border, int precision, void* stream);
int iane_hip_convolve_separable_passes_workspace_bytes(const iane_view_desc* src, const iane_view_desc* taps,
	int n_taps, const iane_view_desc* dst, int border, int precision, size_t* bytes);
int iane_hip_convolve_separable_passes(const iane_view_desc* src, const iane_view_desc* taps,
	int n_taps, const iane_view_desc* dst, int border, int precision, void* workspace, size_t workspace_bytes,
	void* stream);

#ifdef __cplusplus
}
#endif

#endif /* IANE_HIP_H */
