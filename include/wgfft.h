/*
 * wgfft.h -- C-ABI of the MI355X-native batched complex64 Stockham FFT.
 *
 * Drop-in boundary for wgls-rs-fft's GpuFft path.  Data is interleaved f32
 * (re, im) -- the layout of num_complex::Complex<f32>, so a Rust
 * Vec<Complex<f32>> (or a flattened batch of them) is passed as float*.
 * Transform b of a batch starts at complex offset b * n.
 *
 * Conventions
 *   forward  X_k = sum_n x_n exp(-2 pi i k n / N)           (unnormalised)
 *   inverse  x_n = (1/N) sum_k X_k exp(+2 pi i k n / N)      (1/N scaled)
 *   Return value: WGFFT_OK (0) or a negative WGFFT_ERR_* code; the message
 *   of the last failure on the calling thread is wgfft_last_error().
 *   Nothing aborts or throws across this boundary.
 *   Host entry points are synchronous (they use the handle's own stream).
 *   Device entry points take device pointers (16-byte aligned) and a
 *   hipStream_t passed as void* (NULL = HIP's default null stream) and are
 *   stream-ordered: they return once the work is enqueued.
 *   A handle may be used by one thread at a time (calls are serialised by an
 *   internal mutex); distinct handles are independent.  Device calls of one
 *   handle on different streams may run concurrently, except that those using
 *   the handle's scratch buffers (global passes for n > 16384, strided
 *   layouts, VJP reductions) are ordered after one another: while they all
 *   use one stream, by that stream; the first such call on a second stream
 *   waits once for the device (hipDeviceSynchronize; inside a stream capture
 *   it fails instead, so make one call on a new stream before capturing on
 *   it) and from then on each such call records an event that the next one
 *   on another stream waits for.  A mixed-length plan has one scratch area
 *   of its own: its calls on different streams must be ordered by the
 *   caller.
 */
#ifndef WGFFT_H
#define WGFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wgfft_handle wgfft_t;
typedef struct wgfft_ragged_plan wgfft_ragged_t;

enum {
    WGFFT_OK = 0,
    WGFFT_ERR_INVALID_LENGTH = -1,  /* "Transform length must be a non-zero power of two" */
    WGFFT_ERR_RAGGED_BATCH = -2,    /* "All input vectors in a batch must have the same length" */
    WGFFT_ERR_NO_DEVICE = -3,       /* no HIP device / device unusable */
    WGFFT_ERR_OUT_OF_MEMORY = -4,   /* device or host allocation failed or exceeded the cap */
    WGFFT_ERR_HIP = -5,             /* a HIP runtime call failed */
    WGFFT_ERR_INVALID_ARGUMENT = -6 /* null pointer, misaligned buffer, bad direction, ... */
};

#define WGFFT_FORWARD (-1)
#define WGFFT_INVERSE (+1)

/* ---- device and handle ------------------------------------------------ */

/* 1 if a HIP device is usable, else 0.
 * Replaces GpuFft::is_gpu_available (src/lib.rs:381-389). */
int wgfft_device_available(void);

/* Create a handle on HIP device `device` (its own stream, plan cache and
 * scratch).  Fails with WGFFT_ERR_NO_DEVICE when there is none.
 * Replaces GpuFft::new (src/lib.rs:285-330). */
int wgfft_create(wgfft_t** out, int device);

/* Release a handle and everything it owns (NULL is a no-op).
 * Replaces the drop of GpuFft and its SizeCache map (src/lib.rs:75-82). */
void wgfft_destroy(wgfft_t* h);

/* Message of the last failure on this thread ("" if none).  Replaces the
 * Box<dyn Error> strings of src/lib.rs:508,530. */
const char* wgfft_last_error(void);

/* Executor name.  Replaces FftExecutor::name (src/lib.rs:85-87). */
const char* wgfft_name(void);

/* ---- validation (pure host logic, no device needed) ------------------- */

/* Validate a batch of `count` signal lengths the way GpuFft does before any
 * device work: empty batch -> WGFFT_OK with *n_out = 0; every length equal
 * to lens[0] else WGFFT_ERR_RAGGED_BATCH; each a non-zero power of two else
 * WGFFT_ERR_INVALID_LENGTH (checked per signal in order, so the first
 * offending signal decides the code).
 * Replaces transform_batch_internal's checks (src/lib.rs:522-533) and
 * validate_input_size (src/lib.rs:506-511). */
int wgfft_validate_lengths(const uint64_t* lens, uint64_t count, uint64_t* n_out);

/* ---- batched transforms ------------------------------------------------ */

/* Batched FFT of host buffers: in/out hold batch*n complex64 (may alias).
 * dir = WGFFT_FORWARD or WGFFT_INVERSE (inverse is 1/N scaled).  batch == 0
 * is a no-op.  Replaces GpuFft::fft / GpuFft::ifft (src/lib.rs:433-438,
 * 483-488) and transform_batch_internal (src/lib.rs:517-568). */
int wgfft_exec_host(wgfft_t* h, const float* in, float* out, uint64_t n, uint64_t batch, int dir);

/* The same with every signal in a buffer of its own, as GpuFft::fft takes
 * and returns them (&[Vec<Complex<f32>>] in, Vec<Vec<Complex<f32>>> out):
 * in_rows[b] and out_rows[b] each point at n complex64 values.  The staging
 * copies gather and scatter the rows, so the batch is never flattened into
 * one buffer.  Rows must not overlap one another; out_rows[b] may equal
 * in_rows[b].  Replaces transform_batch_internal's flatten and re-boxing
 * (src/lib.rs:539-544, 563-565) together with the transform. */
int wgfft_exec_host_rows(wgfft_t* h, const float* const* in_rows, float* const* out_rows, uint64_t n, uint64_t batch,
                         int dir);

/* Same on device buffers, enqueued on `stream`.  in == out is allowed.
 * Replaces execute_compute_pass (src/lib.rs:582-628) with data resident. */
int wgfft_exec_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir,
                      void* stream);

/* The same transform allocating nothing: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  The intermediates of lengths above 16384 live in a static
 * 64 MiB device array that every handle on a GPU shares (each GPU the library
 * runs on holds it until the process ends): at most 2^23 for a two-pass plan,
 * 2^22 for three passes or more; larger lengths fail with
 * WGFFT_ERR_INVALID_ARGUMENT, and larger batches run in chunks that fit.
 * The handle's plan for n must exist (a plan allocates its twiddle tables):
 * make it first with wgfft_plan_describe or any other call of length n, else
 * WGFFT_ERR_INVALID_ARGUMENT.  Calls of this entry point that overlap on one
 * GPU (other streams, other handles) race on the shared array: order them.
 * The array keeps the last call's intermediates; nothing in it needs to
 * start at zero. */
int wgfft_exec_device_noalloc(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir,
                              void* stream);

/* Batched FFT of strided device layouts: element j of signal b is at complex
 * offset b*idist + j*istride of d_in (b*odist + j*ostride of d_out); 8-byte
 * aligned buffers; the output layout must not overlap itself.  The reference
 * lists configurable strides and distances as future work (REQUIREMENTS.md:14,
 * "Stride Management").  Unit stride with distance n is the plain device call.
 * Padded rows -- unit strides, even distances of at least n, 16-byte aligned
 * buffers, input and output disjoint (or the same rows, in place) -- are
 * transformed in one pass for n <= 16384.  Other layouts are gathered into a
 * contiguous scratch buffer, transformed in place there and scattered, on the
 * caller's stream (adjacent signals, distance 1 -- the columns of a matrix --
 * through 32 x 32 tiles that keep both sides coalesced). */
int wgfft_exec_strided_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch,
                              uint64_t istride, uint64_t idist, uint64_t ostride, uint64_t odist, int dir,
                              void* stream);

/* Windowed forward Y_b = FFT(w * x_b), w real (n floats) shared by the batch,
 * multiplied at load (no extra pass).  Replaces apply_hann_window + fft
 * (tests/test_utils.rs:39-50, examples/windowed_fft.rs:14-23). */
int wgfft_exec_windowed_host(wgfft_t* h, const float* x, const float* w, float* y, uint64_t n, uint64_t batch);
int wgfft_exec_windowed_device(wgfft_t* h, const float* d_x, const float* d_w, float* d_y, uint64_t n,
                               uint64_t batch, void* stream);

/* ---- differentiation of Y = FFT(w * x) (no counterpart in the reference) */

/* Vector-Jacobian product for a real loss, gradients in the
 * (dL/dRe + i dL/dIm) convention:  gu = F^H gY,  gx = w * gu,
 * gw_n = sum_b Re(conj(gu_bn) x_bn).  d_gw (n floats) is overwritten. */
int wgfft_vjp_windowed_device(wgfft_t* h, const float* d_x, const float* d_w, const float* d_gy, float* d_gx,
                              float* d_gw, uint64_t n, uint64_t batch, void* stream);

/* Fused forward + adjoint for L = sum_b sum_k m_k |Y_bk|^2, Y = FFT(w * x):
 * one pass over x writes gx (and Y when d_y != NULL), gw (n floats) and the
 * scalar loss (d_loss, one float).  The mask m is real (n floats). */
int wgfft_fwd_adjoint_masked_device(wgfft_t* h, const float* d_x, const float* d_w, const float* d_mask,
                                    float* d_y, float* d_gx, float* d_gw, float* d_loss, uint64_t n,
                                    uint64_t batch, void* stream);

/* Spectral filter y = IFFT(FFT(x + noise) * s), 1/N scaled, s real (n
 * floats): the reference pipeline's Gaussian noise (when sigma != 0; the
 * noise of wgfft_add_noise_device), FFT, spectral Hann window and IFFT + 1/N
 * stages (examples/parallel_pipeline.rs:29-105, run at :586-642) in one pass
 * over memory for N = 512 .. 16384, else on device as noise, forward,
 * multiply and inverse launches.  d_x == d_y is allowed; d_x is not changed
 * unless it is d_y. */
int wgfft_filter_device(wgfft_t* h, const float* d_x, const float* d_s, float* d_y, uint64_t n, uint64_t batch,
                        uint32_t seed, float sigma, void* stream);

/* The pipeline demo's Gaussian noise stage (examples/parallel_pipeline.rs:
 * 29-63, NOISE_WGSL), in place on d_x (8-byte aligned): sample t of signal b
 * gets sigma * sqrt(-2 ln u1) * (cos, sin)(2 pi u2) from the demo's hash of
 * (t, b, seed); the demo uses sigma = 0.1, seed 0. */
int wgfft_add_noise_device(wgfft_t* h, float* d_x, uint64_t n, uint64_t batch, uint32_t seed, float sigma,
                           void* stream);

/* ---- mixed-length batches (no counterpart in the reference: it needs one
 *      GpuFft call per length, src/lib.rs:527-531) ---------------------- */

/* Plan `count` transforms of lengths 2^log2n[i] (6 <= log2n[i] <= 16) stored
 * contiguously at complex offsets offsets[i] (host arrays; offsets even).
 * The plan uploads a device work queue; it stays valid until destroyed with
 * wgfft_ragged_destroy or until its handle is destroyed: wgfft_destroy also
 * frees every plan still open on the handle, after which such a plan must
 * not be used or passed to wgfft_ragged_destroy. */
int wgfft_ragged_create(wgfft_t* h, const uint64_t* offsets, const uint32_t* log2n, uint64_t count,
                        wgfft_ragged_t** out);
void wgfft_ragged_destroy(wgfft_ragged_t* p);

/* One persistent launch over the whole queue (mode 0), the same launch
 * captured once into a hipGraph and replayed (mode 1), or one launch per
 * distinct length (mode 2, the comparison baseline). */
int wgfft_ragged_exec_device(wgfft_ragged_t* p, const float* d_in, float* d_out, int dir, int mode, void* stream);

/* Total complex samples covered by a mixed-length plan (0 for NULL). */
uint64_t wgfft_ragged_samples(const wgfft_ragged_t* p);

/* The plan's kernel choice (tile prefetch, workgroups), made by timing on
 * its first out-of-place call. */
int wgfft_ragged_describe(const wgfft_ragged_t* p, char* desc, uint64_t desc_len);

/* ---- measurement and tuning ------------------------------------------- */

/* Device-only timing of wgfft_exec_device on resident buffers: `warmup`
 * synchronised calls, then `iters` back-to-back calls bracketed by HIP
 * events on the handle's stream; *sec_per_iter = elapsed / iters.
 * Replaces GpuFftTrait::benchmark_gpu_only (src/lib.rs:48-68,109-143). */
int wgfft_bench_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir,
                       int warmup, int iters, double* sec_per_iter);

/* Time every compiled kernel variant for length n on a scratch batch of
 * `batch` transforms and make the fastest the handle's choice for n, for
 * every batch size (an untuned plan of 512 .. 2048 runs batches of at most
 * 4 transforms per CU one transform per workgroup; a tuned or pinned one
 * does not).  Writes the chosen variant's description into desc (if
 * non-NULL). */
int wgfft_autotune(wgfft_t* h, uint64_t n, uint64_t batch, char* desc, uint64_t desc_len);

/* Description of the variant the handle currently uses for length n (and,
 * for an untuned plan of 512 .. 2048, the one its small batches use). */
int wgfft_plan_describe(wgfft_t* h, uint64_t n, char* desc, uint64_t desc_len);

/* Number of compiled single-kernel (LDS-resident) variants for length n;
 * 0 for lengths served by the register kernel or by global passes. */
int wgfft_variant_count(uint64_t n);

/* Make compiled variant `index` (0 <= index < wgfft_variant_count(n)) the
 * handle's choice for length n, as a tuning run would; writes its
 * description into desc (if non-NULL).  Lets callers pin a variant and tests
 * check every one. */
int wgfft_select_variant(wgfft_t* h, uint64_t n, int index, char* desc, uint64_t desc_len);

/* Global passes of the plan for length n (0 for single-kernel lengths), and
 * the compiled candidates for pass `pass` (radix, E, G, LDS-staged or direct
 * I/O, with or without the register prefetch of the next tile). */
int wgfft_pass_count(uint64_t n);
int wgfft_pass_variant_count(uint64_t n, int pass);

/* Make candidate `index` the kernel of global pass `pass` of length n for
 * this handle; writes the plan description into desc (if non-NULL). */
int wgfft_select_pass_variant(wgfft_t* h, uint64_t n, int pass, int index, char* desc, uint64_t desc_len);

/* Give length n (> 16384) the radix split 2^log2r[0] x ... x 2^log2r[passes-1]
 * (first pass first), each pass on its default kernel, as a tuning run would;
 * the radices must multiply to n and each needs a compiled kernel at its
 * stride.  Writes the plan description into desc (if non-NULL).  Lets
 * callers pin a split and tests run three-pass chains at small lengths. */
int wgfft_select_split(wgfft_t* h, uint64_t n, const int* log2r, int passes, char* desc, uint64_t desc_len);

/* Number of compiled fused forward+adjoint / VJP variants for length n (0 when
 * the length is served by the unfused path), and: make variant `index` the
 * handle's choice for length n, as a tuning run would, writing its
 * description into desc (if non-NULL).  No reference counterpart: the
 * adjoint has none (the reference has no differentiation path). */
int wgfft_adjoint_variant_count(uint64_t n);
int wgfft_select_adjoint_variant(wgfft_t* h, uint64_t n, int index, char* desc, uint64_t desc_len);

/* Fill nfloats floats with i.i.d. U[-1,1) = (u32 >> 8) * 2^-23 - 1 drawn
 * from PCG32 seeded through SplitMix64(seed) (input synthesis, host). */
void wgfft_synth_uniform(float* out, uint64_t nfloats, uint64_t seed);

/* Cap on the bytes a handle may hold in device scratch (0 = no cap); an
 * allocation beyond it fails with WGFFT_ERR_OUT_OF_MEMORY.  Host-side knob
 * for exercising the out-of-memory path. */
void wgfft_set_alloc_cap(uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* WGFFT_H */
