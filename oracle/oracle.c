/*
 * oracle.c -- CPU checker for the batched complex64 Stockham FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, the smoke check and the bench's cpu_baseline
 * leg do, and only as the checker (or, for cpu_baseline, as the timed CPU
 * restatement of the reference algorithm).
 *
 * Contents
 *   1. Input synthesis: SplitMix64-seeded PCG32, uniform[-1,1) and Box-Muller
 *      Gaussian, the multi-sinusoid large-N signal, ragged length tables,
 *      window + mask for the windowed/VJP workload.
 *   2. Restatement of the reference f32 algorithm (wgls-rs-fft GpuFft):
 *        stage plan           src/lib.rs:725-732, 836-861
 *        f32-angle twiddles   src/lib.rs:761-769
 *        R4 stage             src/shaders.rs:18-62
 *        R2 stage             src/shaders.rs:79-105
 *        ping-pong A<->B      src/lib.rs:863-889, result_in_b :899
 *        inverse conj pre     src/lib.rs:571-579
 *        inverse conj + 1/N   src/lib.rs:661-669
 *      Built with -ffp-contract=off so every f32 operation rounds as written.
 *   3. FP64 oracle: direct DFT with exact index reduction, iterative radix-2
 *      FFT with FP64 twiddles, windowed forward, analytic VJP, masked loss.
 *
 * Pinning (tests/test_oracle.py): numpy pocketfft FP64 golden vectors,
 * closed-form known answers (impulse, constant, complex exponential, shifted
 * impulse), Parseval, linearity, tiny-N brute force, the reference's own test
 * inputs and tolerances (tests/gpu_fft_test.rs:8-40, tests/ifft_test.rs:22-107,
 * tests/windowed_fft_test.rs:10-133), PCG32/SplitMix64 published outputs, and
 * FP64 central finite differences for the VJP.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* 1. PRNG and synthetic inputs                                        */
/* ------------------------------------------------------------------ */

uint64_t oracle_splitmix64_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

typedef struct { uint64_t state, inc; } pcg32_t;

static uint32_t pcg32_next(pcg32_t *r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

static void pcg32_srandom(pcg32_t *r, uint64_t initstate, uint64_t initseq) {
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r->state += initstate;
    pcg32_next(r);
}

/* PCG32 seeded directly (for the published pcg32-demo known answers). */
void oracle_pcg32_raw(uint64_t initstate, uint64_t initseq, uint32_t *out, uint64_t count) {
    pcg32_t r;
    pcg32_srandom(&r, initstate, initseq);
    for (uint64_t i = 0; i < count; ++i) out[i] = pcg32_next(&r);
}

static void seed_stream(pcg32_t *r, uint64_t seed) {
    uint64_t sm = seed;
    uint64_t st = oracle_splitmix64_next(&sm);
    uint64_t sq = oracle_splitmix64_next(&sm);
    pcg32_srandom(r, st, sq);
}

static inline double u01_24(uint32_t u) { return (double)(u >> 8) * (1.0 / 16777216.0); }

static inline double gauss(pcg32_t *r) {
    double u1 = ((double)(pcg32_next(r) >> 8) + 1.0) * (1.0 / 16777216.0);
    double u2 = u01_24(pcg32_next(r));
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* nfloats values uniform in [-1,1): (u32 >> 8) * 2^-23 - 1, exact in f32. */
void oracle_synth_uniform(float *out, uint64_t nfloats, uint64_t seed) {
    pcg32_t r;
    seed_stream(&r, seed);
    for (uint64_t i = 0; i < nfloats; ++i)
        out[i] = (float)((double)(pcg32_next(&r) >> 8) * (1.0 / 8388608.0) - 1.0);
}

void oracle_synth_gauss(double *out, uint64_t count, uint64_t seed) {
    pcg32_t r;
    seed_stream(&r, seed);
    for (uint64_t i = 0; i < count; ++i) out[i] = gauss(&r);
}

/*
 * Large-N signal: transform b holds 32 unit sinusoids with f_j = U[0,1)*N and
 * phase phi_j = 2*pi*U[0,1) drawn from stream (seed + b), plus
 * 0.1*(g_re + i g_im) where the Gaussians of sample block c (65536 samples)
 * come from stream (seed + b + (c + 1) * 2^32).  The phase argument f_j*n is
 * reduced mod N in FP64 at each block start and advanced inside a block by
 * the FP64 rotation exp(2 pi i f_j / N).  Blocks are generated in parallel.
 */
#define SIN_BLOCK 65536ull
typedef struct {
    float *out;
    uint64_t n, seed, b;
    const double *f, *ph;
    uint64_t c0, c1;
} sin_job_t;

static void *sin_worker(void *arg) {
    sin_job_t *j = (sin_job_t *)arg;
    uint64_t n = j->n;
    double cr[32], ci[32], rr[32], ri[32];
    for (int t = 0; t < 32; ++t) {
        double a = 2.0 * M_PI * j->f[t] / (double)n;
        rr[t] = cos(a);
        ri[t] = sin(a);
    }
    for (uint64_t c = j->c0; c < j->c1; ++c) {
        uint64_t i0 = c * SIN_BLOCK, i1 = i0 + SIN_BLOCK < n ? i0 + SIN_BLOCK : n;
        pcg32_t r;
        seed_stream(&r, j->seed + j->b + ((c + 1) << 32));
        for (int t = 0; t < 32; ++t) {
            double a = fmod(j->f[t] * (double)i0, (double)n);
            double th = 2.0 * M_PI * a / (double)n + j->ph[t];
            cr[t] = cos(th);
            ci[t] = sin(th);
        }
        float *o = j->out + 2 * n * j->b;
        for (uint64_t i = i0; i < i1; ++i) {
            double re = 0.0, im = 0.0;
            for (int t = 0; t < 32; ++t) {
                re += cr[t];
                im += ci[t];
                double nr = cr[t] * rr[t] - ci[t] * ri[t];
                ci[t] = cr[t] * ri[t] + ci[t] * rr[t];
                cr[t] = nr;
            }
            double gr = gauss(&r), gi = gauss(&r);
            o[2 * i] = (float)(re + 0.1 * gr);
            o[2 * i + 1] = (float)(im + 0.1 * gi);
        }
    }
    return NULL;
}

void oracle_synth_sinusoids(float *out, uint64_t n, uint64_t batch, uint64_t seed, int threads) {
    double *f = (double *)malloc(sizeof(double) * 32 * batch);
    double *ph = (double *)malloc(sizeof(double) * 32 * batch);
    for (uint64_t b = 0; b < batch; ++b) {
        pcg32_t r;
        seed_stream(&r, seed + b);
        for (int t = 0; t < 32; ++t) {
            f[32 * b + t] = u01_24(pcg32_next(&r)) * (double)n;
            ph[32 * b + t] = 2.0 * M_PI * u01_24(pcg32_next(&r));
        }
    }
    uint64_t blocks = (n + SIN_BLOCK - 1) / SIN_BLOCK;
    uint64_t jobs_n = blocks * batch;
    if (threads < 1) threads = 1;
    sin_job_t *jobs = (sin_job_t *)calloc(jobs_n, sizeof(sin_job_t));
    for (uint64_t b = 0, k = 0; b < batch; ++b)
        for (uint64_t c = 0; c < blocks; ++c, ++k) {
            jobs[k].out = out; jobs[k].n = n; jobs[k].seed = seed; jobs[k].b = b;
            jobs[k].f = f + 32 * b; jobs[k].ph = ph + 32 * b; jobs[k].c0 = c; jobs[k].c1 = c + 1;
        }
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (uint64_t k0 = 0; k0 < jobs_n; k0 += (uint64_t)threads) {
        int started = 0;
        for (int t = 0; t < threads && k0 + t < jobs_n; ++t, ++started)
            pthread_create(&th[t], NULL, sin_worker, &jobs[k0 + t]);
        for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    }
    free(th);
    free(jobs);
    free(f);
    free(ph);
}

/* Selected DFT bins in FP64 by direct summation with exact index reduction
 * (k*j mod N in integers): out[i] = sum_j x_j exp(-2 pi i bins[i] j / N). */
void oracle_dft_bins64(const float *x, uint64_t n, const uint64_t *bins, uint64_t count, double *out) {
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t k = bins[i] % n, idx = 0;
        long double re = 0.0L, im = 0.0L;
        for (uint64_t j = 0; j < n; ++j) {
            double a = 2.0 * M_PI * ((double)idx / (double)n);
            double c = cos(a), s = sin(a);
            double xr = x[2 * j], xi = x[2 * j + 1];
            re += xr * c + xi * s;
            im += xi * c - xr * s;
            idx += k;
            if (idx >= n) idx -= n;
        }
        out[2 * i] = (double)re;
        out[2 * i + 1] = (double)im;
    }
}

/* Ragged batch: log2 lengths k ~ U{kmin..kmax} via (u32 >> 8) % span. */
void oracle_synth_ragged_log2(uint32_t *log2n, uint64_t count, uint32_t kmin, uint32_t kmax,
                              uint64_t seed) {
    pcg32_t r;
    seed_stream(&r, seed);
    uint32_t span = kmax - kmin + 1u;
    for (uint64_t i = 0; i < count; ++i) log2n[i] = kmin + (pcg32_next(&r) >> 8) % span;
}

/* Window w_n = 0.5(1-cos(2 pi n/(N-1))) + 0.01 g_n (FP64, rounded to f32) and
 * spectral mask m_k ~ Bernoulli(0.3) drawn as U[0,1) < 0.3, from one stream:
 * N gaussians first, then N mask draws. */
void oracle_synth_window_mask(float *w, float *mask, uint64_t n, uint64_t seed) {
    pcg32_t r;
    seed_stream(&r, seed);
    for (uint64_t i = 0; i < n; ++i) {
        double hann = n > 1 ? 0.5 * (1.0 - cos(2.0 * M_PI * (double)i / (double)(n - 1))) : 1.0;
        w[i] = (float)(hann + 0.01 * gauss(&r));
    }
    for (uint64_t i = 0; i < n; ++i) mask[i] = u01_24(pcg32_next(&r)) < 0.3 ? 1.0f : 0.0f;
}

/* ------------------------------------------------------------------ */
/* 2. Restatement of the reference f32 Stockham R4/R2 path             */
/* ------------------------------------------------------------------ */

/* src/lib.rs:764-768: angle = -TAU * (j as f32) / (n as f32), f32 cos/sin. */
static void ref_twiddles(float *tw, uint64_t n) {
    const float tau = 6.28318530717958647692f; /* std::f32::consts::TAU */
    for (uint64_t j = 0; j < n; ++j) {
        float angle = -tau * (float)j / (float)n;
        tw[2 * j] = cosf(angle);
        tw[2 * j + 1] = sinf(angle);
    }
}

/* src/shaders.rs:14-16 cmul(a, b) with a = twiddle, b = sample. */
static inline void cmul_ref(float ar, float ai, float br, float bi, float *cr, float *ci) {
    float t0 = ar * br;
    float t1 = ai * bi;
    float t2 = ar * bi;
    float t3 = ai * br;
    *cr = t0 - t1;
    *ci = t2 + t3;
}

/* One R4 stage, src/shaders.rs:18-62, for all tid < n/4 of one signal. */
static void ref_r4_stage(const float *src, float *dst, const float *tw, uint32_t n, uint32_t p) {
    uint32_t quarter_n = n >> 2;
    uint32_t four_p = p << 2;
    uint32_t stride = quarter_n / p;
    for (uint32_t tid = 0; tid < quarter_n; ++tid) {
        uint32_t k = tid % p, j = tid / p;
        uint32_t i0 = j * p + k, i1 = i0 + quarter_n, i2 = i1 + quarter_n, i3 = i2 + quarter_n;
        float x0r = src[2 * i0], x0i = src[2 * i0 + 1];
        float x1r = src[2 * i1], x1i = src[2 * i1 + 1];
        float x2r = src[2 * i2], x2i = src[2 * i2 + 1];
        float x3r = src[2 * i3], x3i = src[2 * i3 + 1];
        uint32_t t = k * stride;
        cmul_ref(tw[2 * t], tw[2 * t + 1], x1r, x1i, &x1r, &x1i);
        cmul_ref(tw[4 * t], tw[4 * t + 1], x2r, x2i, &x2r, &x2i);
        cmul_ref(tw[6 * t], tw[6 * t + 1], x3r, x3i, &x3r, &x3i);
        float s02r = x0r + x2r, s02i = x0i + x2i, d02r = x0r - x2r, d02i = x0i - x2i;
        float s13r = x1r + x3r, s13i = x1i + x3i, d13r = x1r - x3r, d13i = x1i - x3i;
        uint32_t d = 2u * (j * four_p + k);
        dst[d] = s02r + s13r;
        dst[d + 1] = s02i + s13i;
        dst[d + 2 * p] = d02r + d13i;
        dst[d + 2 * p + 1] = d02i - d13r;
        dst[d + 4 * p] = s02r - s13r;
        dst[d + 4 * p + 1] = s02i - s13i;
        dst[d + 6 * p] = d02r - d13i;
        dst[d + 6 * p + 1] = d02i + d13r;
    }
}

/* One R2 stage, src/shaders.rs:79-105. */
static void ref_r2_stage(const float *src, float *dst, const float *tw, uint32_t n, uint32_t p) {
    uint32_t half_n = n >> 1, two_p = p + p;
    for (uint32_t tid = 0; tid < half_n; ++tid) {
        uint32_t k = tid % p, j = tid / p;
        uint32_t i1 = j * p + k, i2 = i1 + half_n;
        float x1r = src[2 * i1], x1i = src[2 * i1 + 1];
        float x2r = src[2 * i2], x2i = src[2 * i2 + 1];
        uint32_t t = k * (half_n / p);
        float tr, ti;
        cmul_ref(tw[2 * t], tw[2 * t + 1], x2r, x2i, &tr, &ti);
        uint32_t d = 2u * (j * two_p + k);
        dst[d] = x1r + tr;
        dst[d + 1] = x1i + ti;
        dst[d + 2 * p] = x1r - tr;
        dst[d + 2 * p + 1] = x1i - ti;
    }
}

/* Forward transform of one signal held in buffer a; returns the buffer holding
 * the result (a or b), following the ping-pong of src/lib.rs:863-889. */
static float *ref_signal(float *a, float *b, const float *tw, uint32_t n) {
    uint32_t log_n = 0;
    while ((1u << log_n) < n) ++log_n;
    uint32_t num_r4 = log_n / 2;
    int has_r2 = log_n % 2 == 1;
    float *src = a, *dst = b;
    for (uint32_t s = 0; s < num_r4; ++s) {
        ref_r4_stage(src, dst, tw, n, 1u << (2 * s));
        float *t = src; src = dst; dst = t;
    }
    if (has_r2) {
        ref_r2_stage(src, dst, tw, n, 1u << (2 * num_r4));
        float *t = src; src = dst; dst = t;
    }
    return src;
}

typedef struct {
    const float *in;
    float *out;
    const float *tw;
    uint32_t n;
    int inverse;
    uint64_t b0, b1;
} ref_job_t;

static void *ref_worker(void *arg) {
    ref_job_t *j = (ref_job_t *)arg;
    uint32_t n = j->n;
    float *a = (float *)malloc(sizeof(float) * 2 * n);
    float *b = (float *)malloc(sizeof(float) * 2 * n);
    const float scale = 1.0f / (float)n; /* src/lib.rs:662 */
    for (uint64_t s = j->b0; s < j->b1; ++s) {
        const float *x = j->in + 2 * (uint64_t)n * s;
        /* prepare_input_data, src/lib.rs:571-579 */
        for (uint32_t i = 0; i < n; ++i) {
            a[2 * i] = x[2 * i];
            a[2 * i + 1] = j->inverse ? -x[2 * i + 1] : x[2 * i + 1];
        }
        float *res = ref_signal(a, b, j->tw, n);
        float *y = j->out + 2 * (uint64_t)n * s;
        if (j->inverse) {
            /* apply_inverse_transform_postprocessing, src/lib.rs:661-669 */
            for (uint32_t i = 0; i < n; ++i) {
                y[2 * i] = res[2 * i] * scale;
                y[2 * i + 1] = -res[2 * i + 1] * scale;
            }
        } else {
            memcpy(y, res, sizeof(float) * 2 * n);
        }
    }
    free(a);
    free(b);
    return NULL;
}

/* Reference f32 path over a batch, split across `threads` pthreads. */
int oracle_ref_fft_f32(const float *in, float *out, uint64_t n, uint64_t batch, int inverse,
                       int threads) {
    if (n == 0 || (n & (n - 1)) != 0 || n > (1ull << 31)) return -1;
    if (batch == 0) return 0;
    float *tw = (float *)malloc(sizeof(float) * 2 * n);
    if (!tw) return -2;
    ref_twiddles(tw, n);
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > batch) threads = (int)batch;
    ref_job_t *jobs = (ref_job_t *)calloc((size_t)threads, sizeof(ref_job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        jobs[t].in = in; jobs[t].out = out; jobs[t].tw = tw; jobs[t].n = (uint32_t)n;
        jobs[t].inverse = inverse;
        jobs[t].b0 = batch * (uint64_t)t / (uint64_t)threads;
        jobs[t].b1 = batch * (uint64_t)(t + 1) / (uint64_t)threads;
    }
    if (threads == 1) {
        ref_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, ref_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    }
    free(th);
    free(jobs);
    free(tw);
    return 0;
}

/* The reference's own f32 twiddle table (for direct inspection in tests). */
void oracle_ref_twiddles(float *tw, uint64_t n) { ref_twiddles(tw, n); }

/* ------------------------------------------------------------------ */
/* 3. FP64 oracle                                                      */
/* ------------------------------------------------------------------ */

/* Direct DFT, sign = -1 forward, +1 inverse (unnormalised).  Index product
 * j*k is reduced mod n exactly before forming the angle. */
void oracle_dft64(const double *in, double *out, uint64_t n, uint64_t batch, int sign) {
    double *c = (double *)malloc(sizeof(double) * n);
    double *s = (double *)malloc(sizeof(double) * n);
    for (uint64_t j = 0; j < n; ++j) {
        double th = 2.0 * M_PI * (double)j / (double)n;
        c[j] = cos(th);
        s[j] = sin(th) * (double)sign;
    }
    for (uint64_t b = 0; b < batch; ++b) {
        const double *x = in + 2 * n * b;
        double *y = out + 2 * n * b;
        for (uint64_t k = 0; k < n; ++k) {
            double re = 0.0, im = 0.0;
            uint64_t idx = 0;
            for (uint64_t j = 0; j < n; ++j) {
                double xr = x[2 * j], xi = x[2 * j + 1];
                re += xr * c[idx] - xi * s[idx];
                im += xr * s[idx] + xi * c[idx];
                idx += k;
                if (idx >= n) idx -= n;
            }
            y[2 * k] = re;
            y[2 * k + 1] = im;
        }
    }
    free(c);
    free(s);
}

/* Iterative radix-2 DIT FFT in FP64 (bit reversal + butterflies), twiddles
 * from a FP64 table cos/sin(2 pi j / n). sign = -1 forward, +1 inverse. */
static void fft64_one(double *x, uint64_t n, const double *c, const double *s) {
    for (uint64_t i = 1, j = 0; i < n; ++i) {
        uint64_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr; x[2 * j + 1] = ti;
        }
    }
    for (uint64_t len = 2; len <= n; len <<= 1) {
        uint64_t half = len >> 1, step = n / len;
        for (uint64_t i = 0; i < n; i += len) {
            for (uint64_t k = 0; k < half; ++k) {
                double wr = c[k * step], wi = s[k * step];
                double *u = x + 2 * (i + k), *v = x + 2 * (i + k + half);
                double tr = v[0] * wr - v[1] * wi;
                double ti = v[0] * wi + v[1] * wr;
                v[0] = u[0] - tr; v[1] = u[1] - ti;
                u[0] += tr; u[1] += ti;
            }
        }
    }
}

typedef struct {
    double *x;
    uint64_t n, b0, b1;
    const double *c, *s;
} fft64_job_t;

static void *fft64_worker(void *arg) {
    fft64_job_t *j = (fft64_job_t *)arg;
    for (uint64_t b = j->b0; b < j->b1; ++b) fft64_one(j->x + 2 * j->n * b, j->n, j->c, j->s);
    return NULL;
}

/* One transform split over threads (for batches smaller than the thread
 * count): every stage of fft64_one's butterflies, each computed exactly as
 * there, is divided between the threads (by blocks, or by k within each block
 * when there are fewer blocks than threads), with a barrier between stages,
 * so the result is bit-identical to the one-thread transform.  Thread t also
 * fills its share of the twiddle table and of the bit-reversal swaps (a pair
 * i < rev(i) is swapped only by the owner of i). */
typedef struct {
    double *x;
    uint64_t n;
    double *c, *s;
    int sign, t, nt;
    pthread_barrier_t *bar;
} fft64_par_t;

static void *fft64_par_worker(void *arg) {
    fft64_par_t *p = (fft64_par_t *)arg;
    const uint64_t n = p->n, T = (uint64_t)p->nt, t = (uint64_t)p->t;
    double *x = p->x, *c = p->c, *s = p->s;
    int log2n = 0;
    while ((1ull << log2n) < n) ++log2n;
    for (uint64_t j = (n / 2 + 1) * t / T; j < (n / 2 + 1) * (t + 1) / T; ++j) {
        double th = 2.0 * M_PI * (double)j / (double)n;
        c[j] = cos(th);
        s[j] = sin(th) * (double)p->sign;
    }
    for (uint64_t i = n * t / T; i < n * (t + 1) / T; ++i) {
        uint64_t r = 0;
        for (int b = 0; b < log2n; ++b) r |= ((i >> b) & 1ull) << (log2n - 1 - b);
        if (i < r) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * r]; x[2 * i + 1] = x[2 * r + 1];
            x[2 * r] = tr; x[2 * r + 1] = ti;
        }
    }
    pthread_barrier_wait(p->bar);
    for (uint64_t len = 2; len <= n; len <<= 1) {
        const uint64_t half = len >> 1, step = n / len, blocks = n / len;
        uint64_t b0 = 0, b1 = blocks, k0 = 0, k1 = half;
        if (blocks >= T) {
            b0 = blocks * t / T;
            b1 = blocks * (t + 1) / T;
        } else {
            k0 = half * t / T;
            k1 = half * (t + 1) / T;
        }
        for (uint64_t blk = b0; blk < b1; ++blk) {
            const uint64_t i = blk * len;
            for (uint64_t k = k0; k < k1; ++k) {
                double wr = c[k * step], wi = s[k * step];
                double *u = x + 2 * (i + k), *v = x + 2 * (i + k + half);
                double tr = v[0] * wr - v[1] * wi;
                double ti = v[0] * wi + v[1] * wr;
                v[0] = u[0] - tr; v[1] = u[1] - ti;
                u[0] += tr; u[1] += ti;
            }
        }
        pthread_barrier_wait(p->bar);
    }
    return NULL;
}

static int fft64_one_parallel(double *x, uint64_t n, int sign, int threads) {
    double *c = (double *)malloc(sizeof(double) * (n / 2 + 1));
    double *s = (double *)malloc(sizeof(double) * (n / 2 + 1));
    fft64_par_t *jobs = (fft64_par_t *)calloc((size_t)threads, sizeof(fft64_par_t));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    pthread_barrier_t bar;
    if (!c || !s || !jobs || !th || pthread_barrier_init(&bar, NULL, (unsigned)threads) != 0) {
        free(c); free(s); free(jobs); free(th);
        return -2;
    }
    for (int t = 0; t < threads; ++t) {
        jobs[t].x = x; jobs[t].n = n; jobs[t].c = c; jobs[t].s = s; jobs[t].sign = sign;
        jobs[t].t = t; jobs[t].nt = threads; jobs[t].bar = &bar;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, fft64_par_worker, &jobs[t]);
    fft64_par_worker(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&bar);
    free(th);
    free(jobs);
    free(c);
    free(s);
    return 0;
}

/* In-place batched FP64 FFT over `threads` pthreads: whole transforms per
 * thread when the batch has at least as many transforms as threads (or the
 * transforms are short), else each transform split over the threads. */
int oracle_fft64_inplace(double *x, uint64_t n, uint64_t batch, int sign, int threads) {
    if (n == 0 || (n & (n - 1)) != 0) return -1;
    if (threads > 1 && batch < (uint64_t)threads && n >= (1ull << 16)) {
        for (uint64_t b = 0; b < batch; ++b) {
            int rc = fft64_one_parallel(x + 2 * n * b, n, sign, threads);
            if (rc) return rc;
        }
        return 0;
    }
    double *c = (double *)malloc(sizeof(double) * (n / 2 + 1));
    double *s = (double *)malloc(sizeof(double) * (n / 2 + 1));
    for (uint64_t j = 0; j <= n / 2; ++j) {
        /* symmetric evaluation keeps the table accurate near pi/2 multiples */
        double th = 2.0 * M_PI * (double)j / (double)n;
        c[j] = cos(th);
        s[j] = sin(th) * (double)sign;
    }
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > batch) threads = (int)(batch ? batch : 1);
    fft64_job_t *jobs = (fft64_job_t *)calloc((size_t)threads, sizeof(fft64_job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        jobs[t].x = x; jobs[t].n = n; jobs[t].c = c; jobs[t].s = s;
        jobs[t].b0 = batch * (uint64_t)t / (uint64_t)threads;
        jobs[t].b1 = batch * (uint64_t)(t + 1) / (uint64_t)threads;
    }
    if (threads == 1) {
        fft64_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, fft64_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    }
    free(th);
    free(jobs);
    free(c);
    free(s);
    return 0;
}

/* Y_b = DFT(w * x_b) in FP64 (w real, shared by the batch). */
int oracle_windowed_fwd64(const float *x, const float *w, double *y, uint64_t n, uint64_t batch,
                          int threads) {
    for (uint64_t b = 0; b < batch; ++b)
        for (uint64_t i = 0; i < n; ++i) {
            y[2 * (n * b + i)] = (double)w[i] * (double)x[2 * (n * b + i)];
            y[2 * (n * b + i) + 1] = (double)w[i] * (double)x[2 * (n * b + i) + 1];
        }
    return oracle_fft64_inplace(y, n, batch, -1, threads);
}

/*
 * Vector-Jacobian product of Y = DFT(w * x) for a real loss, complex gradients
 * in the (dL/dRe + i dL/dIm) convention:
 *   gu = F^H gY  (unnormalised inverse-sign DFT),
 *   gx = w * gu,
 *   gw_n = sum_b Re(conj(gu_bn) x_bn).
 */
int oracle_vjp64(const float *x, const float *w, const float *gy, double *gx, double *gw,
                 uint64_t n, uint64_t batch, int threads) {
    double *gu = (double *)malloc(sizeof(double) * 2 * n * batch);
    if (!gu) return -2;
    for (uint64_t i = 0; i < 2 * n * batch; ++i) gu[i] = (double)gy[i];
    oracle_fft64_inplace(gu, n, batch, +1, threads);
    for (uint64_t i = 0; i < n; ++i) gw[i] = 0.0;
    for (uint64_t b = 0; b < batch; ++b)
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t e = n * b + i;
            double ur = gu[2 * e], ui = gu[2 * e + 1];
            gx[2 * e] = (double)w[i] * ur;
            gx[2 * e + 1] = (double)w[i] * ui;
            gw[i] += ur * (double)x[2 * e] + ui * (double)x[2 * e + 1];
        }
    free(gu);
    return 0;
}

/* L = sum_b sum_k m_k |Y_bk|^2 with Y = DFT(w*x); returns L and writes
 * gY = 2 m Y (FP64) so tests can chain it into oracle_vjp64. */
double oracle_masked_loss64(const float *x, const float *w, const float *mask, double *gy,
                            uint64_t n, uint64_t batch, int threads) {
    double *y = (double *)malloc(sizeof(double) * 2 * n * batch);
    oracle_windowed_fwd64(x, w, y, n, batch, threads);
    double loss = 0.0;
    for (uint64_t b = 0; b < batch; ++b)
        for (uint64_t k = 0; k < n; ++k) {
            uint64_t e = n * b + k;
            double m = (double)mask[k];
            loss += m * (y[2 * e] * y[2 * e] + y[2 * e + 1] * y[2 * e + 1]);
            if (gy) {
                gy[2 * e] = 2.0 * m * y[2 * e];
                gy[2 * e + 1] = 2.0 * m * y[2 * e + 1];
            }
        }
    free(y);
    return loss;
}
