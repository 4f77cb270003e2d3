/*
 * sanitize_check.c -- drives every function of oracle.c on small inputs
 * under AddressSanitizer and UndefinedBehaviorSanitizer (test
 * infrastructure, built by `make -C oracle sanitize`).  Each call is also
 * checked against a closed form or an invariant, so the driver fails on a
 * wrong answer as well as on a memory or undefined-behaviour error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint64_t oracle_splitmix64_next(uint64_t *s);
void oracle_pcg32_raw(uint64_t initstate, uint64_t initseq, uint32_t *out, uint64_t count);
void oracle_synth_uniform(float *out, uint64_t nfloats, uint64_t seed);
void oracle_synth_gauss(double *out, uint64_t count, uint64_t seed);
void oracle_synth_sinusoids(float *out, uint64_t n, uint64_t batch, uint64_t seed, int threads);
void oracle_dft_bins64(const float *x, uint64_t n, const uint64_t *bins, uint64_t count, double *out);
void oracle_synth_ragged_log2(uint32_t *log2n, uint64_t count, uint32_t kmin, uint32_t kmax, uint64_t seed);
void oracle_synth_window_mask(float *w, float *mask, uint64_t n, uint64_t seed);
int oracle_ref_fft_f32(const float *in, float *out, uint64_t n, uint64_t batch, int inverse, int threads);
void oracle_ref_twiddles(float *tw, uint64_t n);
void oracle_dft64(const double *in, double *out, uint64_t n, uint64_t batch, int sign);
int oracle_fft64_inplace(double *x, uint64_t n, uint64_t batch, int sign, int threads);
int oracle_windowed_fwd64(const float *x, const float *w, double *y, uint64_t n, uint64_t batch, int threads);
int oracle_vjp64(const float *x, const float *w, const float *gy, double *gx, double *gw, uint64_t n,
                 uint64_t batch, int threads);
double oracle_masked_loss64(const float *x, const float *w, const float *mask, double *gy, uint64_t n,
                            uint64_t batch, int threads);

static int failures = 0;
#define CHECK(c, what)                                            \
    do {                                                          \
        int ok_ = (c);                                            \
        printf("%s %s\n", ok_ ? "PASS" : "FAIL", what);           \
        if (!ok_) ++failures;                                     \
    } while (0)

static double max_abs_diff(const double *a, const double *b, size_t len) {
    double m = 0.0;
    for (size_t i = 0; i < len; ++i) m = fmax(m, fabs(a[i] - b[i]));
    return m;
}

int main(void) {
    uint64_t s = 0;
    CHECK(oracle_splitmix64_next(&s) == 0xE220A8397B1DCDAFull, "splitmix64 first output");
    uint32_t pcg[2];
    oracle_pcg32_raw(42, 54, pcg, 2);
    CHECK(pcg[0] == 0xA15C02B7u && pcg[1] == 0x7B47F409u, "pcg32 first outputs");

    enum { N = 256, B = 3 };
    float *x = malloc(sizeof(float) * 2 * N * B), *y = malloc(sizeof(float) * 2 * N * B);
    oracle_synth_uniform(x, 2 * N * B, 7);
    int in_range = 1;
    for (int i = 0; i < 2 * N * B; ++i) in_range &= (x[i] >= -1.0f && x[i] < 1.0f);
    CHECK(in_range, "uniform in [-1, 1)");
    double g[64];
    oracle_synth_gauss(g, 64, 3);
    CHECK(isfinite(g[63]), "gauss finite");
    float *sin = malloc(sizeof(float) * 2 * 70000 * 2);
    oracle_synth_sinusoids(sin, 70000, 2, 11, 3); /* two blocks per signal, partial last block */
    CHECK(isfinite(sin[2 * 70000 * 2 - 1]), "sinusoids over a partial block");
    uint32_t lg[100];
    oracle_synth_ragged_log2(lg, 100, 6, 16, 5);
    int lg_ok = 1;
    for (int i = 0; i < 100; ++i) lg_ok &= (lg[i] >= 6 && lg[i] <= 16);
    CHECK(lg_ok, "ragged lengths in range");
    float w[N], m[N];
    oracle_synth_window_mask(w, m, N, 9);
    CHECK(fabsf(w[N / 2] - 1.0f) < 0.1f && (m[0] == 0.0f || m[0] == 1.0f), "window and mask");

    /* reference restatement: impulse -> ones, threads do not change bits, round trip */
    float *imp = calloc(2 * N, sizeof(float));
    imp[0] = 1.0f;
    CHECK(oracle_ref_fft_f32(imp, y, N, 1, 0, 1) == 0 && fabsf(y[2 * 77] - 1.0f) < 1e-6f && fabsf(y[2 * 77 + 1]) < 1e-6f,
          "restatement impulse");
    float *y2 = malloc(sizeof(float) * 2 * N * B), *r = malloc(sizeof(float) * 2 * N * B);
    oracle_ref_fft_f32(x, y, N, B, 0, 1);
    oracle_ref_fft_f32(x, y2, N, B, 0, 3);
    CHECK(memcmp(y, y2, sizeof(float) * 2 * N * B) == 0, "restatement threads bit-identical");
    oracle_ref_fft_f32(y, r, N, B, 1, 2);
    double rt = 0.0;
    for (int i = 0; i < 2 * N * B; ++i) rt = fmax(rt, fabs((double)r[i] - (double)x[i]));
    CHECK(rt < 1e-5, "restatement round trip");
    CHECK(oracle_ref_fft_f32(x, y, 12, 1, 0, 1) != 0, "restatement rejects a non-power of two");
    float tw[2 * 64];
    oracle_ref_twiddles(tw, 64);
    CHECK(fabsf(tw[0] - 1.0f) < 1e-7f && fabsf(tw[2 * 16 + 1] + 1.0f) < 1e-6f, "f32-angle twiddles");

    /* FP64: radix-2 FFT (batched and split over threads) vs direct DFT and bins */
    double *xd = malloc(sizeof(double) * 2 * N * B), *yd = malloc(sizeof(double) * 2 * N * B);
    double *zd = malloc(sizeof(double) * 2 * N * B);
    for (int i = 0; i < 2 * N * B; ++i) xd[i] = x[i];
    oracle_dft64(xd, yd, N, B, -1);
    memcpy(zd, xd, sizeof(double) * 2 * N * B);
    oracle_fft64_inplace(zd, N, B, -1, 2);
    CHECK(max_abs_diff(yd, zd, 2 * N * B) < 1e-10, "fft64 vs dft64");
    enum { NB = 1 << 16 };
    double *big = malloc(sizeof(double) * 2 * NB), *big2 = malloc(sizeof(double) * 2 * NB);
    float *bigf = malloc(sizeof(float) * 2 * NB);
    oracle_synth_uniform(bigf, 2 * NB, 13);
    for (int i = 0; i < 2 * NB; ++i) big[i] = big2[i] = bigf[i];
    oracle_fft64_inplace(big, NB, 1, -1, 1);
    oracle_fft64_inplace(big2, NB, 1, -1, 5); /* one transform split over five threads */
    CHECK(memcmp(big, big2, sizeof(double) * 2 * NB) == 0, "fft64 split over threads bit-identical");
    uint64_t bins[3] = {0, 1, NB - 1};
    double bv[6];
    oracle_dft_bins64(bigf, NB, bins, 3, bv);
    CHECK(fabs(bv[2] - big[2]) < 1e-8 && fabs(bv[5] - big[2 * (NB - 1) + 1]) < 1e-8, "dft bins vs fft64");
    CHECK(oracle_fft64_inplace(zd, 12, 1, -1, 1) != 0, "fft64 rejects a non-power of two");

    /* windowed forward, VJP, masked loss: the adjoint identity sum gy conj... via the loss gradient */
    double *wy = malloc(sizeof(double) * 2 * N * B), *gy = malloc(sizeof(double) * 2 * N * B);
    double *gx = malloc(sizeof(double) * 2 * N * B), gw[N];
    oracle_windowed_fwd64(x, w, wy, N, B, 2);
    double loss = oracle_masked_loss64(x, w, m, gy, N, B, 2);
    double want = 0.0;
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < N; ++k) {
            const double re = wy[2 * (N * b + k)], im = wy[2 * (N * b + k) + 1];
            want += m[k] * (re * re + im * im);
        }
    CHECK(fabs(loss - want) <= 1e-9 * fmax(1.0, want), "masked loss");
    float *gyf = malloc(sizeof(float) * 2 * N * B);
    for (int i = 0; i < 2 * N * B; ++i) gyf[i] = (float)gy[i];
    CHECK(oracle_vjp64(x, w, gyf, gx, gw, N, B, 2) == 0 && isfinite(gw[N - 1]), "vjp");
    /* <gx, x> = sum_n w_n Re(conj(gu) x) summed over the batch = sum gw_n w_n ... */
    double lhs = 0.0, rhs = 0.0;
    for (int b = 0; b < B; ++b)
        for (int i = 0; i < N; ++i) lhs += gx[2 * (N * b + i)] * x[2 * (N * b + i)] + gx[2 * (N * b + i) + 1] * x[2 * (N * b + i) + 1];
    for (int i = 0; i < N; ++i) rhs += gw[i] * w[i];
    CHECK(fabs(lhs - rhs) <= 1e-9 * fmax(1.0, fabs(rhs)), "vjp: <gx, x> = <gw, w>");
    CHECK(fabs(lhs - 2.0 * loss) <= 1e-6 * fmax(1.0, loss), "vjp: Euler identity <grad, x> = 2 L");

    free(x); free(y); free(y2); free(r); free(sin); free(imp); free(xd); free(yd); free(zd);
    free(big); free(big2); free(bigf); free(wy); free(gy); free(gx); free(gyf);
    printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
