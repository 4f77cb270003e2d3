// gpu_fft_tests.cpp -- the reference's own tests restated against the C++
// mirror (tests/gpu_fft_test.rs, tests/ifft_test.rs, tests/windowed_fft_test.rs,
// src/lib.rs:207-227), checked against an in-file FP64 DFT.
//   gpu_fft_tests cpu   error contract without a device
//   gpu_fft_tests gpu   transforms on the GPU
// Prints one line per check and exits non-zero on the first failure.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_fft.hpp"

using wgls_rs_fft::Batch;
using wgls_rs_fft::check;
using wgls_rs_fft::Complex;
using wgls_rs_fft::FftError;
using wgls_rs_fft::GpuFft;

static int failures = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        bool ok_ = (cond);                                                 \
        std::printf("%s %s\n", ok_ ? "PASS" : "FAIL", what);               \
        if (!ok_) ++failures;                                              \
    } while (0)

// Direct FP64 DFT (sign -1 forward, +1 inverse without scale), exact index reduction.
static std::vector<std::complex<double>> dft(const std::vector<Complex>& x, int sign) {
    const size_t n = x.size();
    std::vector<std::complex<double>> y(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc = 0.0;
        for (size_t j = 0; j < n; ++j) {
            const double a = 2.0 * M_PI * (double)((k * j) % n) / (double)n;
            acc += std::complex<double>(x[j]) * std::complex<double>(std::cos(a), sign * std::sin(a));
        }
        y[k] = acc;
    }
    return y;
}

static std::vector<Complex> make_input(size_t n) {  // tests/gpu_fft_test.rs:8-15
    std::vector<Complex> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = Complex(std::sin((float)i * 0.1f), std::cos((float)i * 0.07f));
    return v;
}

static std::vector<Complex> generate_test_signal(size_t n) {  // tests/test_utils.rs:21-34
    std::vector<Complex> v(n);
    const float pi = 3.14159265358979323846f;
    for (size_t i = 0; i < n; ++i) {
        const float t = (float)i / (float)n;
        v[i] = Complex(0.7f * std::sin(2.0f * pi * 10.0f * t) + 0.3f * std::sin(2.0f * pi * 50.0f * t), 0.0f);
    }
    return v;
}

static void apply_hann_window(std::vector<Complex>& s) {  // tests/test_utils.rs:39-50
    const size_t n = s.size();
    const float pi = 3.14159265358979323846f;
    for (size_t i = 0; i < n; ++i) s[i] *= 0.5f * (1.0f - std::cos(2.0f * pi * (float)i / (float)(n - 1)));
}

static float max_diff(const std::vector<Complex>& a, const std::vector<Complex>& b) {
    float m = 0.f;
    for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::abs(a[i] - b[i]));
    return m;
}

// Host logic that needs no device: validation, argument checks on every
// entry point, input synthesis, registry queries (the sanitizer build runs
// these on the CPU).
static void run_host_logic() {
    uint64_t n = 7;
    CHECK(wgfft_validate_lengths(nullptr, 0, &n) == WGFFT_OK && n == 0, "empty batch validates to n = 0");
    uint64_t ok[3] = {1024, 1024, 1024};
    CHECK(wgfft_validate_lengths(ok, 3, &n) == WGFFT_OK && n == 1024, "equal power-of-two lengths");
    CHECK(wgfft_validate_lengths(nullptr, 2, &n) == WGFFT_ERR_INVALID_ARGUMENT, "NULL length table");
    float buf[16];
    CHECK(wgfft_exec_host(nullptr, buf, buf, 8, 1, WGFFT_FORWARD) == WGFFT_ERR_INVALID_ARGUMENT, "exec_host NULL handle");
    CHECK(wgfft_exec_device(nullptr, buf, buf, 8, 1, WGFFT_FORWARD, nullptr) == WGFFT_ERR_INVALID_ARGUMENT,
          "exec_device NULL handle");
    CHECK(wgfft_exec_device_noalloc(nullptr, buf, buf, 8, 1, WGFFT_FORWARD, nullptr) == WGFFT_ERR_INVALID_ARGUMENT,
          "exec_device_noalloc NULL handle");
    CHECK(wgfft_exec_strided_device(nullptr, buf, buf, 8, 1, 1, 8, 1, 8, WGFFT_FORWARD, nullptr) ==
              WGFFT_ERR_INVALID_ARGUMENT,
          "exec_strided NULL handle");
    CHECK(wgfft_vjp_windowed_device(nullptr, buf, buf, buf, buf, buf, 8, 1, nullptr) == WGFFT_ERR_INVALID_ARGUMENT,
          "vjp NULL handle");
    CHECK(wgfft_fwd_adjoint_masked_device(nullptr, buf, buf, buf, nullptr, buf, buf, buf, 8, 1, nullptr) ==
              WGFFT_ERR_INVALID_ARGUMENT,
          "fused adjoint NULL handle");
    CHECK(wgfft_ragged_create(nullptr, nullptr, nullptr, 0, nullptr) == WGFFT_ERR_INVALID_ARGUMENT,
          "ragged_create NULL handle");
    CHECK(wgfft_ragged_samples(nullptr) == 0, "ragged_samples NULL plan");
    wgfft_ragged_destroy(nullptr);
    wgfft_destroy(nullptr);
    int split[3] = {7, 7, 7};
    CHECK(wgfft_select_split(nullptr, 1u << 21, split, 3, nullptr, 0) == WGFFT_ERR_INVALID_ARGUMENT,
          "select_split NULL handle");
    double sec = 0;
    CHECK(wgfft_bench_device(nullptr, buf, buf, 8, 1, WGFFT_FORWARD, 0, 1, &sec) == WGFFT_ERR_INVALID_ARGUMENT,
          "bench NULL handle");
    std::vector<float> a(1000), b(1000);
    wgfft_synth_uniform(a.data(), a.size(), 42);
    wgfft_synth_uniform(b.data(), b.size(), 42);
    bool in_range = true;
    for (float v : a) in_range &= (v >= -1.f && v < 1.f);
    CHECK(a == b && in_range, "synth_uniform deterministic in [-1, 1)");
    CHECK(wgfft_variant_count(4096) >= 1 && wgfft_variant_count(12) == 0 && wgfft_pass_count(1u << 22) == 2 &&
              wgfft_pass_variant_count(1u << 22, 5) == 0 && wgfft_adjoint_variant_count(1u << 20) == 0,
          "registry queries");
    CHECK(std::string(wgfft_name()).find("Stockham") != std::string::npos, "name");
}

static int run_cpu() {
    run_host_logic();
    CHECK(!GpuFft::is_gpu_available(), "no device in this process");
    try {
        GpuFft::create();
        CHECK(false, "create() without a device throws");
    } catch (const FftError& e) {
        CHECK(e.code() == WGFFT_ERR_NO_DEVICE, "create() without a device throws WGFFT_ERR_NO_DEVICE");
    }
    uint64_t n = 0;
    uint64_t ragged[2] = {8, 16}, bad[1] = {12};
    CHECK(wgfft_validate_lengths(ragged, 2, &n) == WGFFT_ERR_RAGGED_BATCH &&
              std::string(wgfft_last_error()) == "All input vectors in a batch must have the same length",
          "ragged batch message");
    CHECK(wgfft_validate_lengths(bad, 1, &n) == WGFFT_ERR_INVALID_LENGTH &&
              std::string(wgfft_last_error()) == "Transform length must be a non-zero power of two",
          "bad length message");
    return failures ? 1 : 0;
}

static int run_gpu() {
    GpuFft fft = GpuFft::create();
    CHECK(!fft.name().empty(), "name()");
    // tests/gpu_fft_test.rs:17-42
    {
        auto x = make_input(1024);
        Batch out = fft.fft({x});
        auto ref = dft(x, -1);
        float m = 0.f;
        for (size_t i = 0; i < 1024; ++i) m = std::max(m, (float)std::abs(std::complex<double>(out[0][i]) - ref[i]));
        CHECK(out.size() == 1 && out[0].size() == 1024 && m < 1e-3f, "gpu_fft_matches_rustfft (1e-3 vs FP64)");
    }
    // tests/ifft_test.rs:22-107
    {
        auto x = make_input(1024);
        Batch spec = fft.fft({x});
        Batch rec = fft.ifft(spec);
        CHECK(max_diff(rec[0], x) < 1e-3f, "ifft roundtrip");
        Batch spec2 = fft.fft(rec);
        CHECK(max_diff(spec2[0], spec[0]) < 1e-3f, "ifft->fft roundtrip");
    }
    // src/lib.rs:207-227
    {
        std::vector<Complex> x(1024);
        for (size_t i = 0; i < 1024; ++i) x[i] = Complex((float)i * 0.1f, 0.f);
        Batch rec = fft.ifft(fft.fft({x}));
        CHECK(max_diff(rec[0], x) < 1e-4f, "roundtrip_consistency (1e-4)");
    }
    // tests/windowed_fft_test.rs:10-133 (N = 2048)
    {
        auto sig = generate_test_signal(2048);
        auto win = sig;
        apply_hann_window(win);
        Batch spec = fft.fft({win});
        CHECK(max_diff(fft.ifft(spec)[0], win) < 1e-3f, "windowed roundtrip");
        std::vector<float> p(2048);
        for (size_t i = 0; i < 2048; ++i) p[i] = std::norm(spec[0][i]);
        bool p10 = false, p50 = false;
        for (size_t i = 1; i < 1024; ++i)
            if (p[i] > p[i - 1] && p[i] > p[i + 1] && p[i] > 10.f) {
                p10 |= (i >= 8 && i < 12);
                p50 |= (i >= 45 && i < 55);
            }
        CHECK(p10 && p50, "frequency detection at 10 and 50");
        std::vector<float> w(2048);
        const float pi = 3.14159265358979323846f;
        for (size_t i = 0; i < 2048; ++i) w[i] = 0.5f * (1.0f - std::cos(2.0f * pi * (float)i / 2047.0f));
        Batch fused = fft.fft_windowed({sig}, w);
        CHECK(max_diff(fused[0], spec[0]) < 1e-3f, "fft_windowed equals window-then-fft");
    }
    // host paths of the runtime under load: the pinned-staging pipeline
    // (>= 8 MB), a global-pass length, a pinned split, the autotuner, the
    // mixed-length plan (freed with its handle), padded rows and the
    // no-allocation entry point
    {
        const size_t n = 4096, batch = 1024;  // 32 MB: staged through pinned chunks
        Batch in(batch, make_input(n));
        Batch out = fft.fft(in);
        Batch back = fft.ifft(out);
        CHECK(out.size() == batch && max_diff(back[batch - 1], in[batch - 1]) < 1e-3f, "host pipeline 4096 x 1024");
        Batch big(2, make_input(1u << 20));
        Batch bo = fft.ifft(fft.fft(big));
        CHECK(max_diff(bo[1], big[1]) < 1e-3f, "multipass 2^20 round trip");
        char desc[256];
        int split[3] = {7, 7, 7};
        CHECK(wgfft_select_split(fft.handle(), 1u << 21, split, 3, desc, sizeof desc) == WGFFT_OK, "select_split");
        Batch s3(1, make_input(1u << 21));
        CHECK(max_diff(fft.ifft(fft.fft(s3))[0], s3[0]) < 1e-3f, "three-pass split round trip");
        CHECK(wgfft_autotune(fft.handle(), 1024, 512, desc, sizeof desc) == WGFFT_OK, "autotune 1024");
        CHECK(wgfft_autotune(fft.handle(), 1u << 17, 4, desc, sizeof desc) == WGFFT_OK, "autotune 2^17");
    }
    // the rows entry behind fft()/ifft() (wgfft_exec_host_rows): bit-identical
    // to the flat entry on the same signals, in place too, and for transforms
    // larger than a staging chunk
    {
        const size_t n = 256, batch = 700;  // 1.4 MB: one staged chunk, copies split over the copy threads
        Batch in(batch);
        for (size_t b = 0; b < batch; ++b) {
            in[b] = make_input(n);
            in[b][b % n] = Complex((float)b, -1.0f);
        }
        Batch out = fft.fft(in);
        std::vector<Complex> flat, fo(n * batch);
        for (const auto& v : in) flat.insert(flat.end(), v.begin(), v.end());
        check(wgfft_exec_host(fft.handle(), reinterpret_cast<const float*>(flat.data()),
                              reinterpret_cast<float*>(fo.data()), n, batch, WGFFT_FORWARD));
        bool same = out.size() == batch;
        for (size_t b = 0; b < batch && same; ++b) same = std::equal(out[b].begin(), out[b].end(), fo.begin() + b * n);
        CHECK(same, "rows entry equals the flat entry (256 x 700)");
        std::vector<float*> rows(batch);
        for (size_t b = 0; b < batch; ++b) rows[b] = reinterpret_cast<float*>(in[b].data());
        check(wgfft_exec_host_rows(fft.handle(), rows.data(), rows.data(), n, batch, WGFFT_FORWARD));
        same = true;
        for (size_t b = 0; b < batch && same; ++b) same = in[b] == out[b];
        CHECK(same, "rows entry in place");
        Batch huge(2, make_input(1u << 23));  // each transform larger than a staging chunk: row by row
        Batch hb = fft.ifft(fft.fft(huge));
        CHECK(max_diff(hb[1], huge[1]) < 1e-3f, "rows entry 2^23 round trip");
    }
    {
        wgfft_t* h = nullptr;
        check(wgfft_create(&h, 0));
        uint64_t offs[3] = {0, 64, 64 + 1024};
        uint32_t lg[3] = {6, 10, 16};
        wgfft_ragged_t* p = nullptr;
        wgfft_ragged_t* q = nullptr;
        CHECK(wgfft_ragged_create(h, offs, lg, 3, &p) == WGFFT_OK && wgfft_ragged_samples(p) == 64 + 1024 + 65536,
              "ragged plan");
        CHECK(wgfft_ragged_create(h, offs, lg, 3, &q) == WGFFT_OK, "second ragged plan");
        wgfft_ragged_destroy(q);
        wgfft_destroy(h);  // frees p with the handle
        CHECK(true, "handle destroyed with an open plan");
    }
    // error contract on a live executor
    CHECK(fft.fft({}).empty(), "empty batch -> empty result");
    try {
        fft.fft({std::vector<Complex>(8), std::vector<Complex>(16)});
        CHECK(false, "ragged batch throws");
    } catch (const FftError& e) {
        CHECK(e.code() == WGFFT_ERR_RAGGED_BATCH, "ragged batch throws");
    }
    try {
        fft.ifft({std::vector<Complex>(12)});
        CHECK(false, "non-power-of-two throws");
    } catch (const FftError& e) {
        CHECK(e.code() == WGFFT_ERR_INVALID_LENGTH, "non-power-of-two throws");
    }
    return failures ? 1 : 0;
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "gpu") {
        run_host_logic();
        return run_gpu();
    }
    return run_cpu();
}
