// gpu_fft.hpp -- C++ host mirror of wgls-rs-fft's GpuFft / FftExecutor over
// the C-ABI of include/wgfft.h (header-only; link libwgfft.so).
//
// Same names, argument meaning and error behaviour as the reference:
//   FftExecutor {name, fft, ifft, as_any}            src/lib.rs:32-45
//   GpuFft::new()  -> throws FftError when no GPU     src/lib.rs:285-330
//   GpuFft::is_gpu_available()                        src/lib.rs:381-389
//   fft / ifft: batch of equal-length power-of-two
//   signals in, new batch out; inverse 1/N scaled     src/lib.rs:433-488
//   empty batch -> empty result; ragged batch and
//   bad lengths -> the reference's messages          src/lib.rs:506-533
// Rust's Result<_, Box<dyn Error>> becomes a thrown FftError carrying the
// C-ABI status code and the same message text.
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wgfft.h"

namespace wgls_rs_fft {

using Complex = std::complex<float>;  // layout-compatible with interleaved (re, im) f32
using Batch = std::vector<std::vector<Complex>>;

class FftError : public std::runtime_error {
public:
    FftError(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const noexcept { return code_; }

private:
    int code_;
};

inline void check(int rc) {
    if (rc != WGFFT_OK) throw FftError(rc, wgfft_last_error());
}

class FftExecutor {
public:
    virtual ~FftExecutor() = default;
    virtual std::string name() const = 0;
    virtual Batch fft(const Batch& inputs) = 0;
    virtual Batch ifft(const Batch& inputs) = 0;
    virtual const void* as_any() const = 0;
};

class GpuFft final : public FftExecutor {
public:
    static GpuFft create(int device = 0) {
        wgfft_t* h = nullptr;
        check(wgfft_create(&h, device));
        return GpuFft(h);
    }
    static bool is_gpu_available() { return wgfft_device_available() != 0; }

    GpuFft(GpuFft&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    GpuFft& operator=(GpuFft&& o) noexcept {
        if (this != &o) {
            wgfft_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
        }
        return *this;
    }
    GpuFft(const GpuFft&) = delete;
    GpuFft& operator=(const GpuFft&) = delete;
    ~GpuFft() override { wgfft_destroy(h_); }

    std::string name() const override { return wgfft_name(); }
    Batch fft(const Batch& inputs) override { return transform(inputs, WGFFT_FORWARD); }
    Batch ifft(const Batch& inputs) override { return transform(inputs, WGFFT_INVERSE); }
    const void* as_any() const override { return this; }

    // FFT(w * x) per signal with a real window of the signal length.
    Batch fft_windowed(const Batch& inputs, const std::vector<float>& window) {
        uint64_t n = 0;
        std::vector<Complex> flat = flatten(inputs, &n);
        if (inputs.empty()) return {};
        if (window.size() != n) throw FftError(WGFFT_ERR_INVALID_ARGUMENT, "window length must equal the signal length");
        std::vector<Complex> out(flat.size());
        check(wgfft_exec_windowed_host(h_, reinterpret_cast<const float*>(flat.data()), window.data(),
                                       reinterpret_cast<float*>(out.data()), n, inputs.size()));
        return split(out, n, inputs.size());
    }

    wgfft_t* handle() const { return h_; }

private:
    explicit GpuFft(wgfft_t* h) : h_(h) {}

    static std::vector<Complex> flatten(const Batch& inputs, uint64_t* n_out) {
        std::vector<uint64_t> lens(inputs.size());
        for (size_t i = 0; i < inputs.size(); ++i) lens[i] = inputs[i].size();
        check(wgfft_validate_lengths(lens.data(), lens.size(), n_out));
        std::vector<Complex> flat;
        flat.reserve(inputs.size() * *n_out);
        for (const auto& v : inputs) flat.insert(flat.end(), v.begin(), v.end());
        return flat;
    }
    static Batch split(const std::vector<Complex>& flat, uint64_t n, size_t batch) {
        Batch out(batch);
        for (size_t b = 0; b < batch; ++b) out[b].assign(flat.begin() + b * n, flat.begin() + (b + 1) * n);
        return out;
    }
    Batch transform(const Batch& inputs, int dir) {
        std::vector<uint64_t> lens(inputs.size());
        for (size_t i = 0; i < inputs.size(); ++i) lens[i] = inputs[i].size();
        uint64_t n = 0;
        check(wgfft_validate_lengths(lens.data(), lens.size(), &n));
        if (inputs.empty()) return {};
        // every signal where it lies: the library's staging copies gather and
        // scatter the rows (wgfft_exec_host_rows), nothing is flattened
        Batch out(inputs.size(), std::vector<Complex>(n));
        std::vector<const float*> in_rows(inputs.size());
        std::vector<float*> out_rows(inputs.size());
        for (size_t b = 0; b < inputs.size(); ++b) {
            in_rows[b] = reinterpret_cast<const float*>(inputs[b].data());
            out_rows[b] = reinterpret_cast<float*>(out[b].data());
        }
        check(wgfft_exec_host_rows(h_, in_rows.data(), out_rows.data(), n, inputs.size(), dir));
        return out;
    }

    wgfft_t* h_ = nullptr;
};

}  // namespace wgls_rs_fft
