// noise.hip -- the reference pipeline's Gaussian noise stage on device.
//
// examples/parallel_pipeline.rs:29-63 (NOISE_WGSL) adds, in place, to sample
// t of signal b: r (cos u2, sin u2) with r = sigma sqrt(-2 ln u1) (Box-Muller),
// u1 = max(h1 / 2^32, 1e-7), u2 = 2 pi h2 / 2^32, h1 = H(t 1234567 + b 7654321
// + seed), h2 = H(h1 + 999983) in wrapping 32-bit arithmetic, where H is the
// PCG-RXS-M-XS integer hash (O'Neill, "PCG: A Family of Simple Fast
// Space-Efficient Statistically Good Algorithms for Random Number
// Generation", 2014; as a GPU hash in Jarzynski and Olano, "Hash Functions
// for GPU Rendering", JCGT 2020).  The demo uses sigma = 0.1 and seed 0.
// The same index mapping and hash are kept so that a signal gets the same
// noise pattern as in the demo (pipeline_noise in fft_device.hpp, shared with
// the fused filter, which adds it to the samples it loads).  The log, sine and
// cosine are the hardware approximations (v_log_f32, v_sin_f32, v_cos_f32),
// as a WGSL implementation may use (WGSL bounds sin and cos only to an absolute
// 2^-11); the samples agree with an FP64 evaluation of the same formula to
// 1e-5 absolute at sigma = 0.1.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/wgfft.h"
#include "runtime_internal.hpp"

namespace wgfft {

// dst may be src (in place): no restrict qualifiers
__global__ void __launch_bounds__(256) k_add_noise(const float2* src, float2* dst, uint32_t log2n, uint64_t total,
                                                   uint32_t seed, float sigma) {
    const uint64_t mask = (1ull << log2n) - 1;
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u)
        dst[e] = cadd(src[e], pipeline_noise((uint32_t)(e & mask), (uint32_t)(e >> log2n), seed, sigma));
}

hipError_t launch_add_noise(const float2* src, float2* dst, uint64_t n, uint64_t batch, uint32_t seed, float sigma,
                            hipStream_t s) {
    uint32_t log2n = 0;
    while ((1ull << log2n) < n) ++log2n;
    const uint64_t total = n * batch;
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_add_noise, dim3(blocks), dim3(256), 0, s, src, dst, log2n, total, seed, sigma);
    return hipGetLastError();
}

}  // namespace wgfft

using namespace wgfft;

extern "C" int wgfft_add_noise_device(wgfft_t* h, float* d_x, uint64_t n, uint64_t batch, uint32_t seed,
                                      float sigma, void* stream) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    if (batch && n > (~0ull) / 8 / batch) return fail(WGFFT_ERR_INVALID_ARGUMENT, "n * batch overflows");
    if (batch == 0) return WGFFT_OK;
    if (!d_x) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((uintptr_t)d_x & 7u) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer must be 8-byte aligned");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = launch_add_noise((const float2*)d_x, (float2*)d_x, n, batch, seed, sigma, (hipStream_t)stream)) !=
        hipSuccess)
        return hip_fail(e, "noise launch");
    return WGFFT_OK;
}
