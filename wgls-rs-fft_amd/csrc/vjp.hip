// vjp.hip -- differentiation of Y = FFT(w * x) (no counterpart in the
// reference, which has no adjoint).
//
//   gu = F^H gY            conjugate-twiddle Stockham pass, no 1/N scale
//   gx = w * gu
//   gw_n = sum_b Re(conj(gu_bn) x_bn)
//
// The VJP entry runs the adjoint transform with the same LDS kernels as the
// inverse (unit scale) and finishes with one elementwise pass that forms gx
// in place and per-chunk partial sums of gw, reduced deterministically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "../../include/wgfft.h"
#include "kernels_adjoint.hpp"
#include "runtime_internal.hpp"

namespace wgfft {

std::vector<AdjointVariant>& adjoint_registry() {
    static std::vector<AdjointVariant> v;
    return v;
}

// The registry default of a length for a role: the first variant claiming
// the role, else the first registered.
const AdjointVariant* find_adjoint(int log2n, int role) {
    for (const auto& v : adjoint_registry())
        if (v.log2n == log2n && (v.roles & role)) return &v;
    for (const auto& v : adjoint_registry())
        if (v.log2n == log2n) return &v;
    return nullptr;
}

const AdjointVariant* pick_adjoint(Handle* h, int log2n, int role) {
    auto& m = role == kRoleVjp ? h->tuned_vjp : role == kRoleFilter ? h->tuned_filter : h->tuned_adj;
    auto it = m.find(log2n);
    return it != m.end() ? it->second : find_adjoint(log2n, role);
}

int Handle::adjoint_blocks_per_cu(const AdjointVariant* v, int kind) {
    auto key = std::make_pair(v, kind);
    auto it = adj_occ_cache.find(key);
    if (it != adj_occ_cache.end()) return it->second;
    const int b = (kind < 2 || kind == 4) ? v->blocks_per_cu(kind != 0, kind == 4) : v->filter_blocks_per_cu(kind == 3);
    adj_occ_cache[key] = b;
    return b;
}

const std::vector<const AdjointVariant*> adjoint_variants(int log2n) {
    std::vector<const AdjointVariant*> out;
    for (const auto& v : adjoint_registry())
        if (v.log2n == log2n) out.push_back(&v);
    return out;
}

const char* adjoint_describe(const AdjointVariant* v, char* buf, uint64_t len) {
    if (buf && len)
        std::snprintf(buf, len, "fused N=%d plan=%s E=%d keep_x=%d%s%s%s%s%s%s%s", 1 << v->log2n, v->plan, v->E,
                      (v->opt & kAdjKeepX) ? 1 : 0, (v->opt & kAdjTwLds) ? " tw-lds" : "",
                      (v->opt & kAdjSplit2) ? " split2" : "",
                      ((v->opt >> 8) & 15) ? (" waves>=" + std::to_string((v->opt >> 8) & 15)).c_str() : "",
                      ((v->opt >> 12) & 63) ? (" chunk=" + std::to_string((v->opt >> 12) & 63)).c_str() : "",
                      ((v->opt >> 18) & 1) ? " prefetch" : "", ((v->opt >> 18) & 2) ? " cached-x" : "", ((v->opt >> 18) & 12) == 12 ? " keep-x"
                                                    : ((v->opt >> 18) & 4) ? " keep-x0"
                                                    : ((v->opt >> 18) & 8) ? " keep-x1" : "");
    return buf;
}

constexpr int kRowChunks = 64;  // batch chunks of the gw partial sums

// gx <- w * gu (in place), part[c][i] = sum over the c-th batch chunk.
__global__ void __launch_bounds__(256)
k_vjp_finish(const float2* __restrict__ x, const float* __restrict__ w, float2* __restrict__ gx,
             float* __restrict__ part, uint64_t n, uint64_t batch, uint64_t rows_per_chunk) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t c = blockIdx.y;
    if (i >= n) return;
    const float wi = w[i];
    const uint64_t b0 = c * rows_per_chunk;
    const uint64_t b1 = min(batch, b0 + rows_per_chunk);
    float acc = 0.f;
    for (uint64_t b = b0; b < b1; ++b) {
        const uint64_t e = b * n + i;
        float2 u = gx[e];
        float2 xv = x[e];
        acc += u.x * xv.x + u.y * xv.y;
        gx[e] = make_float2(wi * u.x, wi * u.y);
    }
    part[c * n + i] = acc;
}

// Column sums of a rows x n matrix in two fixed-order stages: blocks of 64
// columns x 4 row lanes reduce a chunk of rows each (coalesced 256-byte row
// segments), then k_sum_rows folds the chunk sums.
__global__ void __launch_bounds__(256)
k_sum_rows_chunks(const float* __restrict__ part, float* __restrict__ out, uint64_t n, uint64_t rows,
                  uint64_t rows_per_chunk) {
    __shared__ float red[4][64];
    const uint32_t lane = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const uint64_t col = (uint64_t)blockIdx.x * 64 + lane;
    const uint64_t r0 = (uint64_t)blockIdx.y * rows_per_chunk;
    const uint64_t r1 = min(rows, r0 + rows_per_chunk);
    float acc = 0.f;
    if (col < n)
        for (uint64_t r = r0 + rl; r < r1; r += 4) acc += part[r * n + col];
    red[rl][lane] = acc;
    __syncthreads();
    if (rl == 0 && col < n) out[(uint64_t)blockIdx.y * n + col] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

__global__ void __launch_bounds__(256)
k_sum_rows(const float* __restrict__ part, float* __restrict__ out, uint64_t n, uint64_t rows) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float acc = 0.f;
    for (uint64_t r = 0; r < rows; ++r) acc += part[r * n + i];
    out[i] = acc;
}

// gY = 2 m Y in place; per-block partial sums of m |Y|^2.
__global__ void __launch_bounds__(256)
k_mask_grad(float2* __restrict__ y, const float* __restrict__ mask, float* __restrict__ part, uint64_t n,
            uint64_t total) {
    __shared__ float red[256];
    float acc = 0.f;
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const float m = mask[e & (n - 1)];
        float2 v = y[e];
        acc += m * (v.x * v.x + v.y * v.y);
        y[e] = make_float2(2.f * m * v.x, 2.f * m * v.y);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_all(const float* __restrict__ part, float* __restrict__ out, int count) {
    __shared__ float red[256];
    float acc = 0.f;
    for (int i = threadIdx.x; i < count; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Column sums of the rows x ld partial matrix in one launch: a block is
// COLS columns x 1024 / COLS row lanes, each lane striding the rows with
// eight loads in flight; lanes combine in a fixed order through LDS (no
// cross-block ordering, so no device-scope fences).  Columns < n go to gw;
// when loss != NULL, column n holds per-row loss partials summed into
// loss[0].  Measured on partial matrices of 256-1024 rows: 32 columns (129
// blocks at N = 4096) for N <= 4096 (768 x 4100: 3.1 against 3.8 us with
// 64), 64 columns above (512 x 8196: 4.1 against 5.1 us with 32).
template <int COLS>
__global__ void __launch_bounds__(1024)
k_reduce_cols(const float* __restrict__ part, uint64_t ld, uint64_t rows, uint64_t n, float* __restrict__ gw,
              float* __restrict__ loss) {
    constexpr int LANES = 1024 / COLS;
    __shared__ float red[LANES][COLS];
    const uint32_t lane = threadIdx.x % COLS, rl = threadIdx.x / COLS;
    const uint64_t col = (uint64_t)blockIdx.x * COLS + lane;
    const uint64_t ncols = loss ? n + 1 : n;
    float acc = 0.f;
    if (col < ncols) {
        uint64_t r = rl;
        for (; r + 7 * LANES < rows; r += 8 * LANES) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = part[(r + (uint64_t)LANES * i) * ld + col];
            acc += ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
        }
        for (; r < rows; r += LANES) acc += part[r * ld + col];
    }
    red[rl][lane] = acc;
    __syncthreads();
    for (int w = LANES / 2; w > 0; w >>= 1) {
        if ((int)rl < w) red[rl][lane] += red[rl + w][lane];
        __syncthreads();
    }
    if (rl == 0 && col < ncols) {
        if (col < n) gw[col] = red[0][lane];
        else loss[0] = red[0][lane];
    }
}

// out[i] = sum_r part[r][i] (deterministic order).
static int reduce_rows(Handle* h, const float* part, uint64_t rows, uint64_t n, float* out, hipStream_t s) {
    const uint64_t chunks = std::min<uint64_t>(rows, 16);
    const uint64_t per = (rows + chunks - 1) / chunks;
    int rc;
    if ((rc = h->ensure(h->red2, chunks * n * sizeof(float)))) return rc;
    dim3 grid((unsigned)((n + 63) / 64), (unsigned)chunks);
    hipLaunchKernelGGL(k_sum_rows_chunks, grid, dim3(256), 0, s, part, (float*)h->red2.p, n, rows, per);
    hipLaunchKernelGGL(k_sum_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const float*)h->red2.p, out, n,
                       chunks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "row reduction launch");
    return WGFFT_OK;
}

// gx, gw from gu (already in d_gx).
static int vjp_finish(Handle* h, const float* d_x, const float* d_w, float* d_gx, float* d_gw, uint64_t n,
                      uint64_t batch, hipStream_t s) {
    const uint64_t chunks = std::min<uint64_t>(batch, kRowChunks);
    const uint64_t rows = (batch + chunks - 1) / chunks;
    int rc;
    if ((rc = h->ensure(h->red, chunks * n * sizeof(float)))) return rc;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)chunks);
    hipLaunchKernelGGL(k_vjp_finish, grid, dim3(256), 0, s, (const float2*)d_x, d_w, (float2*)d_gx,
                       (float*)h->red.p, n, batch, rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "vjp finish launch");
    return reduce_rows(h, (const float*)h->red.p, chunks, n, d_gw, s);
}

// One fused launch (plus the fixed-order reductions).  fwd = true: loss
// L = sum m |FFT(w x)|^2 with gY = 2 m Y formed in registers; fwd = false:
// the given gY.
int fused_adjoint(Handle* h, const AdjointVariant* v, bool fwd, const float* d_x, const float* d_w,
                         const float* d_mask, const float* d_gy, float* d_y, float* d_gx, float* d_gw,
                         float* d_loss, uint64_t n, uint64_t batch, const float2* tw, hipStream_t s) {
    // partial matrix: one row of n + 4 floats per workgroup (gw partials,
    // then the loss partial in column n; rows kept 16-byte aligned)
    const uint64_t ld = n + 4;
    int grid = h->adjoint_blocks_per_cu(v, fwd ? (d_y ? 4 : 1) : 0) * h->num_cus;
    if ((uint64_t)grid > batch) grid = (int)batch;
    const uint64_t rows = (uint64_t)grid;
    int rc;
    if ((rc = h->ensure(h->red, rows * ld * sizeof(float)))) return rc;
    AdjointArgs a{(const float2*)d_x, d_w, d_mask, (const float2*)d_gy, (float2*)d_y, (float2*)d_gx, tw,
                  (float*)h->red.p, ld, batch, grid, s};
    hipError_t e = v->launch(fwd, a);
    if (e != hipSuccess) return hip_fail(e, "adjoint kernel launch");
    const uint64_t ncols = fwd ? n + 1 : n;
    if (n <= 4096)
        hipLaunchKernelGGL(k_reduce_cols<32>, dim3((unsigned)((ncols + 31) / 32)), dim3(1024), 0, s,
                           (const float*)h->red.p, ld, rows, n, d_gw, fwd ? d_loss : nullptr);
    else
        hipLaunchKernelGGL(k_reduce_cols<64>, dim3((unsigned)((ncols + 63) / 64)), dim3(1024), 0, s,
                           (const float*)h->red.p, ld, rows, n, d_gw, fwd ? d_loss : nullptr);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adjoint reduction launch");
    return WGFFT_OK;
}

// Zero fill by a kernel (not hipMemsetAsync), so the empty-batch results can
// be recorded into a captured graph like every other launch.
__global__ void __launch_bounds__(256) k_zero_fill(float* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) p[i] = 0.f;
}

static hipError_t zero_fill(float* p, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_fill, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// y[b][k] *= s[k]: the unfused spectral multiply for lengths without a fused
// filter kernel.
__global__ void __launch_bounds__(256) k_spectral_mul(float2* __restrict__ y, const float* __restrict__ sp,
                                                      uint64_t n, uint64_t total) {
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u) {
        const float m = sp[e & (n - 1)];
        const float2 v = y[e];
        y[e] = make_float2(m * v.x, m * v.y);
    }
}

}  // namespace wgfft

using namespace wgfft;

static int vjp_checks(wgfft_t* h, uint64_t n, uint64_t batch) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    if (batch && n > (~0ull) / 8 / batch) return fail(WGFFT_ERR_INVALID_ARGUMENT, "n * batch overflows");
    return WGFFT_OK;
}

extern "C" int wgfft_vjp_windowed_device(wgfft_t* h, const float* d_x, const float* d_w, const float* d_gy,
                                         float* d_gx, float* d_gw, uint64_t n, uint64_t batch, void* stream) {
    int rc;
    if ((rc = vjp_checks(h, n, batch))) return rc;
    if (!d_x || !d_w || !d_gy || !d_gx || !d_gw) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((((uintptr_t)d_x | (uintptr_t)d_gy | (uintptr_t)d_gx | (uintptr_t)d_w) & 15u) != 0)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffers must be 16-byte aligned");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;  // NULL: HIP default stream
    ScratchUse su(h, s);
    if (su.e != hipSuccess) return hip_fail(su.e, "stream ordering of the scratch buffers");
    if (batch == 0) {
        if ((e = zero_fill(d_gw, n, s)) != hipSuccess) return hip_fail(e, "zero fill of gw");
        return WGFFT_OK;
    }
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind == Plan::kBatched) {
        const AdjointVariant* v = pick_adjoint(h, p->log2n, kRoleVjp);
        if (v)
            return fused_adjoint(h, v, false, d_x, d_w, nullptr, d_gy, nullptr, d_gx, d_gw, nullptr, n, batch, p->tw,
                                 s);
    }
    // gu = F^H gY: inverse-sign transform with unit scale
    if ((rc = h->run(p, kInverse, (const float2*)d_gy, (float2*)d_gx, nullptr, batch, 1.0f, s))) return rc;
    return vjp_finish(h, d_x, d_w, d_gx, d_gw, n, batch, s);
}

extern "C" int wgfft_fwd_adjoint_masked_device(wgfft_t* h, const float* d_x, const float* d_w, const float* d_mask,
                                               float* d_y, float* d_gx, float* d_gw, float* d_loss, uint64_t n,
                                               uint64_t batch, void* stream) {
    int rc;
    if ((rc = vjp_checks(h, n, batch))) return rc;
    if (!d_x || !d_w || !d_mask || !d_gx || !d_gw || !d_loss)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((((uintptr_t)d_x | (uintptr_t)d_gx | (uintptr_t)d_w | (uintptr_t)d_y) & 15u) != 0)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffers must be 16-byte aligned");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;  // NULL: HIP default stream
    ScratchUse su(h, s);
    if (su.e != hipSuccess) return hip_fail(su.e, "stream ordering of the scratch buffers");
    if (batch == 0) {
        if ((e = zero_fill(d_gw, n, s)) != hipSuccess || (e = zero_fill(d_loss, 1, s)) != hipSuccess)
            return hip_fail(e, "zero fill of gw and loss");
        return WGFFT_OK;
    }
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind == Plan::kBatched) {
        const AdjointVariant* v = pick_adjoint(h, p->log2n, kRoleFused);
        if (v)
            return fused_adjoint(h, v, true, d_x, d_w, d_mask, nullptr, d_y, d_gx, d_gw, d_loss, n, batch, p->tw, s);
    }
    // lengths without a fused kernel: forward, mask, adjoint as separate passes
    const uint64_t total = n * batch;
    // Y (into d_y or scratch), then gY = 2 m Y in the scratch copy
    if ((rc = h->ensure(h->io_aux, total * sizeof(float2)))) return rc;
    float2* gy = (float2*)h->io_aux.p;
    if ((rc = h->run(p, kWindowed, (const float2*)d_x, gy, d_w, batch, 1.0f, s))) return rc;
    if (d_y) {
        e = hipMemcpyAsync(d_y, gy, total * sizeof(float2), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(Y)");
    }
    const int blocks = (int)std::min<uint64_t>((total + 255) / 256, 4096);
    if ((rc = h->ensure(h->lossbuf, blocks * sizeof(float) + 256))) return rc;
    hipLaunchKernelGGL(k_mask_grad, dim3(blocks), dim3(256), 0, s, gy, d_mask, (float*)h->lossbuf.p, n, total);
    hipLaunchKernelGGL(k_sum_all, dim3(1), dim3(256), 0, s, (const float*)h->lossbuf.p, d_loss, blocks);
    if ((rc = h->run(p, kInverse, gy, (float2*)d_gx, nullptr, batch, 1.0f, s))) return rc;
    return vjp_finish(h, d_x, d_w, d_gx, d_gw, n, batch, s);
}

extern "C" int wgfft_adjoint_variant_count(uint64_t n) {
    if (n == 0 || (n & (n - 1))) return 0;
    int log2n = 0;
    while ((1ull << log2n) < n) ++log2n;
    return (int)adjoint_variants(log2n).size();
}

extern "C" int wgfft_select_adjoint_variant(wgfft_t* h, uint64_t n, int index, char* desc, uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    int log2n = 0;
    while ((1ull << log2n) < n) ++log2n;
    const auto vs = adjoint_variants(log2n);
    if (index < 0 || index >= (int)vs.size()) return fail(WGFFT_ERR_INVALID_ARGUMENT, "variant index out of range");
    std::lock_guard<std::mutex> lk(h->mu);
    h->tuned_adj[log2n] = h->tuned_vjp[log2n] = h->tuned_filter[log2n] = vs[(size_t)index];
    adjoint_describe(vs[(size_t)index], desc, desc_len);
    return WGFFT_OK;
}

// y = F^{-1}(s * F(x + noise)), 1/N scaled: the reference pipeline's noise
// (sigma != 0), FFT, spectral window and IFFT + 1/N stages
// (examples/parallel_pipeline.rs:29-105,586-642).
// One fused pass where the length has fused kernels (N = 512 .. 16384), else
// forward, multiply and inverse launches on device.
extern "C" int wgfft_filter_device(wgfft_t* h, const float* d_x, const float* d_s, float* d_y, uint64_t n,
                                   uint64_t batch, uint32_t seed, float sigma, void* stream) {
    int rc;
    if ((rc = vjp_checks(h, n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!d_x || !d_s || !d_y) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((((uintptr_t)d_x | (uintptr_t)d_y | (uintptr_t)d_s) & 15u) != 0)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffers must be 16-byte aligned");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;  // NULL: HIP default stream
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind == Plan::kBatched) {
        const AdjointVariant* v = pick_adjoint(h, p->log2n, kRoleFilter);
        if (v) {
            int grid = h->adjoint_blocks_per_cu(v, sigma != 0.f ? 3 : 2) * h->num_cus;
            if ((uint64_t)grid > batch) grid = (int)batch;
            FilterArgs a{(const float2*)d_x, d_s, (float2*)d_y, p->tw, batch, grid, s, seed, sigma};
            if ((e = v->filter(a)) != hipSuccess) return hip_fail(e, "filter kernel launch");
            return WGFFT_OK;
        }
    }
    const float2* src = (const float2*)d_x;
    if (sigma != 0.f) {  // y = x + noise, then the forward in place
        if ((e = launch_add_noise(src, (float2*)d_y, n, batch, seed, sigma, s)) != hipSuccess)
            return hip_fail(e, "noise launch");
        src = (const float2*)d_y;
    }
    if ((rc = h->run(p, kForward, src, (float2*)d_y, nullptr, batch, 1.0f, s))) return rc;
    const uint64_t total = n * batch;
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_spectral_mul, dim3(blocks), dim3(256), 0, s, (float2*)d_y, d_s, n, total);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "spectral multiply launch");
    return h->run(p, kInverse, (const float2*)d_y, (float2*)d_y, nullptr, batch, 1.0f / (float)n, s);
}
