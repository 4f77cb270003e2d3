// runtime.hip -- host runtime behind include/wgfft.h: handles, per-length
// plans (kernel variant + twiddle tables), device scratch, validation and the
// C-ABI entry points.  Mirrors wgls-rs-fft's GpuFft runtime (src/lib.rs):
// the per-N SizeCache becomes a Plan, the ping-pong buffers become either one
// LDS-resident launch (N <= 16384) or a chain of global passes, and the
// conjugate-and-scale inverse of src/lib.rs:571-579,661-669 is done in-kernel
// with conjugated twiddles and a 1/N scale at the final store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wgfft.h"
#include "kernels.hpp"
#include "runtime_internal.hpp"

namespace wgfft {

std::vector<BatchedVariant>& batched_registry() {
    static std::vector<BatchedVariant> v;
    return v;
}
std::vector<GlobalPassVariant>& global_pass_registry() {
    static std::vector<GlobalPassVariant> v;
    return v;
}

thread_local std::string g_last_error;
static std::atomic<uint64_t> g_alloc_cap{0};

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(WGFFT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

const char* kMsgLength = "Transform length must be a non-zero power of two";
const char* kMsgRagged = "All input vectors in a batch must have the same length";

static bool is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }
static int log2u(uint64_t n) {
    int l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

hipError_t Handle::scratch_begin(hipStream_t st) {
    if (!scratch_pending || st == scratch_stream) return hipSuccess;  // one stream orders its own work
    if (scratch_recorded) return hipStreamWaitEvent(st, scratch_ev, 0);
    // the other stream's last use has no event after it: wait for the device
    // once (not possible inside a capture), then record after every use
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipErrorStreamCaptureUnsupported;
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    scratch_pending = false;
    scratch_events = true;
    return hipSuccess;
}

// The event only orders device work between streams of this device, so it
// releases at device scope (the default system-scope fence adds ~1 us of GPU
// idle time per record).
void Handle::scratch_end(hipStream_t st) {
    scratch_stream = st;
    scratch_pending = true;
    scratch_recorded = false;
    if (!scratch_events) return;
    if (!scratch_ev &&
        hipEventCreateWithFlags(&scratch_ev, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
        (void)hipGetLastError();
        scratch_ev = nullptr;
        return;
    }
    if (hipEventRecord(scratch_ev, st) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    scratch_recorded = true;
}

hipError_t Handle::scratch_quiesce() {
    if (!scratch_pending || scratch_stream == stream || scratch_recorded) return hipSuccess;
    const hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) scratch_pending = false;
    return e;
}

void Handle::scratch_settled(hipStream_t st) {
    if (scratch_pending && st == scratch_stream) {
        scratch_pending = false;
        scratch_recorded = false;
    }
}

int Handle::ensure(DevBuf& b, uint64_t bytes) {
    if (b.bytes >= bytes && b.p) return WGFFT_OK;
    uint64_t cap = g_alloc_cap.load();
    uint64_t after = held - b.bytes + bytes;
    if (cap && after > cap)
        return fail(WGFFT_ERR_OUT_OF_MEMORY, "device scratch of " + std::to_string(after) +
                                                 " bytes exceeds the allocation cap of " + std::to_string(cap));
    if (b.p) {
        hipFree(b.p);
        held -= b.bytes;
        b.p = nullptr;
        b.bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(WGFFT_ERR_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    b.p = p;
    b.bytes = bytes;
    held += bytes;
    return WGFFT_OK;
}

void Handle::release(DevBuf& b) {
    if (b.p) {
        hipFree(b.p);
        held -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

// W_N^j = exp(-2 pi i j / N) for j < count, FP64 then rounded to f32.
static std::vector<float2> twiddle_table(uint64_t N, uint64_t count, uint64_t step) {
    std::vector<float2> t(count);
    for (uint64_t j = 0; j < count; ++j) {
        uint64_t e = (j * step) % N;
        double a = 2.0 * M_PI * ((double)e / (double)N);
        t[j] = make_float2((float)std::cos(a), (float)-std::sin(a));
    }
    return t;
}

int Handle::upload_table(const std::vector<float2>& host, float2** dev) {
    void* p = nullptr;
    uint64_t bytes = host.size() * sizeof(float2);
    uint64_t cap = g_alloc_cap.load();
    if (cap && held + bytes > cap) return fail(WGFFT_ERR_OUT_OF_MEMORY, "twiddle table exceeds the allocation cap");
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(WGFFT_ERR_OUT_OF_MEMORY, std::string("hipMalloc(twiddles) failed: ") + hipGetErrorString(e));
    }
    held += bytes;
    e = hipMemcpy(p, host.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(p);
        held -= bytes;
        return hip_fail(e, "hipMemcpy(twiddles)");
    }
    tables.push_back({p, bytes});
    *dev = (float2*)p;
    return WGFFT_OK;
}

// Default kernel variant of each single-kernel length: the autotuner's pick
// on MI355X at about 2^24 samples per call (7% to 58% faster than the first
// registration, which remains the fallback); at 2^13 the split kernel, which
// is 3% behind at 8192 x 2048 but 24% ahead at 8192 x 1024.  An entry made by
// wgfft_autotune or read from the tune file overrides it.
namespace {
struct DefaultBatched {
    int log2n;
    const char* plan;
    int E, G, flags;
};
const DefaultBatched kDefaultBatched[] = {
    {5, "8x4", 16, 128, 0},       {6, "8x8", 16, 64, 0},          {7, "4x4x8", 16, 32, 0},
    {8, "8x4x8", 16, 16, 0},      {9, "8x8x8", 16, 8, 0},         {10, "16x4x16", 16, 4, kFlagPrio},
    {11, "16x8x16", 16, 2, kFlagPrio},    {12, "16x16x16", 16, 1, kFlagNtStore | kFlagPrio | kFlagHalfPrefetch | kFlagQuarterLoad},
    {13, "2x16x16x16", 16, 1, kFlagSplit2},                        {14, "2x16x8x4x16", 16, 1, kFlagSplit2 | kFlagCachedLoad},
};

// Default multipass plans of 2^15 .. 2^26, picked the same way (1.2x to 5.8x
// faster than the fewest-pass split with the first-registered variants):
// per pass the radix, sub-FFT plan, column group and pipeline kind (the
// registry's prefetch code: 0 staged, 3 direct, 5 16-column rounds).
struct DefaultPass {
    int log2r;
    const char* plan;
    int G, kind;
};
struct DefaultSplit {
    int log2n, passes;
    DefaultPass p[3];
};
const DefaultSplit kDefaultSplits[] = {
    {15, 2, {{8, "16x16", 32, 3}, {7, "4x8x4", 32, 3}}},
    {16, 2, {{8, "16x16", 32, 3}, {8, "16x16", 32, 3}}},
    {17, 2, {{9, "8x8x8", 16, 3}, {8, "16x16", 32, 3}}},
    {18, 2, {{9, "8x8x8", 16, 3}, {9, "8x8x8", 16, 3}}},
    {19, 2, {{10, "16x4x16", 16, 5}, {9, "8x8x8", 16, 3}}},
    {20, 2, {{10, "16x4x16", 16, 5}, {10, "16x4x16", 16, 5}}},
    {21, 2, {{11, "16x8x16", 16, 5}, {10, "16x4x16", 16, 5}}},
    {22, 2, {{11, "16x8x16", 16, 5}, {11, "16x8x16", 16, 5}}},
    {23, 2, {{12, "16x16x16", 4, 0}, {11, "16x8x16", 16, 5}}},
    {24, 3, {{8, "16x16", 32, 3}, {8, "16x16", 32, 3}, {8, "16x16", 32, 3}}},
    {25, 3, {{9, "8x8x8", 16, 3}, {8, "16x16", 32, 3}, {8, "16x16", 16, 3}}},
    {26, 3, {{9, "16x16x2", 16, 0}, {9, "16x16x2", 16, 0}, {8, "16x16", 32, 3}}},
};

const DefaultSplit* default_split(int log2n) {
    for (const auto& d : kDefaultSplits)
        if (d.log2n == log2n) return &d;
    return nullptr;
}

// One transform per workgroup for small batches of 512 .. 2048: measured
// faster than the defaults (tuned at 2^24 samples per call) for every batch
// up to 1024 and equal or faster up to 4096 at 1024 and 2048 (N=1024 x 8:
// 3.4 against 4.1 us per call; 512 x 1024: 3.5 against 4.5 us).
const DefaultBatched kSmallBatch[] = {
    {9, "8x8x8", 8, 1, 0}, {10, "16x4x16", 16, 1, kFlagPrio}, {11, "16x8x16", 16, 1, kFlagPrio}};
}  // namespace

const BatchedVariant* small_batch_variant(int log2n) {
    for (const auto& d : kSmallBatch)
        if (d.log2n == log2n) return find_variant(log2n, d.plan, d.E, d.G, d.flags);
    return nullptr;
}

const BatchedVariant* default_variant(int log2n) {
    for (const auto& d : kDefaultBatched)
        if (d.log2n == log2n)
            if (const BatchedVariant* v = find_variant(log2n, d.plan, d.E, d.G, d.flags)) return v;
    for (const auto& v : batched_registry())
        if (v.log2n == log2n) return &v;
    return nullptr;
}

const BatchedVariant* find_variant(int log2n, const char* plan, int E, int G, int flags) {
    for (const auto& v : batched_registry())
        if (v.log2n == log2n && v.E == E && v.G == G && v.flags == flags && std::strcmp(v.plan, plan) == 0) return &v;
    return nullptr;
}

// Global-pass variants usable for a pass of radix 2^log2r at stride L over
// nr butterflies (G <= nr, and G <= L after the first pass); the first is
// the default.
std::vector<const GlobalPassVariant*> global_pass_candidates(int log2r, uint64_t L, uint64_t nr) {
    std::vector<const GlobalPassVariant*> out;
    for (const auto& v : global_pass_registry()) {
        if (v.log2r != log2r) continue;
        if ((uint64_t)v.G > nr) continue;
        if (L > 1 && (uint64_t)v.G > L) continue;
        out.push_back(&v);
    }
    return out;
}

// Balanced radix split (log2 of each pass radix, larger first) of 2^log2n
// into passes of at most 2^maxr.
static std::vector<int> balanced_split(int log2n, int maxr) {
    std::vector<int> r;
    const int passes = (log2n + maxr - 1) / maxr;
    int rem = log2n;
    for (int i = 0; i < passes; ++i) {
        const int ri = (rem + (passes - i) - 1) / (passes - i);
        r.push_back(ri);
        rem -= ri;
    }
    return r;
}

// Default global-pass split of 2^log2n: the measured one where there is one,
// else radices up to 2^9 (the measured splits above 2^23 are all of that
// kind: more passes, but tiles small enough for 128- to 256-byte runs).
std::vector<int> multipass_radices(int log2n) {
    if (const DefaultSplit* d = default_split(log2n)) {
        std::vector<int> r;
        for (int i = 0; i < d->passes; ++i) r.push_back(d->p[i].log2r);
        return r;
    }
    return balanced_split(log2n, 9);
}

// Splits the autotuner times: the default, the fewest passes (radices up to
// 2^13) and the ones with radices up to 2^9 and 2^8 (one or two more passes
// over memory, but tiles small enough for several workgroups per CU and
// 128-byte column runs).
std::vector<std::vector<int>> multipass_splits(int log2n) {
    std::vector<std::vector<int>> out;
    for (int maxr : {0, 13, 9, 8}) {
        std::vector<int> r = maxr ? balanced_split(log2n, maxr) : multipass_radices(log2n);
        bool dup = false;
        for (const auto& o : out) dup |= (o == r);
        bool ok = true;
        uint64_t L = 1;
        for (int ri : r) {
            ok &= !global_pass_candidates(ri, L, (1ull << log2n) >> ri).empty();
            L <<= ri;
        }
        if (!dup && ok) out.push_back(r);
    }
    return out;
}

// Default variant of pass i of split r: the measured one when r is the
// measured split of this length; otherwise the widest column group (the
// column-run width bounds these passes), preferring for a pass after the
// first the 16-column rounds, then direct tile I/O.
static const GlobalPassVariant* default_pass_variant(int log2n, const std::vector<int>& r, size_t i, uint64_t L,
                                                     const std::vector<const GlobalPassVariant*>& cands) {
    if (const DefaultSplit* d = default_split(log2n)) {
        bool same = (int)r.size() == d->passes;
        for (int q = 0; same && q < d->passes; ++q) same = r[q] == d->p[q].log2r;
        if (same) {
            const DefaultPass& want = d->p[i];
            for (const GlobalPassVariant* v : cands)
                if (v->G == want.G && v->prefetch == want.kind && std::strcmp(v->plan, want.plan) == 0) return v;
        }
    }
    auto rank = [L](const GlobalPassVariant* v) {
        const int k = v->prefetch;
        const int pref = L > 1 ? (k == 5 ? 3 : k == 3 ? 2 : k == 0 ? 1 : 0) : (k == 0 ? 3 : k == 3 ? 2 : k == 5 ? 1 : 0);
        return v->G * 4 + pref;
    };
    const GlobalPassVariant* best = cands.front();
    for (const GlobalPassVariant* v : cands)
        if (rank(v) > rank(best)) best = v;
    return best;
}

// Give a multipass plan the radix split r (default variant per pass) and
// upload the sub-FFT tables it needs.
int Handle::set_split(Plan* p, const std::vector<int>& r) {
    int rc;
    std::vector<const GlobalPassVariant*> vars;
    uint64_t L = 1;
    for (int ri : r) {
        auto cands = global_pass_candidates(ri, L, p->n >> ri);
        if (cands.empty()) return fail(WGFFT_ERR_INVALID_LENGTH, "no global-pass variant for radix 2^" + std::to_string(ri));
        vars.push_back(default_pass_variant(p->log2n, r, vars.size(), L, cands));
        if (!p->tw_r.count(ri)) {
            float2* t = nullptr;
            if ((rc = upload_table(twiddle_table(1ull << ri, 1ull << ri, 1), &t))) return rc;
            p->tw_r[ri] = t;
        }
        L <<= ri;
    }
    p->pass_log2r = r;
    p->pass_var = vars;
    return WGFFT_OK;
}

int Handle::get_plan(uint64_t n, Plan** out) {
    auto it = plans.find(n);
    if (it != plans.end()) {
        *out = &it->second;
        return WGFFT_OK;
    }
    Plan p;
    p.n = n;
    p.log2n = log2u(n);
    int rc;
    if (p.log2n <= 4) {
        p.kind = Plan::kSmall;
    } else if (p.log2n <= 14) {
        p.kind = Plan::kBatched;
        const bool pinned = tuned.count(p.log2n) != 0;
        p.var = pinned ? tuned[p.log2n] : default_variant(p.log2n);
        if (!p.var) return fail(WGFFT_ERR_INVALID_LENGTH, "no kernel variant for this length");
        if (!pinned) p.var_small = small_batch_variant(p.log2n);
        if (pinned && tuned_win.count(p.log2n)) p.var_win = tuned_win[p.log2n];
        if ((rc = upload_table(twiddle_table(n, n, 1), &p.tw))) return rc;
    } else {
        if (p.log2n > 39) return fail(WGFFT_ERR_INVALID_LENGTH, "transform length above 2^39 is not supported");
        p.kind = Plan::kMultiPass;
        if ((rc = set_split(&p, multipass_radices(p.log2n)))) return rc;
        p.h = (uint32_t)((p.log2n + 1) / 2);
        uint64_t lo_n = 1ull << p.h, hi_n = 1ull << (p.log2n - p.h);
        if ((rc = upload_table(twiddle_table(n, lo_n, 1), &p.tw_lo))) return rc;
        if ((rc = upload_table(twiddle_table(n, hi_n, lo_n), &p.tw_hi))) return rc;
    }
    auto res = plans.emplace(n, p);
    *out = &res.first->second;
    return WGFFT_OK;
}

int Handle::blocks_per_cu(const BatchedVariant* v, Mode m, bool cache) {
    auto key = std::make_pair(v, (int)m);
    auto it = occ_cache.find(key);
    if (it != occ_cache.end()) return it->second;
    int b = v->blocks_per_cu(m);
    if (cache) occ_cache[key] = b;
    return b;
}

// Enqueue one batched transform of `batch` length-n signals.  With fixed
// scratch (fixed != nullptr, fixed_bytes of device memory) nothing is
// allocated: the multipass intermediates are carved from it and the batch is
// cut into chunks that fit; the caller has checked that one transform fits.
// idist / odist (complex elements between transforms, 0 = n) are honoured
// by the single-kernel lengths only (n <= 16384).
int Handle::run(Plan* p, Mode mode, const float2* in, float2* out, const float* win, uint64_t batch,
                float scale, hipStream_t s, const BatchedVariant* force, float2* fixed, uint64_t fixed_bytes,
                uint64_t idist, uint64_t odist) {
    if (batch == 0) return WGFFT_OK;
    hipError_t e = hipSuccess;
    if (p->kind == Plan::kMultiPass && ((idist && idist != p->n) || (odist && odist != p->n)))
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "signal distances other than n need n <= 16384");
    if (p->kind == Plan::kSmall) {
        BatchedArgs a{in, out, nullptr, win, batch, scale, 0, s, idist, odist};
        e = launch_small(mode, p->log2n, a);
        if (e != hipSuccess) return hip_fail(e, "small-length kernel launch");
        return WGFFT_OK;
    }
    if (p->kind == Plan::kBatched) {
        const BatchedVariant* v = force                                                      ? force
                                  : p->var_small && batch <= kSmallBatchPerCu * (uint64_t)num_cus ? p->var_small
                                  : mode == kWindowed && p->var_win                              ? p->var_win
                                                                                                 : p->var;
        int grid = blocks_per_cu(v, mode, fixed == nullptr) * num_cus;
        BatchedArgs a{in, out, p->tw, win, batch, scale, grid, s, idist, odist};
        e = v->launch(mode, a);
        if (e != hipSuccess) return hip_fail(e, "batched kernel launch");
        return WGFFT_OK;
    }
    // multi-pass: chunk the batch so the scratch intermediates stay bounded
    const uint64_t tbytes = p->n * sizeof(float2);
    const int P = (int)p->pass_var.size();
    const bool two = multipass_needs_two_buffers(P);
    float2 *buf_a, *buf_b;
    uint64_t chunk;
    if (fixed) {
        const uint64_t per = two ? fixed_bytes / 2 : fixed_bytes;
        chunk = std::min<uint64_t>(per / tbytes, batch);
        if (chunk == 0) return fail(WGFFT_ERR_INVALID_ARGUMENT, "the transform's intermediates exceed the fixed scratch");
        buf_a = fixed;
        buf_b = two ? fixed + (per / sizeof(float2)) : nullptr;
        return run_passes(p, mode, in, out, win, batch, scale, s, buf_a, buf_b, chunk);
    }
    chunk = std::min<uint64_t>(std::max<uint64_t>(1, mp_chunk_bytes / tbytes), batch);
    int rc;
    ScratchUse su(this, s);
    if (su.e != hipSuccess) return hip_fail(su.e, "stream ordering of the scratch buffers");
    if ((rc = ensure(scratch, chunk * tbytes))) return rc;
    if (two && (rc = ensure(scratch2, chunk * tbytes))) return rc;
    buf_a = (float2*)scratch.p;
    buf_b = two ? (float2*)scratch2.p : nullptr;
    return run_passes(p, mode, in, out, win, batch, scale, s, buf_a, buf_b, chunk);
}

// The intermediates of a chain of P passes: pass i < P - 1 writes buf_a when
// i is even and buf_b when odd, so three passes or more need both (two
// passes, in place or not, read the input and write the output directly).
bool multipass_needs_two_buffers(int passes) { return passes >= 3; }

// The global passes of a multipass plan over `batch` transforms, `chunk`
// transforms at a time through the intermediates buf_a (and buf_b).
int Handle::run_passes(Plan* p, Mode mode, const float2* in, float2* out, const float* win, uint64_t batch,
                       float scale, hipStream_t s, float2* buf_a, float2* buf_b, uint64_t chunk) {
    const int P = (int)p->pass_var.size();
    for (uint64_t b0 = 0; b0 < batch; b0 += chunk) {
        const uint64_t nb = std::min(chunk, batch - b0);
        const float2* src = in + b0 * p->n;
        float2* dst_final = out + b0 * p->n;
        uint64_t L = 1;
        for (int i = 0; i < P; ++i) {
            float2* dst = (i == P - 1) ? dst_final : (i % 2 == 0) ? buf_a : buf_b;
            const GlobalPassVariant* v = p->pass_var[i];
            GlobalPassArgs a{src, dst, p->tw_r[p->pass_log2r[i]], p->tw_hi, p->tw_lo, win, (uint32_t)p->log2n, p->h, L,
                             nb, (i == P - 1) ? scale : 1.0f, num_cus, s};
            Mode m = mode;
            if (mode == kWindowed && i > 0) m = kForward;
            const hipError_t e = v->launch(m, a);
            if (e != hipSuccess) return hip_fail(e, "global pass launch");
            src = dst;
            L <<= p->pass_log2r[i];
        }
    }
    return WGFFT_OK;
}

// Strided layouts: element j of signal b at base[b dist + j stride] (complex
// elements).  Gathered into a contiguous buffer, transformed in place there,
// scattered out; n is a power of two.
__global__ void __launch_bounds__(256) k_gather(const float2* __restrict__ src, float2* __restrict__ dst,
                                                uint32_t log2n, uint64_t total, uint64_t stride, uint64_t dist) {
    const uint64_t mask = (1ull << log2n) - 1;
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u)
        dst[e] = src[(e >> log2n) * dist + (e & mask) * stride];
}

__global__ void __launch_bounds__(256) k_scatter(const float2* __restrict__ src, float2* __restrict__ dst,
                                                 uint32_t log2n, uint64_t total, uint64_t stride, uint64_t dist) {
    const uint64_t mask = (1ull << log2n) - 1;
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256u)
        dst[(e >> log2n) * dist + (e & mask) * stride] = src[e];
}

// Adjacent signals (distance 1, any element stride): column layouts.  The
// element-by-element kernels above then read (gather) or write (scatter) one
// 8-byte value per row touched.  Here a 32 x 32 tile of (signal, element)
// pairs moves along the signals on the strided side (256-byte row segments)
// and along the elements on the contiguous side, transposed through LDS.
// Tiles are numbered in one dimension (n / 32 exceeds the grid's y limit at
// large n) and visited grid-stride.
constexpr uint32_t kTile = 32;

__global__ void __launch_bounds__(256) k_gather_adjacent(const float2* __restrict__ src, float2* __restrict__ dst,
                                                         uint64_t n, uint64_t batch, uint64_t stride) {
    __shared__ float2 tile[kTile][kTile + 1];
    const uint64_t tb = (batch + kTile - 1) / kTile, tiles = tb * ((n + kTile - 1) / kTile);
    const uint32_t lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t b0 = (t % tb) * kTile, j0 = (t / tb) * kTile;
        for (uint32_t r = ly; r < kTile; r += 256 / kTile) {  // src[b + j stride], b fastest
            const uint64_t b = b0 + lx, j = j0 + r;
            if (b < batch && j < n) tile[r][lx] = src[b + j * stride];
        }
        __syncthreads();
        for (uint32_t r = ly; r < kTile; r += 256 / kTile) {  // dst[b n + j], j fastest
            const uint64_t b = b0 + r, j = j0 + lx;
            if (b < batch && j < n) dst[b * n + j] = tile[lx][r];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_scatter_adjacent(const float2* __restrict__ src, float2* __restrict__ dst,
                                                          uint64_t n, uint64_t batch, uint64_t stride) {
    __shared__ float2 tile[kTile][kTile + 1];
    const uint64_t tb = (batch + kTile - 1) / kTile, tiles = tb * ((n + kTile - 1) / kTile);
    const uint32_t lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t b0 = (t % tb) * kTile, j0 = (t / tb) * kTile;
        for (uint32_t r = ly; r < kTile; r += 256 / kTile) {  // src[b n + j], j fastest
            const uint64_t b = b0 + r, j = j0 + lx;
            if (b < batch && j < n) tile[r][lx] = src[b * n + j];
        }
        __syncthreads();
        for (uint32_t r = ly; r < kTile; r += 256 / kTile) {  // dst[b + j stride], b fastest
            const uint64_t b = b0 + lx, j = j0 + r;
            if (b < batch && j < n) dst[b + j * stride] = tile[lx][r];
        }
        __syncthreads();
    }
}

}  // namespace wgfft

using namespace wgfft;

// Device scratch of wgfft_exec_device_noalloc: the multipass intermediates of
// callers that cannot allocate.  Each GPU the library runs on holds one copy
// (64 MiB) until the process ends.  Host code never passes this name to a
// kernel or a copy (on the host it is only a stand-in for the symbol): it
// uses the device address hipGetSymbolAddress returns for the current device,
// looked up once per device when a handle is created there.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) float2 scratch[kNoAllocScratchBytes / sizeof(float2)];

constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static float2* g_noalloc_ptr[kMaxDevices];        // device address of `scratch` per device
static std::string g_noalloc_err[kMaxDevices];    // why the lookup failed, if it did

// With `device` current: find this device's copy of the static scratch.
static void noalloc_lookup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        g_noalloc_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_noalloc_ptr[device] = (float2*)ptr;
}

static float2* noalloc_ptr(int device, std::string* why) {
    if (device < 0 || device >= kMaxDevices) {
        *why = "device index beyond the static scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) *why = g_noalloc_err[device];
    return g_noalloc_ptr[device];
}

static bool g_device_probed = false;
static int g_device_count = 0;
static std::mutex g_probe_mu;

static int device_count() {
    std::lock_guard<std::mutex> lk(g_probe_mu);
    if (!g_device_probed) {
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) {
            (void)hipGetLastError();
            c = 0;
        }
        g_device_count = c;
        g_device_probed = true;
    }
    return g_device_count;
}

// Device entry points run on the caller's stream; NULL is HIP's default
// (null) stream, as for any HIP library call.
static hipStream_t pick_stream(wgfft_t*, void* stream) { return (hipStream_t)stream; }

static int check_dir(int dir) {
    if (dir != WGFFT_FORWARD && dir != WGFFT_INVERSE)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "direction must be WGFFT_FORWARD (-1) or WGFFT_INVERSE (+1)");
    return WGFFT_OK;
}

static int check_len(uint64_t n) {
    if (!is_pow2(n)) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    return WGFFT_OK;
}

static int check_aligned(const void* p, const char* what) {
    if (((uintptr_t)p & 15u) != 0)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, std::string(what) + " must be 16-byte aligned");
    return WGFFT_OK;
}

static int check_size(uint64_t n, uint64_t batch) {
    if (batch && n > (~0ull) / sizeof(float2) / batch)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "n * batch overflows the address space");
    return WGFFT_OK;
}

extern "C" {

int wgfft_device_available(void) { return device_count() > 0 ? 1 : 0; }

const char* wgfft_last_error(void) { return g_last_error.c_str(); }

const char* wgfft_name(void) { return "MI355X Stockham (LDS-resident radix-16/8/4/2)"; }

void wgfft_set_alloc_cap(uint64_t bytes) { g_alloc_cap.store(bytes); }

int wgfft_validate_lengths(const uint64_t* lens, uint64_t count, uint64_t* n_out) {
    if (n_out) *n_out = 0;
    if (count == 0) return WGFFT_OK;
    if (!lens) return fail(WGFFT_ERR_INVALID_ARGUMENT, "lens is NULL");
    uint64_t n = lens[0];
    for (uint64_t i = 0; i < count; ++i) {
        if (lens[i] != n) return fail(WGFFT_ERR_RAGGED_BATCH, kMsgRagged);
        if (!is_pow2(lens[i])) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    }
    if (n_out) *n_out = n;
    return WGFFT_OK;
}

int wgfft_create(wgfft_t** out, int device) {
    if (!out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int c = device_count();
    if (c <= 0) return fail(WGFFT_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= c) return fail(WGFFT_ERR_NO_DEVICE, "device index out of range");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    wgfft_t* h = new (std::nothrow) wgfft_t();
    if (!h) return fail(WGFFT_ERR_OUT_OF_MEMORY, "handle allocation failed");
    h->device = device;
    h->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    noalloc_lookup(device);  // with `device` current
    if (const char* env = std::getenv("WGFFT_MP_CHUNK_BYTES")) h->mp_chunk_bytes = std::strtoull(env, nullptr, 10);
    if (const char* env = std::getenv("WGFFT_HOST_PIPE_MIN_BYTES")) h->host_pipe_min_bytes = std::strtoull(env, nullptr, 10);
    if (const char* env = std::getenv("WGFFT_HOST_CHUNK_MIN_BYTES"))
        h->host_chunk_min_bytes = std::max<uint64_t>(1ull << 16, std::strtoull(env, nullptr, 10));
    if (const char* env = std::getenv("WGFFT_HOST_CHUNK_BYTES")) {
        h->host_chunk_bytes = std::max<uint64_t>(1ull << 20, std::strtoull(env, nullptr, 10));
        h->host_chunk_fixed = true;
    }
    load_tune_file(h);
    *out = h;
    return WGFFT_OK;
}

void wgfft_destroy(wgfft_t* h) {
    if (!h) return;
    {
        std::lock_guard<std::mutex> lk(h->mu);
        hipSetDevice(h->device);
        hipStreamSynchronize(h->stream);
        ragged_destroy_all(h);
        for (auto& t : h->tables) hipFree(t.p);
        h->tables.clear();
        h->free_pipe();
        for (DevBuf* b : {&h->io_in, &h->io_out, &h->io_w, &h->io_aux, &h->scratch, &h->scratch2, &h->red, &h->red2, &h->lossbuf})
            h->release(*b);
        if (h->scratch_ev) hipEventDestroy(h->scratch_ev);
        hipStreamDestroy(h->stream);
    }
    delete h;
}

int wgfft_exec_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir,
                      void* stream) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_dir(dir)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!d_in || !d_out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((rc = check_aligned(d_in, "d_in")) || (rc = check_aligned(d_out, "d_out"))) return rc;
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    bool inv = dir == WGFFT_INVERSE;
    return h->run(p, inv ? kInverse : kForward, (const float2*)d_in, (float2*)d_out, nullptr, batch,
                  inv ? 1.0f / (float)n : 1.0f, pick_stream(h, stream));
}

int wgfft_exec_device_noalloc(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir,
                              void* stream) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_dir(dir)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!d_in || !d_out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((rc = check_aligned(d_in, "d_in")) || (rc = check_aligned(d_out, "d_out"))) return rc;
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::string why;
    float2* sp = noalloc_ptr(h->device, &why);
    if (!sp) return fail(WGFFT_ERR_HIP, "the static scratch is not available on this device: " + why);
    // a plan allocates its twiddle tables, so it must exist already
    auto it = h->plans.find(n);
    if (it == h->plans.end())
        return fail(WGFFT_ERR_INVALID_ARGUMENT,
                    "no plan for this length on the handle: make it first with a call that may allocate "
                    "(wgfft_plan_describe, or any transform of this length)");
    Plan* p = &it->second;
    if (p->kind == Plan::kMultiPass) {
        const uint64_t need =
            p->n * sizeof(float2) * (multipass_needs_two_buffers((int)p->pass_var.size()) ? 2 : 1);
        if (need > kNoAllocScratchBytes)
            return fail(WGFFT_ERR_INVALID_ARGUMENT,
                        "length " + std::to_string(n) + " needs " + std::to_string(need) +
                            " bytes of intermediates per transform, more than the " +
                            std::to_string(kNoAllocScratchBytes) +
                            "-byte static scratch (at most 2^23 for a two-pass plan, 2^22 for three passes)");
    }
    const bool inv = dir == WGFFT_INVERSE;
    return h->run(p, inv ? kInverse : kForward, (const float2*)d_in, (float2*)d_out, nullptr, batch,
                  inv ? 1.0f / (float)n : 1.0f, pick_stream(h, stream), nullptr, sp, kNoAllocScratchBytes);
}

int wgfft_exec_strided_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch,
                              uint64_t istride, uint64_t idist, uint64_t ostride, uint64_t odist, int dir,
                              void* stream) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_dir(dir)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!d_in || !d_out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if (istride == 0 || ostride == 0 || (batch > 1 && (idist == 0 || odist == 0)))
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "strides and distances must be positive");
    if (((uintptr_t)d_in | (uintptr_t)d_out) & 7u)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffers must be 8-byte aligned");
    if (istride == 1 && idist == n && ostride == 1 && odist == n && ((((uintptr_t)d_in | (uintptr_t)d_out) & 15u) == 0))
        return wgfft_exec_device(h, d_in, d_out, n, batch, dir, stream);
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    hipStream_t s = pick_stream(h, stream);
    const bool inv = dir == WGFFT_INVERSE;
    // Padded rows (unit element stride, distances of at least n): the
    // single-kernel lengths read and write them in place of the contiguous
    // rows, one pass over memory.  Needs 16-byte aligned rows (even distances)
    // and either disjoint input and output ranges or the same rows in place.
    const bool rows_ok = batch == 1 || (idist >= n && odist >= n && ((idist | odist) & 1) == 0);
    if (istride == 1 && ostride == 1 && p->kind != Plan::kMultiPass && rows_ok &&
        (((uintptr_t)d_in | (uintptr_t)d_out) & 15u) == 0) {
        const uint64_t in_end = (uint64_t)(uintptr_t)d_in + ((batch - 1) * idist + n) * sizeof(float2);
        const uint64_t out_end = (uint64_t)(uintptr_t)d_out + ((batch - 1) * odist + n) * sizeof(float2);
        const bool disjoint = in_end <= (uint64_t)(uintptr_t)d_out || out_end <= (uint64_t)(uintptr_t)d_in;
        const bool same = (const void*)d_in == (const void*)d_out && idist == odist;
        if (disjoint || same)
            return h->run(p, inv ? kInverse : kForward, (const float2*)d_in, (float2*)d_out, nullptr, batch,
                          inv ? 1.0f / (float)n : 1.0f, s, nullptr, nullptr, 0, batch == 1 ? n : idist,
                          batch == 1 ? n : odist);
    }
    const uint64_t total = n * batch;
    if ((rc = h->ensure(h->io_aux, total * sizeof(float2)))) return rc;
    ScratchUse su(h, s);
    if (su.e != hipSuccess) return hip_fail(su.e, "stream ordering of the scratch buffers");
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
    // the adjacent-signal kernels visit 32 x 32 tiles grid-stride
    const uint64_t tiles = ((batch + kTile - 1) / kTile) * ((n + kTile - 1) / kTile);
    const unsigned tblocks = (unsigned)std::min<uint64_t>(tiles, 8192);
    float2* aux = (float2*)h->io_aux.p;
    const uint32_t log2n = (uint32_t)log2u(n);
    if (idist == 1 && batch > 1)
        hipLaunchKernelGGL(k_gather_adjacent, dim3(tblocks), dim3(256), 0, s, (const float2*)d_in, aux, n, batch,
                           istride);
    else
        hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, s, (const float2*)d_in, aux, log2n, total, istride,
                           idist);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "gather launch");
    if ((rc = h->run(p, inv ? kInverse : kForward, aux, aux, nullptr, batch, inv ? 1.0f / (float)n : 1.0f, s)))
        return rc;
    if (odist == 1 && batch > 1)
        hipLaunchKernelGGL(k_scatter_adjacent, dim3(tblocks), dim3(256), 0, s, (const float2*)aux, (float2*)d_out, n,
                           batch, ostride);
    else
        hipLaunchKernelGGL(k_scatter, dim3(blocks), dim3(256), 0, s, (const float2*)aux, (float2*)d_out, log2n, total,
                           ostride, odist);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "scatter launch");
    return WGFFT_OK;
}

int wgfft_exec_host(wgfft_t* h, const float* in, float* out, uint64_t n, uint64_t batch, int dir) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_dir(dir)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!in || !out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "host buffer is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = h->scratch_quiesce()) != hipSuccess) return hip_fail(e, "stream ordering of the scratch buffers");
    const uint64_t bytes = n * batch * sizeof(float2);
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    bool inv = dir == WGFFT_INVERSE;
    if (h->host_staged(n)) {
        rc = h->run_host(in, out, n, batch, [&](const float2* d_in, float2* d_out, uint64_t b) {
            return h->run(p, inv ? kInverse : kForward, d_in, d_out, nullptr, b, inv ? 1.0f / (float)n : 1.0f,
                          h->stream);
        });
        if (rc == WGFFT_OK) h->scratch_settled(h->stream);
        return rc;
    }
    if ((rc = h->ensure(h->io_in, bytes)) || (rc = h->ensure(h->io_out, bytes))) return rc;
    if ((e = hipMemcpyAsync(h->io_in.p, in, bytes, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(H2D)");
    if ((rc = h->run(p, inv ? kInverse : kForward, (const float2*)h->io_in.p, (float2*)h->io_out.p, nullptr, batch,
                     inv ? 1.0f / (float)n : 1.0f, h->stream)))
        return rc;
    if ((e = hipMemcpyAsync(out, h->io_out.p, bytes, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(D2H)");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    h->scratch_settled(h->stream);
    return WGFFT_OK;
}

int wgfft_exec_host_rows(wgfft_t* h, const float* const* in_rows, float* const* out_rows, uint64_t n, uint64_t batch,
                         int dir) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_dir(dir)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!in_rows || !out_rows) return fail(WGFFT_ERR_INVALID_ARGUMENT, "row pointer array is NULL");
    for (uint64_t b = 0; b < batch; ++b)
        if (!in_rows[b] || !out_rows[b]) return fail(WGFFT_ERR_INVALID_ARGUMENT, "host buffer is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = h->scratch_quiesce()) != hipSuccess) return hip_fail(e, "stream ordering of the scratch buffers");
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    const bool inv = dir == WGFFT_INVERSE;
    const float scale = inv ? 1.0f / (float)n : 1.0f;
    const uint64_t tb = n * sizeof(float2);
    if (h->host_staged(n)) {  // the staging copies gather and scatter the rows
        rc = h->run_host(nullptr, nullptr, n, batch,
                         [&](const float2* d_in, float2* d_out, uint64_t b) {
                             return h->run(p, inv ? kInverse : kForward, d_in, d_out, nullptr, b, scale, h->stream);
                         },
                         in_rows, out_rows);
        if (rc == WGFFT_OK) h->scratch_settled(h->stream);
        return rc;
    }
    // transforms larger than a staging chunk: one at a time, each row copied directly
    if ((rc = h->ensure(h->io_in, tb)) || (rc = h->ensure(h->io_out, tb))) return rc;
    for (uint64_t b = 0; b < batch; ++b) {
        if ((e = hipMemcpyAsync(h->io_in.p, in_rows[b], tb, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
            return hip_fail(e, "hipMemcpyAsync(H2D)");
        if ((rc = h->run(p, inv ? kInverse : kForward, (const float2*)h->io_in.p, (float2*)h->io_out.p, nullptr, 1,
                         scale, h->stream)))
            return rc;
        if ((e = hipMemcpyAsync(out_rows[b], h->io_out.p, tb, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
            return hip_fail(e, "hipMemcpyAsync(D2H)");
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    h->scratch_settled(h->stream);
    return WGFFT_OK;
}

int wgfft_exec_windowed_device(wgfft_t* h, const float* d_x, const float* d_w, float* d_y, uint64_t n,
                               uint64_t batch, void* stream) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!d_x || !d_w || !d_y) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((rc = check_aligned(d_x, "d_x")) || (rc = check_aligned(d_y, "d_y")) || (rc = check_aligned(d_w, "d_w")))
        return rc;
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    return h->run(p, kWindowed, (const float2*)d_x, (float2*)d_y, d_w, batch, 1.0f, pick_stream(h, stream));
}

int wgfft_exec_windowed_host(wgfft_t* h, const float* x, const float* w, float* y, uint64_t n, uint64_t batch) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    int rc;
    if ((rc = check_len(n)) || (rc = check_size(n, batch))) return rc;
    if (batch == 0) return WGFFT_OK;
    if (!x || !w || !y) return fail(WGFFT_ERR_INVALID_ARGUMENT, "host buffer is NULL");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    if ((e = h->scratch_quiesce()) != hipSuccess) return hip_fail(e, "stream ordering of the scratch buffers");
    const uint64_t bytes = n * batch * sizeof(float2);
    if ((rc = h->ensure(h->io_w, n * sizeof(float) + 16))) return rc;
    Plan* p;
    if ((rc = h->get_plan(n, &p))) return rc;
    if ((e = hipMemcpyAsync(h->io_w.p, w, n * sizeof(float), hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(H2D)");
    if (h->host_staged(n)) {
        rc = h->run_host(x, y, n, batch, [&](const float2* d_in, float2* d_out, uint64_t b) {
            return h->run(p, kWindowed, d_in, d_out, (const float*)h->io_w.p, b, 1.0f, h->stream);
        });
        if (rc == WGFFT_OK) h->scratch_settled(h->stream);
        return rc;
    }
    if ((rc = h->ensure(h->io_in, bytes)) || (rc = h->ensure(h->io_out, bytes))) return rc;
    if ((e = hipMemcpyAsync(h->io_in.p, x, bytes, hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(H2D)");
    if ((rc = h->run(p, kWindowed, (const float2*)h->io_in.p, (float2*)h->io_out.p, (const float*)h->io_w.p, batch,
                     1.0f, h->stream)))
        return rc;
    if ((e = hipMemcpyAsync(y, h->io_out.p, bytes, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(D2H)");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    h->scratch_settled(h->stream);
    return WGFFT_OK;
}

int wgfft_bench_device(wgfft_t* h, const float* d_in, float* d_out, uint64_t n, uint64_t batch, int dir, int warmup,
                       int iters, double* sec_per_iter) {
    if (!h || !sec_per_iter) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle or output is NULL");
    if (iters < 1) return fail(WGFFT_ERR_INVALID_ARGUMENT, "iters must be >= 1");
    int rc;
    hipError_t e;
    {
        // host-blocking: its calls on the handle's stream are ordered after
        // any pending use of the scratch on another stream here
        std::lock_guard<std::mutex> lk(h->mu);
        if ((e = hipSetDevice(h->device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
        if ((e = h->scratch_quiesce()) != hipSuccess) return hip_fail(e, "stream ordering of the scratch buffers");
    }
    for (int i = 0; i < warmup; ++i) {
        if ((rc = wgfft_exec_device(h, d_in, d_out, n, batch, dir, (void*)h->stream))) return rc;
        if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    }
    hipEvent_t a, b;
    if ((e = hipEventCreate(&a)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        hipEventDestroy(a);
        return hip_fail(e, "hipEventCreate");
    }
    rc = WGFFT_OK;
    hipEventRecord(a, h->stream);
    for (int i = 0; i < iters && rc == WGFFT_OK; ++i)
        rc = wgfft_exec_device(h, d_in, d_out, n, batch, dir, (void*)h->stream);
    hipEventRecord(b, h->stream);
    if (rc == WGFFT_OK) {
        if ((e = hipEventSynchronize(b)) != hipSuccess) {
            rc = hip_fail(e, "hipEventSynchronize");
        } else {
            float ms = 0.f;
            hipEventElapsedTime(&ms, a, b);
            *sec_per_iter = (double)ms * 1e-3 / iters;
            std::lock_guard<std::mutex> lk(h->mu);
            h->scratch_settled(h->stream);
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc;
}

void wgfft_synth_uniform(float* out, uint64_t nfloats, uint64_t seed) { synth_uniform(out, nfloats, seed); }

}  // extern "C"
