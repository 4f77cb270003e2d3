// tune.hip -- kernel-variant autotuner, plan description, tune-file
// persistence and host input synthesis.
//
// The autotuner times every compiled LDS variant for a length (radix split,
// elements per thread, transforms per workgroup) on a resident scratch batch
// and keeps the fastest for the handle.  With WGFFT_TUNE_FILE set, choices
// are read at handle creation and appended when a tuning run finishes
// (lines: "log2n plan E G flags").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/wgfft.h"
#include "kernels_adjoint.hpp"
#include "runtime_internal.hpp"

namespace wgfft {

static uint64_t splitmix64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Pcg32 {
    uint64_t state = 0, inc = 1;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
};

void synth_uniform(float* out, uint64_t nfloats, uint64_t seed) {
    uint64_t sm = seed;
    uint64_t st = splitmix64_next(&sm);
    uint64_t sq = splitmix64_next(&sm);
    Pcg32 r;
    r.seed(st, sq);
    for (uint64_t i = 0; i < nfloats; ++i) out[i] = (float)((double)(r.next() >> 8) * (1.0 / 8388608.0) - 1.0);
}

// Uniform [-1, 1) noise for the tuning batch (an integer hash of the index;
// zero-filled inputs draw less power and time optimistically).
__global__ void k_fill_noise(float* __restrict__ x, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = (uint32_t)i * 0x9E3779B1u ^ (uint32_t)(i >> 32);
        v ^= v >> 16;
        v *= 0x85EBCA6Bu;
        v ^= v >> 13;
        v *= 0xC2B2AE35u;
        v ^= v >> 16;
        x[i] = (float)(v >> 8) * (1.0f / 8388608.0f) - 1.0f;
    }
}

static int fill_noise(Handle* h, void* p, uint64_t bytes) {
    const uint64_t n = bytes / sizeof(float);
    hipLaunchKernelGGL(k_fill_noise, dim3(1024), dim3(256), 0, h->stream, (float*)p, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? WGFFT_OK : hip_fail(e, "noise fill");
}

// Timer for candidate launches on the handle's stream.  The shader clock
// ramps up over tens of milliseconds of continuous load, so settle() runs a
// launch back to back before a sweep, and each candidate is timed as the mean
// of three bursts of at least ~2 ms.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        hipEventCreate(&a);
        hipEventCreate(&b);
    }
    ~Timer() {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
    // launch() returns a WGFFT status; a burst stops at the first failure
    // and returns it, so a candidate that did not run is never timed
    template <class F>
    int burst(F&& launch, int iters, float* ms) {
        hipEventRecord(a, s);
        int rc = WGFFT_OK;
        for (int i = 0; i < iters && rc == WGFFT_OK; ++i) rc = launch();
        hipEventRecord(b, s);
        const hipError_t e = hipEventSynchronize(b);
        if (rc != WGFFT_OK) return rc;
        if (e != hipSuccess) return hip_fail(e, "autotune timing");
        hipEventElapsedTime(ms, a, b);
        return WGFFT_OK;
    }
    template <class F>
    int settle(F&& launch, float target_ms = 40.f) {
        float ms = 0.f, total = 0.f;
        int iters = 4, rc;
        while (total < target_ms) {
            if ((rc = burst(launch, iters, &ms))) return rc;
            total += ms;
            if (ms < 2.f && iters < 4096) iters *= 2;
        }
        return WGFFT_OK;
    }
    // per-launch time in ms: the mean over three bursts, the steady rate a
    // caller sees (the best burst favoured a plan 5% slower on average)
    template <class F>
    int time(F&& launch, float* per_launch) {
        float ms = 0.f;
        int rc;
        if ((rc = burst(launch, 2, &ms))) return rc;
        const int iters = std::max(5, std::min(500, (int)(2.f / std::max(ms * 0.5f, 1e-3f))));
        float total = 0.f;
        for (int round = 0; round < 3; ++round) {
            if ((rc = burst(launch, iters, &ms))) return rc;
            total += ms;
        }
        *per_launch = total / (3.f * iters);
        return WGFFT_OK;
    }
};

void load_tune_file(Handle* h) {
    const char* path = std::getenv("WGFFT_TUNE_FILE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    int log2n, E, G, flags;
    char plan[64];
    while (std::fscanf(f, "%d %63s %d %d %d", &log2n, plan, &E, &G, &flags) == 5) {
        const BatchedVariant* v = find_variant(log2n, plan, E, G, flags);
        if (v) h->tuned[log2n] = v;
    }
    std::fclose(f);
}

static void describe(const BatchedVariant* v, char* desc, uint64_t len) {
    if (!desc || !len) return;
    char waves[24] = "";
    if (flag_min_waves(v->flags)) std::snprintf(waves, sizeof waves, " waves>=%d", flag_min_waves(v->flags));
    std::snprintf(desc, len, "lds N=%d plan=%s E=%d G=%d%s%s%s%s%s%s threads=%d lds=%dB", 1 << v->log2n, v->plan, v->E,
                  v->G, (v->flags & kFlagNtStore) ? " nt-store" : "", (v->flags & kFlagPrio) ? " prio" : "",
                  (v->flags & kFlagTwRegs) ? " tw-regs" : (v->flags & kFlagTwLds) ? " tw-lds" : "",
                  (v->flags & kFlagNoPrefetch)     ? " no-prefetch"
                  : (v->flags & kFlagHalfPrefetch) ? ((v->flags & kFlagQuarterLoad) ? " 3/4-prefetch" : " half-prefetch")
                  : (v->flags & kFlagSplit2)       ? " split2"
                                                   : "",
                  (v->flags & kFlagCachedLoad) ? " cached-loads" : "", waves,
                  v->threads, v->lds_bytes);
}

static std::string describe_multipass(const Plan* p) {
    std::string s = "multipass N=" + std::to_string(p->n) + " passes=";
    for (size_t i = 0; i < p->pass_var.size(); ++i) {
        const GlobalPassVariant* v = p->pass_var[i];
        s += (i ? "," : "") + std::string("R") + std::to_string(1 << v->log2r) + "(" + v->plan + ",G" +
             std::to_string(v->G) + (v->prefetch == 3 ? ",direct"
                                      : v->prefetch == 4 ? ",direct,prefetch"
                                      : v->prefetch == 5 ? ",direct,halves"
                                      : v->prefetch == 6 ? ",direct,quarters"
                                      : v->prefetch      ? ",prefetch"
                                                         : "") + ")";
    }
    return s;
}

// Time every radix split the planner offers (multipass_splits) and, for
// each, every combination of global-pass variants (a few per pass) on a
// scratch batch; keep the fastest split and variants.  A combination whose
// launch fails is skipped; when none runs the plan keeps its previous split
// and the first failure is returned.
static int tune_multipass(Handle* h, Plan* p, uint64_t batch, char* desc, uint64_t desc_len) {
    const uint64_t bytes = p->n * batch * sizeof(float2);
    int rc;
    if ((rc = h->ensure(h->io_in, bytes)) || (rc = h->ensure(h->io_out, bytes))) return rc;
    if ((rc = fill_noise(h, h->io_in.p, bytes))) return rc;
    Timer tm(h->stream);
    auto launch = [&] {
        return h->run(p, kForward, (const float2*)h->io_in.p, (float2*)h->io_out.p, nullptr, batch, 1.0f, h->stream);
    };
    const std::vector<int> prev_split = p->pass_log2r;
    const std::vector<const GlobalPassVariant*> prev_vars = p->pass_var;
    if ((rc = tm.settle(launch))) return rc;
    std::vector<int> best_split = p->pass_log2r;
    std::vector<const GlobalPassVariant*> best_vars = p->pass_var;
    float best_ms = 1e30f;
    int first_err = WGFFT_OK;
    std::string first_msg;
    for (const std::vector<int>& split : multipass_splits(p->log2n)) {
        if ((rc = h->set_split(p, split))) return rc;
        std::vector<std::vector<const GlobalPassVariant*>> cands;
        uint64_t L = 1;
        for (size_t i = 0; i < split.size(); ++i) {
            cands.push_back(global_pass_candidates(split[i], L, p->n >> split[i]));
            L <<= split[i];
        }
        std::vector<size_t> idx(cands.size(), 0);
        for (;;) {
            for (size_t i = 0; i < cands.size(); ++i) p->pass_var[i] = cands[i][idx[i]];
            float ms;
            if ((rc = tm.time(launch, &ms))) {
                if (first_err == WGFFT_OK) {
                    first_err = rc;
                    first_msg = g_last_error;
                }
            } else if (ms < best_ms) {
                best_ms = ms;
                best_split = split;
                best_vars = p->pass_var;
            }
            size_t k = 0;
            while (k < idx.size() && ++idx[k] == cands[k].size()) idx[k++] = 0;
            if (k == idx.size()) break;
        }
    }
    if (best_ms >= 1e30f) {
        if ((rc = h->set_split(p, prev_split))) return rc;
        p->pass_var = prev_vars;
        return fail(first_err, "autotune: no candidate ran (" + first_msg + ")");
    }
    if ((rc = h->set_split(p, best_split))) return rc;
    p->pass_var = best_vars;
    if (desc && desc_len) std::snprintf(desc, desc_len, "%s", describe_multipass(p).c_str());
    return WGFFT_OK;
}

}  // namespace wgfft

using namespace wgfft;

extern "C" int wgfft_plan_describe(wgfft_t* h, uint64_t n, char* desc, uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    std::lock_guard<std::mutex> lk(h->mu);
    hipSetDevice(h->device);
    Plan* p;
    int rc;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (!desc || !desc_len) return WGFFT_OK;
    if (p->kind == Plan::kSmall) {
        std::snprintf(desc, desc_len, "register N=%llu one thread per transform", (unsigned long long)n);
    } else if (p->kind == Plan::kBatched) {
        describe(p->var, desc, desc_len);
        if (p->var_small) {
            char small[160];
            describe(p->var_small, small, sizeof small);
            const std::string full = std::string(desc) + "; batches <= " +
                                     std::to_string(kSmallBatchPerCu * (uint64_t)h->num_cus) + ": " + small;
            std::snprintf(desc, desc_len, "%s", full.c_str());
        }
        if (p->var_win) {
            char win[160];
            describe(p->var_win, win, sizeof win);
            const std::string full = std::string(desc) + "; windowed: " + win;
            std::snprintf(desc, desc_len, "%s", full.c_str());
        }
    } else {
        std::snprintf(desc, desc_len, "%s", describe_multipass(p).c_str());
    }
    return WGFFT_OK;
}

// The sweeps of wgfft_autotune, with h->mu held on the handle's device.
static int autotune_locked(Handle* h, uint64_t n, uint64_t batch, char* desc, uint64_t desc_len) {
    Plan* p;
    int rc;
    if ((rc = h->get_plan(n, &p))) return rc;
    const uint64_t bytes = n * batch * sizeof(float2);
    // the sweeps write the handle's scratch buffers: ordered after any device
    // call still using them on another stream
    ScratchUse su(h, h->stream);
    if (su.e != hipSuccess) return hip_fail(su.e, "stream ordering of the scratch buffers");
    if (p->kind == Plan::kMultiPass) return tune_multipass(h, p, batch, desc, desc_len);
    if (p->kind != Plan::kBatched) {
        if (desc && desc_len) std::snprintf(desc, desc_len, "no alternatives for N=%llu", (unsigned long long)n);
        return WGFFT_OK;
    }
    if ((rc = h->ensure(h->io_in, bytes)) || (rc = h->ensure(h->io_out, bytes))) return rc;
    if ((rc = fill_noise(h, h->io_in.p, bytes))) return rc;
    // a window for the windowed entry's sweep (also the fused kernels' w and m below)
    if ((rc = h->ensure(h->io_w, 2 * n * sizeof(float) + 64))) return rc;
    if ((rc = fill_noise(h, h->io_w.p, 2 * n * sizeof(float) + 64))) return rc;
    Timer tm(h->stream);
    const BatchedVariant* best = nullptr;
    float best_ms = 1e30f;
    const BatchedVariant* cur = p->var;
    auto launch = [&] {
        return h->run(p, kForward, (const float2*)h->io_in.p, (float2*)h->io_out.p, nullptr, batch, 1.0f, h->stream,
                      cur);
    };
    auto launch_win = [&] {
        return h->run(p, kWindowed, (const float2*)h->io_in.p, (float2*)h->io_out.p, (const float*)h->io_w.p, batch,
                      1.0f, h->stream, cur);
    };
    if ((rc = tm.settle(launch))) return rc;
    int first_err = WGFFT_OK;
    std::string first_msg;
    std::vector<const BatchedVariant*> cands;
    for (const auto& v : batched_registry())
        if (v.log2n == p->log2n) cands.push_back(&v);
    // two sweeps in opposite orders, each candidate keeping its faster mean
    // (clock drift along one sweep favours whatever it times last; with
    // best-burst scores the sweep kept a plan 5-9% slower than the default
    // at 8192 x 1024).  The windowed entry is timed beside the plain one and
    // keeps its own fastest: its window loads change the register budget, and
    // the plan fastest for one mode is not always so for the other (8192 x
    // 1024: 2x8x8x8x8 forward 22.6 us, windowed 25.2 us; 2x16x16x16 23.7 and
    // 22.9 us).
    std::vector<float> cand_ms(cands.size(), 1e30f), cand_win(cands.size(), 1e30f);
    for (int sweep = 0; sweep < 2; ++sweep)
        for (size_t i = 0; i < cands.size(); ++i) {
            const size_t j = sweep ? cands.size() - 1 - i : i;
            cur = cands[j];
            float ms;
            if ((rc = tm.time(launch, &ms))) {  // a candidate that fails to launch is never chosen
                if (first_err == WGFFT_OK) {
                    first_err = rc;
                    first_msg = g_last_error;
                }
                continue;
            }
            cand_ms[j] = std::min(cand_ms[j], ms);
            if (tm.time(launch_win, &ms) == WGFFT_OK) cand_win[j] = std::min(cand_win[j], ms);
        }
    const BatchedVariant* best_win = nullptr;
    float best_win_ms = 1e30f;
    for (size_t j = 0; j < cands.size(); ++j) {
        if (cand_ms[j] < best_ms) {
            best_ms = cand_ms[j];
            best = cands[j];
        }
        if (cand_win[j] < best_win_ms) {
            best_win_ms = cand_win[j];
            best_win = cands[j];
        }
    }
    if (!best) return fail(first_err ? first_err : WGFFT_ERR_INVALID_LENGTH, "autotune: no variant ran (" + first_msg + ")");
    // fused kernels for this length, timed per role: the fused forward +
    // adjoint, the VJP and the spectral filter each keep their own fastest
    std::string adj_desc;
    const auto adjs = adjoint_variants(p->log2n);
    if (adjs.size() > 1) {
        if ((rc = h->ensure(h->io_aux, bytes))) return rc;
        float* w = (float*)h->io_w.p;
        float* m = w + n;
        float* loss = m + n;
        float* gw = (float*)h->io_out.p;  // scratch: gx goes to io_aux
        const AdjointVariant* av = adjs[0];
        auto fused_launch = [&] {
            return fused_adjoint(h, av, true, (const float*)h->io_in.p, w, m, nullptr, nullptr, (float*)h->io_aux.p, gw,
                                 loss, n, batch, p->tw, h->stream);
        };
        auto vjp_launch = [&] {  // gY: the noise batch itself
            return fused_adjoint(h, av, false, (const float*)h->io_in.p, w, nullptr, (const float*)h->io_in.p, nullptr,
                                 (float*)h->io_aux.p, gw, nullptr, n, batch, p->tw, h->stream);
        };
        auto filter_launch = [&] {
            int grid = h->adjoint_blocks_per_cu(av, 2) * h->num_cus;
            if ((uint64_t)grid > batch) grid = (int)batch;
            FilterArgs fa{(const float2*)h->io_in.p, w, (float2*)h->io_aux.p, p->tw, batch, grid, h->stream, 0u, 0.f};
            const hipError_t e = av->filter(fa);
            return e == hipSuccess ? WGFFT_OK : hip_fail(e, "filter kernel launch");
        };
        const int roles[3] = {kRoleFused, kRoleVjp, kRoleFilter};
        const AdjointVariant* best_role[3] = {nullptr, nullptr, nullptr};
        for (int r = 0; r < 3; ++r) {
            // two interleaved sweeps, each variant keeping its faster time: one
            // sweep's order and clock drift picked a variant 8% slower than the
            // default now and then; a variant that fails to launch keeps 1e30
            std::vector<float> adj_ms(adjs.size(), 1e30f);
            for (int sweep = 0; sweep < 2; ++sweep)
                for (size_t i = 0; i < adjs.size(); ++i) {
                    const size_t j = sweep ? adjs.size() - 1 - i : i;
                    av = adjs[j];
                    float ms;
                    const int trc = roles[r] == kRoleFused ? tm.time(fused_launch, &ms)
                                    : roles[r] == kRoleVjp ? tm.time(vjp_launch, &ms)
                                                           : tm.time(filter_launch, &ms);
                    if (trc != WGFFT_OK) continue;
                    adj_ms[j] = std::min(adj_ms[j], ms);
                }
            float best_ms_r = 1e30f;
            for (size_t j = 0; j < adjs.size(); ++j)
                if (adj_ms[j] < best_ms_r) {
                    best_ms_r = adj_ms[j];
                    best_role[r] = adjs[j];
                }
        }
        char buf[128];
        if (best_role[0]) {
            h->tuned_adj[p->log2n] = best_role[0];
            adj_desc = std::string("; ") + adjoint_describe(best_role[0], buf, sizeof buf);
        }
        if (best_role[1]) {
            h->tuned_vjp[p->log2n] = best_role[1];
            adj_desc += std::string("; vjp ") + best_role[1]->plan + ((best_role[1]->opt & kAdjSplit2) ? " split2" : "");
        }
        if (best_role[2]) {
            h->tuned_filter[p->log2n] = best_role[2];
            adj_desc += std::string("; filter ") + best_role[2]->plan + ((best_role[2]->opt & kAdjSplit2) ? " split2" : "");
        }
    }
    if (best) {
        h->tuned[p->log2n] = best;
        p->var = best;
        p->var_small = nullptr;
        p->var_win = best_win && best_win != best ? best_win : nullptr;
        if (p->var_win)
            h->tuned_win[p->log2n] = p->var_win;
        else
            h->tuned_win.erase(p->log2n);
        describe(best, desc, desc_len);
        if (desc && desc_len) {
            std::string full = desc;
            if (p->var_win) {
                char wbuf[160];
                describe(p->var_win, wbuf, sizeof wbuf);
                full += std::string("; windowed: ") + wbuf;
            }
            full += adj_desc;
            std::snprintf(desc, desc_len, "%s", full.c_str());
        }
        if (const char* path = std::getenv("WGFFT_TUNE_FILE")) {
            if (FILE* f = std::fopen(path, "a")) {
                std::fprintf(f, "%d %s %d %d %d\n", best->log2n, best->plan, best->E, best->G, best->flags);
                std::fclose(f);
            }
        }
    }
    return WGFFT_OK;
}

extern "C" int wgfft_autotune(wgfft_t* h, uint64_t n, uint64_t batch, char* desc, uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    if (batch == 0) return fail(WGFFT_ERR_INVALID_ARGUMENT, "batch must be >= 1");
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    // host-blocking: its sweeps have finished when it returns
    if ((e = h->scratch_quiesce()) != hipSuccess) return hip_fail(e, "stream ordering of the scratch buffers");
    const int rc = autotune_locked(h, n, batch, desc, desc_len);
    if (rc != WGFFT_OK) return rc;
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    h->scratch_settled(h->stream);
    return WGFFT_OK;
}

extern "C" int wgfft_variant_count(uint64_t n) {
    if (n == 0 || (n & (n - 1))) return 0;
    int log2n = 0;
    while ((1ull << log2n) < n) ++log2n;
    int c = 0;
    for (const auto& v : batched_registry()) c += v.log2n == log2n;
    return c;
}

extern "C" int wgfft_select_variant(wgfft_t* h, uint64_t n, int index, char* desc, uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n == 0 || (n & (n - 1))) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    int rc;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind != Plan::kBatched) return fail(WGFFT_ERR_INVALID_ARGUMENT, "length has no selectable variants");
    int i = 0;
    for (const auto& v : batched_registry()) {
        if (v.log2n != p->log2n) continue;
        if (i++ == index) {
            h->tuned[p->log2n] = &v;
            h->tuned_win.erase(p->log2n);
            p->var = &v;
            p->var_small = nullptr;
            p->var_win = nullptr;
            describe(&v, desc, desc_len);
            return WGFFT_OK;
        }
    }
    return fail(WGFFT_ERR_INVALID_ARGUMENT, "variant index out of range");
}

static int log2_exact(uint64_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    int l = 0;
    while ((1ull << l) < n) ++l;
    return l;
}

extern "C" int wgfft_pass_count(uint64_t n) {
    const int l = log2_exact(n);
    if (l <= 14 || l > 39) return 0;
    return (int)multipass_radices(l).size();
}

extern "C" int wgfft_pass_variant_count(uint64_t n, int pass) {
    const int l = log2_exact(n);
    if (l <= 14 || l > 39) return 0;
    const std::vector<int> r = multipass_radices(l);
    if (pass < 0 || pass >= (int)r.size()) return 0;
    uint64_t L = 1;
    for (int i = 0; i < pass; ++i) L <<= r[i];
    return (int)global_pass_candidates(r[pass], L, n >> r[pass]).size();
}

extern "C" int wgfft_select_pass_variant(wgfft_t* h, uint64_t n, int pass, int index, char* desc,
                                         uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (log2_exact(n) < 0) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    int rc;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind != Plan::kMultiPass) return fail(WGFFT_ERR_INVALID_ARGUMENT, "length has no global passes");
    if (pass < 0 || pass >= (int)p->pass_log2r.size()) return fail(WGFFT_ERR_INVALID_ARGUMENT, "pass out of range");
    uint64_t L = 1;
    for (int i = 0; i < pass; ++i) L <<= p->pass_log2r[i];
    const auto c = global_pass_candidates(p->pass_log2r[pass], L, n >> p->pass_log2r[pass]);
    if (index < 0 || index >= (int)c.size()) return fail(WGFFT_ERR_INVALID_ARGUMENT, "variant index out of range");
    p->pass_var[pass] = c[index];
    if (desc && desc_len) std::snprintf(desc, desc_len, "%s", describe_multipass(p).c_str());
    return WGFFT_OK;
}

extern "C" int wgfft_select_split(wgfft_t* h, uint64_t n, const int* log2r, int passes, char* desc,
                                  uint64_t desc_len) {
    if (!h) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle is NULL");
    const int l = log2_exact(n);
    if (l < 0) return fail(WGFFT_ERR_INVALID_LENGTH, kMsgLength);
    if (!log2r || passes < 2) return fail(WGFFT_ERR_INVALID_ARGUMENT, "a split needs at least two radices");
    std::vector<int> r(log2r, log2r + passes);
    int sum = 0;
    uint64_t L = 1;
    for (int ri : r) {
        if (ri < 1 || ri > 13) return fail(WGFFT_ERR_INVALID_ARGUMENT, "radices must be 2^1 .. 2^13");
        sum += ri;
    }
    if (sum != l) return fail(WGFFT_ERR_INVALID_ARGUMENT, "the radices must multiply to n");
    for (int ri : r) {
        if (global_pass_candidates(ri, L, n >> ri).empty())
            return fail(WGFFT_ERR_INVALID_ARGUMENT, "no global-pass kernel for radix 2^" + std::to_string(ri) +
                                                        " at stride " + std::to_string(L));
        L <<= ri;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    Plan* p;
    int rc;
    if ((rc = h->get_plan(n, &p))) return rc;
    if (p->kind != Plan::kMultiPass) return fail(WGFFT_ERR_INVALID_ARGUMENT, "length has no global passes");
    if ((rc = h->set_split(p, r))) return rc;
    if (desc && desc_len) std::snprintf(desc, desc_len, "%s", describe_multipass(p).c_str());
    return WGFFT_OK;
}
