// ragged.hip -- mixed-length batches in one persistent launch (no
// counterpart in the reference, which needs one GpuFft call per length,
// src/lib.rs:527-531).
//
// The host sorts the transforms by length (largest first) and cuts them into
// work units: for N <= 4096 a unit is G transforms of one length that a
// 256-thread workgroup transforms together in LDS (T*G = 256, 16 elements
// per thread); N = 2^13 is one transform in LDS (32 elements per thread);
// N = 2^14 is one transform held in registers as a radix-2 stage and two
// 8192-point halves through the LDS; for N = 2^15, 2^16 a unit is one
// transform that the workgroup does as a two-pass Stockham (radix R1 then
// R2 LDS sub-FFTs) through its own slot of global scratch.  Units never
// depend on each other, so the only shared
// state is the queue head.  A persistent grid (occupancy x CUs) dequeues
// units with one atomic ticket per unit; the workgroup holding the launch's
// last ticket resets the counter, so the launch can be captured once into a
// hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>
#include <numeric>
#include <vector>

#include "../../include/wgfft.h"
#include "kernels.hpp"
#include "runtime_internal.hpp"

namespace wgfft {

struct RaggedUnit {
    uint32_t log2n;
    uint32_t first;  // index into the sorted offset table
    uint32_t count;  // transforms in the unit
    uint32_t pad;
};

constexpr int kRaggedThreads = 256;
// Inputs and outputs are streamed non-temporally (read and written once), so
// they displace less of the cached big-unit scratch.
constexpr bool NTIO = true;
constexpr int kMinLog2 = 6, kMaxLog2 = 16, kDirectMaxLog2 = 13, kSplitLog2 = 14;

// The wave priority is raised while a wave issues global loads and stores, so
// the waves of a CU keep the memory pipeline fed ahead of their LDS and ALU work.
__device__ __forceinline__ void prio_up() { __builtin_amdgcn_s_setprio(2); }
__device__ __forceinline__ void prio_down() { __builtin_amdgcn_s_setprio(0); }

// elements per thread of the direct (LDS-resident) unit for length 2^k
constexpr int direct_e(int log2n) { return log2n == 13 ? 32 : 16; }
// transforms per direct unit for length 2^k
constexpr int direct_g(int log2n) { return kRaggedThreads / ((1 << log2n) / direct_e(log2n)); }

// LDS footprint (complex elements) of every unit kind
template <int N, int E, int... Rs>
constexpr int direct_lds() { return direct_g(ilog2(N)) * LdsLayout<N, E, Rs...>::len; }
template <int R, int E, int... Rs>
constexpr int pass_lds() { return (kRaggedThreads / (R / E)) * GdSlot<R, E, Rs...>::len; }

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int kRaggedLds = cmax(
    cmax(cmax(direct_lds<64, 16, 8, 8>(), direct_lds<128, 16, 4, 4, 8>()),
         cmax(direct_lds<256, 16, 8, 4, 8>(), direct_lds<512, 16, 8, 8, 8>())),
    cmax(cmax(direct_lds<1024, 16, 8, 16, 8>(), direct_lds<2048, 16, 8, 4, 8, 8>()),
         cmax(cmax(direct_lds<4096, 16, 8, 8, 8, 8>(), direct_lds<8192, 32, 16, 16, 16, 2>()),
              cmax(cmax(pass_lds<128, 16, 4, 8, 4>(), pass_lds<256, 16, 16, 16>()),
                   LdsLayoutSingle<8192, 32, 8, 16, 8, 8>::len))));

// Unit handlers take the thread index through launder_lane() once per unit,
// so no handler's index arithmetic is hoisted out of the persistent loop and
// the kernel's register budget stays the largest handler's.
template <int N, int E, bool INV, int... Rs>
__device__ __forceinline__ void ragged_direct(const float2* __restrict__ in, float2* __restrict__ out,
                                              const uint64_t* __restrict__ offs, uint32_t first, uint32_t count,
                                              const TwLds& tl, float2* __restrict__ lds_raw, uint32_t tid) {
    constexpr int T = N / E;
    constexpr int LEN = LdsLayout<N, E, Rs...>::len;
    static_assert(T * direct_g(ilog2(N)) == kRaggedThreads, "direct unit must fill the workgroup");
    const uint32_t g = tid / T, t = tid % T;
    const bool active = g < count;
    const uint64_t off = active ? offs[first + g] : 0;
    float2 d[E];
    prio_up();
    if (active) Io<N, E, Rs...>::template load<NTIO>(d, in + off, t);
    prio_down();
    Engine<N, E, T, INV>::run_with(d, lds_raw + g * LEN, tl.shifted(kMaxLog2 - ilog2(N)), t, Radices<Rs...>{});
    prio_up();
    if (active) Io<N, E, Rs...>::template store<INV, NTIO>(out + off, d, t, 1.0f / (float)N);
    prio_down();
}

// One length-N signal (N = 2^14) in one workgroup without scratch: a
// decimation-in-frequency radix-2 stage in registers, y0[n] = x[n] + x[n + H]
// and y1[n] = (x[n] - x[n + H]) W_N^n (H = N/2), then the two H-point halves
// in turn through the LDS, and X[2k] = Y0[k], X[2k + 1] = Y1[k] stored as
// 16-byte pairs.  Thread t holds single-layout slot i RF + m, index
// t + T i + m H/RF, of both halves (RF the first radix).
template <int N, int E, bool INV, int... Rs>
__device__ __forceinline__ void ragged_split2(const float2* __restrict__ src, float2* __restrict__ dst,
                                              const TwLds& tl, float2* __restrict__ lds, uint32_t tid) {
    constexpr int H = N / 2;
    constexpr int T = H / E;
    static_assert(T == kRaggedThreads, "the halves must fill the workgroup");
    using MP = GdMap<H, E, 1, Rs...>;
    using EN = Engine<H, E, T, INV>;
    const TwLds tn = tl.shifted(kMaxLog2 - ilog2(N));  // W_N
    typedef float f2v __attribute__((ext_vector_type(2)));
    float2 y0[E], y1[E];
    prio_up();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t n = MP::row(e, tid);
        const f2v a = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(src + n));
        const f2v b = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(src + n + H));
        y0[e] = make_float2(a.x, a.y);
        y1[e] = make_float2(b.x, b.y);
    }
    prio_down();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t j = MP::row(e, tid) << tn.sh;  // W_N^n as an index of the 2^16 tables (< 2^15)
        const float2 w = cmul(tn.hi[j >> tn.h], tn.lo[j & ((1u << tn.h) - 1)]);
        const float2 a = y0[e], b = y1[e];
        y0[e] = cadd(a, b);
        y1[e] = twmul<INV>(csub(a, b), w);
    }
    const TwLds th = tl.shifted(kMaxLog2 - ilog2(H));  // W_H
    EN::run_single(y0, lds, th, tid, Radices<Rs...>{});
    WGFFT_LDS_SYNC();  // the first half's last exchange has been read
    const uint32_t t1 = launder_lane(tid);
    EN::run_single(y1, lds, th, t1, Radices<Rs...>{});
    const float sc = INV ? 1.0f / (float)N : 1.0f;
    prio_up();
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t k = MP::row(e, t1);
        const float2 a = cscale(y0[e], sc), b = cscale(y1[e], sc);
        st16<NTIO>(reinterpret_cast<float4*>(dst + 2 * k), make_float4(a.x, a.y, b.x, b.y));
    }
    prio_down();
}

// Two-pass transform of one length-N signal (N = R1 * R2) by one workgroup:
// pass 1 (L = 1) src -> scratch, pass 2 (L = R1) scratch -> dst.
// PF: prefetch the next group's tile into registers while the current one is
// transformed (more registers, more overlap).
template <int LOGN, int R1, int E1, class P1, int R2, int E2, class P2, bool INV, bool PF>
struct RaggedBig;

template <int LOGN, int R1, int E1, int... Rs1, int R2, int E2, int... Rs2, bool INV, bool PF>
struct RaggedBig<LOGN, R1, E1, Radices<Rs1...>, R2, E2, Radices<Rs2...>, INV, PF> {
    static constexpr uint64_t N = 1ull << LOGN;
    static constexpr int G1 = kRaggedThreads / (R1 / E1);
    static constexpr int G2 = kRaggedThreads / (R2 / E2);
    static_assert(R1 * R2 == (1 << LOGN), "radices must multiply to N");
    static_assert(G2 <= R1, "second pass needs G <= L");
    __device__ __forceinline__ static void run(const float2* __restrict__ src, float2* __restrict__ dst,
                                               float2* __restrict__ scratch, const float2* __restrict__ tw_all,
                                               const TwLds& tl, float2* __restrict__ lds, uint32_t tid) {
        const TwLds t1 = tl.shifted(kMaxLog2 - ilog2(R1)), t2 = tl.shifted(kMaxLog2 - ilog2(R2));
        constexpr int M = INV ? kInverse : kForward;
        const TwOneLevel twN{tw_all + N};
        using Side = GdSide<TwOneLevel, E1, M>;
        static_assert(E1 == E2, "both passes share the prefetch registers");
        // pass 1 (L = 1): src -> scratch (no twiddles, so no side loads)
        float2 cur[E1];
        Side sd;
        prio_up();
        if (PF) gd_load<R1, E1, G1, NTIO, Rs1...>(cur, src, N / R1, 0, tid);
        prio_down();
#pragma unroll 1
        for (uint64_t b0 = 0; b0 < N / R1; b0 += G1) {
            if constexpr (PF) {
                float2 nxt[E1];
                prio_up();
                if (b0 + G1 < N / R1) gd_load<R1, E1, G1, NTIO, Rs1...>(nxt, src, N / R1, b0 + G1, tid);
                prio_down();
                gd_compute<R1, E1, G1, M, TwOneLevel, TwLds, false, Rs1...>(cur, sd, scratch, t1, N, 1, b0, 1.0f, lds,
                                                                            tid);
#pragma unroll
                for (int i = 0; i < E1; ++i) cur[i] = nxt[i];
            } else {
                prio_up();
                gd_load<R1, E1, G1, NTIO, Rs1...>(cur, src, N / R1, b0, tid);
                prio_down();
                gd_compute<R1, E1, G1, M, TwOneLevel, TwLds, false, Rs1...>(cur, sd, scratch, t1, N, 1, b0, 1.0f, lds,
                                                                            tid);
            }
            WGFFT_LDS_SYNC();
        }
        __syncthreads();  // scratch stores of the whole workgroup visible to its loads
        // pass 2 (L = R1): scratch -> dst
        const float sc = INV ? 1.0f / (float)N : 1.0f;
        if (PF) {
            gd_load_side<R2, E2, G2, M, TwOneLevel, Rs2...>(sd, twN, nullptr, N, R1, 0, tid);
            prio_up();
            gd_load<R2, E2, G2, false, Rs2...>(cur, scratch, N / R2, 0, tid);
            prio_down();
        }
#pragma unroll 1
        for (uint64_t b0 = 0; b0 < N / R2; b0 += G2) {
            if constexpr (PF) {
                float2 nxt[E2];
                Side nsd;
                if (b0 + G2 < N / R2) {
                    gd_load_side<R2, E2, G2, M, TwOneLevel, Rs2...>(nsd, twN, nullptr, N, R1, b0 + G2, tid);
                    prio_up();
                    gd_load<R2, E2, G2, false, Rs2...>(nxt, scratch, N / R2, b0 + G2, tid);
                    prio_down();
                }
                gd_compute<R2, E2, G2, M, TwOneLevel, TwLds, NTIO, Rs2...>(cur, sd, dst, t2, N, R1, b0, sc, lds, tid);
#pragma unroll
                for (int i = 0; i < E2; ++i) cur[i] = nxt[i];
                sd = nsd;
            } else {
                gd_load_side<R2, E2, G2, M, TwOneLevel, Rs2...>(sd, twN, nullptr, N, R1, b0, tid);
                prio_up();
                gd_load<R2, E2, G2, false, Rs2...>(cur, scratch, N / R2, b0, tid);
                prio_down();
                gd_compute<R2, E2, G2, M, TwOneLevel, TwLds, NTIO, Rs2...>(cur, sd, dst, t2, N, R1, b0, sc, lds, tid);
            }
            WGFFT_LDS_SYNC();
        }
        __syncthreads();  // the next unit reuses this scratch slot
    }
};

template <bool INV, bool PF>
__global__ void __launch_bounds__(kRaggedThreads)
k_ragged(const float2* __restrict__ in, float2* __restrict__ out, const RaggedUnit* __restrict__ units,
         uint32_t u_begin, uint32_t u_end, const uint64_t* __restrict__ offs, const float2* __restrict__ tw_all,
         float2* __restrict__ scratch, uint32_t slot_log2, uint32_t* __restrict__ counters) {
    __shared__ __align__(16) float2 lds[kRaggedLds];
    __shared__ __align__(16) float2 lds_tw[twlds_len(kMaxLog2)];
    __shared__ uint32_t s_unit;
    float2* my_scratch = scratch + ((uint64_t)blockIdx.x << slot_log2);  // 2^slot_log2: the largest two-pass length
    // W_65536 as two LDS tables; every sub-FFT of length R reads it with
    // stride 65536 / R (TwLds::shifted), so no twiddle load of a unit waits
    // behind another unit's global loads
    const TwLds tl = twlds_fill<1 << kMaxLog2>(lds_tw, tw_all + (1u << kMaxLog2), threadIdx.x, kRaggedThreads);

    for (;;) {
        if (threadIdx.x == 0) {
            // one ticket per dequeue; every workgroup draws exactly one ticket past the end
            const uint32_t ticket = __hip_atomic_fetch_add(&counters[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_unit = u_begin + ticket;
            // the last ticket ever drawn in this launch resets the queue for the
            // next launch or graph replay: all other draws are ordered before it
            // on this one location, so no fence is needed
            if (ticket == (u_end - u_begin) + gridDim.x - 1)
                __hip_atomic_exchange(&counters[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t u = s_unit;
        __syncthreads();
        if (u >= u_end) break;
        const RaggedUnit un = units[u];
        const uint32_t tid = launder_lane(threadIdx.x);
        switch (un.log2n) {
            case 6: ragged_direct<64, 16, INV, 8, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 7: ragged_direct<128, 16, INV, 4, 4, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 8: ragged_direct<256, 16, INV, 8, 4, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 9: ragged_direct<512, 16, INV, 8, 8, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 10: ragged_direct<1024, 16, INV, 8, 16, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 11: ragged_direct<2048, 16, INV, 8, 4, 8, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 12: ragged_direct<4096, 16, INV, 8, 8, 8, 8>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 13: ragged_direct<8192, 32, INV, 16, 16, 16, 2>(in, out, offs, un.first, un.count, tl, lds, tid); break;
            case 14: {
                const uint64_t o = offs[un.first];
                ragged_split2<16384, 32, INV, 8, 16, 8, 8>(in + o, out + o, tl, lds, tid);
                break;
            }
            case 15: {
                const uint64_t o = offs[un.first];
                RaggedBig<15, 256, 16, Radices<16, 16>, 128, 16, Radices<4, 8, 4>, INV, PF>::run(in + o, out + o,
                                                                                              my_scratch, tw_all, tl, lds, tid);
                break;
            }
            default: {
                const uint64_t o = offs[un.first];
                RaggedBig<16, 256, 16, Radices<16, 16>, 256, 16, Radices<16, 16>, INV, PF>::run(in + o, out + o,
                                                                                              my_scratch, tw_all, tl, lds, tid);
                break;
            }
        }
        WGFFT_LDS_SYNC();
    }
}

static int ragged_blocks_per_cu(bool pf) {
    int b = 0;
    const void* f = pf ? (const void*)k_ragged<false, true> : (const void*)k_ragged<false, false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, kRaggedThreads, 0) != hipSuccess || b < 1) b = 1;
    return b;
}

}  // namespace wgfft

using namespace wgfft;

struct wgfft_ragged_plan {
    wgfft_t* h = nullptr;
    uint64_t count = 0;
    uint64_t total_samples = 0;
    std::vector<std::pair<uint32_t, uint32_t>> size_ranges;  // unit ranges per length (largest first)
    uint32_t nunits = 0;
    void* d_units = nullptr;
    void* d_offs = nullptr;
    void* d_tw = nullptr;
    void* d_counters = nullptr;
    void* d_scratch = nullptr;
    int grid = 0;      // workgroups of the chosen kernel variant
    int grid_max = 0;  // scratch slots allocated (largest grid of the variants)
    uint32_t slot_log2 = 0;  // scratch slot of 2^slot_log2 complex per workgroup (0: no two-pass length)
    int pf = -1;       // big-unit tile prefetch: -1 undecided (timed on first use)
    struct GraphKey {
        const void* in;
        void* out;
        int dir;
        void* stream;
        bool operator<(const GraphKey& o) const {
            return std::tie(in, out, dir, stream) < std::tie(o.in, o.out, o.dir, o.stream);
        }
    };
    std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

static void ragged_free(wgfft_ragged_t* p) {
    if (!p) return;
    for (auto& kv : p->graphs) {
        hipGraphExecDestroy(kv.second.second);
        hipGraphDestroy(kv.second.first);
    }
    for (void* q : {p->d_units, p->d_offs, p->d_tw, p->d_counters, p->d_scratch})
        if (q) hipFree(q);
    delete p;
}

namespace wgfft {
void ragged_destroy_all(Handle* h) {
    if (!h->ragged_plans.empty()) hipDeviceSynchronize();
    for (wgfft_ragged_t* p : h->ragged_plans) ragged_free(p);
    h->ragged_plans.clear();
}
}  // namespace wgfft

extern "C" int wgfft_ragged_create(wgfft_t* h, const uint64_t* offsets, const uint32_t* log2n, uint64_t count,
                                   wgfft_ragged_t** out) {
    if (out) *out = nullptr;
    if (!h || !out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "handle or out is NULL");
    if (count && (!offsets || !log2n)) return fail(WGFFT_ERR_INVALID_ARGUMENT, "offset or length table is NULL");
    if (count > 0xFFFFFFFFull) return fail(WGFFT_ERR_INVALID_ARGUMENT, "too many transforms");
    // validate lengths, alignment and overlap
    std::vector<uint32_t> by_off(count);
    std::iota(by_off.begin(), by_off.end(), 0u);
    for (uint64_t i = 0; i < count; ++i) {
        if (log2n[i] < (uint32_t)kMinLog2 || log2n[i] > (uint32_t)kMaxLog2)
            return fail(WGFFT_ERR_INVALID_LENGTH, "mixed-length transforms must have 2^6 <= N <= 2^16");
        if (offsets[i] & 1ull) return fail(WGFFT_ERR_INVALID_ARGUMENT, "offsets must be even (16-byte aligned)");
    }
    std::sort(by_off.begin(), by_off.end(), [&](uint32_t a, uint32_t b) { return offsets[a] < offsets[b]; });
    for (uint64_t i = 1; i < count; ++i) {
        const uint32_t a = by_off[i - 1], b = by_off[i];
        if (offsets[a] + (1ull << log2n[a]) > offsets[b]) return fail(WGFFT_ERR_INVALID_ARGUMENT, "transforms overlap");
    }
    // sort by length (largest first), stable in index order
    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return log2n[a] > log2n[b]; });
    std::vector<uint64_t> offs(count);
    std::vector<RaggedUnit> units;
    std::vector<std::pair<uint32_t, uint32_t>> ranges;
    uint64_t samples = 0;
    for (uint64_t i = 0; i < count;) {
        const uint32_t k = log2n[order[i]];
        const uint32_t u0 = (uint32_t)units.size();
        uint64_t j = i;
        while (j < count && log2n[order[j]] == k) {
            offs[j] = offsets[order[j]];
            samples += 1ull << k;
            ++j;
        }
        const uint32_t per = k <= (uint32_t)kDirectMaxLog2 ? (uint32_t)direct_g((int)k) : 1u;
        for (uint64_t s = i; s < j; s += per) {
            const uint32_t c = (uint32_t)std::min<uint64_t>(per, j - s);
            units.push_back(RaggedUnit{k, (uint32_t)s, c, 0u});
        }
        ranges.push_back({u0, (uint32_t)units.size()});
        i = j;
    }

    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto* p = new (std::nothrow) wgfft_ragged_plan();
    if (!p) return fail(WGFFT_ERR_OUT_OF_MEMORY, "plan allocation failed");
    p->h = h;
    p->count = count;
    p->total_samples = samples;
    p->size_ranges = ranges;
    p->nunits = (uint32_t)units.size();
    p->grid_max = std::max(ragged_blocks_per_cu(false), ragged_blocks_per_cu(true)) * h->num_cus;
    p->grid = p->grid_max;
    // twiddles: tw_all[N + j] = W_N^j for N = 2 .. 2^16
    const uint64_t tw_n = 2ull << kMaxLog2;
    std::vector<float2> tw(tw_n);
    for (int k = 0; k <= kMaxLog2; ++k) {
        const uint64_t n = 1ull << k;
        for (uint64_t j = 0; j < n; ++j) {
            double a = 2.0 * M_PI * ((double)j / (double)n);
            tw[n + j] = make_float2((float)std::cos(a), (float)-std::sin(a));
        }
    }
    tw[0] = make_float2(1.f, 0.f);
    // per-workgroup scratch only for the two-pass lengths, sized for the largest present
    for (const auto& u : units) p->slot_log2 = std::max(p->slot_log2, u.log2n > (uint32_t)kSplitLog2 ? u.log2n : 0u);
    const uint64_t scratch_bytes = p->slot_log2 ? ((uint64_t)p->grid_max << p->slot_log2) * sizeof(float2) : 0;
    struct {
        void** dst;
        const void* src;
        uint64_t bytes;
    } uploads[] = {{&p->d_units, units.data(), units.size() * sizeof(RaggedUnit)},
                   {&p->d_offs, offs.data(), offs.size() * sizeof(uint64_t)},
                   {&p->d_tw, tw.data(), tw.size() * sizeof(float2)},
                   {&p->d_counters, nullptr, 64},
                   {&p->d_scratch, nullptr, scratch_bytes}};
    for (auto& u : uploads) {
        const uint64_t b = std::max<uint64_t>(u.bytes, 16);
        if ((e = hipMalloc(u.dst, b)) != hipSuccess) {
            (void)hipGetLastError();
            ragged_free(p);
            return fail(WGFFT_ERR_OUT_OF_MEMORY, std::string("hipMalloc(ragged plan) failed: ") + hipGetErrorString(e));
        }
        if (u.src && u.bytes) e = hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice);
        else if (!u.src) e = hipMemset(*u.dst, 0, b);
        if (e != hipSuccess) {
            ragged_free(p);
            return hip_fail(e, "ragged plan upload");
        }
    }
    h->ragged_plans.insert(p);
    *out = p;
    return WGFFT_OK;
}

extern "C" void wgfft_ragged_destroy(wgfft_ragged_t* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(p->h->mu);
    hipSetDevice(p->h->device);
    hipDeviceSynchronize();
    p->h->ragged_plans.erase(p);
    ragged_free(p);
}

template <bool INV, bool PF>
static void ragged_launch_t(wgfft_ragged_t* p, const float* in, float* out, uint32_t u0, uint32_t u1, int grid,
                            hipStream_t s) {
    hipLaunchKernelGGL((k_ragged<INV, PF>), dim3(grid), dim3(kRaggedThreads), 0, s, (const float2*)in, (float2*)out,
                       (const RaggedUnit*)p->d_units, u0, u1, (const uint64_t*)p->d_offs, (const float2*)p->d_tw,
                       (float2*)p->d_scratch, p->slot_log2, (uint32_t*)p->d_counters);
}

static hipError_t ragged_launch_pf(wgfft_ragged_t* p, const float* in, float* out, bool inv, bool pf, uint32_t u0,
                                   uint32_t u1, int grid, hipStream_t s) {
    if (inv) {
        if (pf) ragged_launch_t<true, true>(p, in, out, u0, u1, grid, s);
        else ragged_launch_t<true, false>(p, in, out, u0, u1, grid, s);
    } else {
        if (pf) ragged_launch_t<false, true>(p, in, out, u0, u1, grid, s);
        else ragged_launch_t<false, false>(p, in, out, u0, u1, grid, s);
    }
    return hipGetLastError();
}

// First out-of-place use: time the kernel variants (with and without the
// big-unit tile prefetch) at each workgroups-per-CU count up to their
// occupancy over the whole queue (each launch rewrites the same output from
// the untouched input; best of three after a warm launch) and keep the
// fastest.  Fewer workgroups keep less of the big units' scratch in flight,
// which can then stay in the memory-side cache.  An in-place first call
// keeps the default variant and full occupancy.
static hipError_t ragged_choose(wgfft_ragged_t* p, const float* in, float* out, bool inv, hipStream_t s) {
    if (p->pf >= 0) return hipSuccess;
    if ((const void*)in == (const void*)out) {
        p->pf = 0;
        p->grid = std::max(ragged_blocks_per_cu(false), 1) * p->h->num_cus;
        return hipSuccess;
    }
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        hipEventDestroy(a);
        return e;
    }
    float best = 1e30f;
    const bool trace = std::getenv("WGFFT_RAGGED_TRACE") != nullptr;  // print the timings to stderr
    // settle the shader clock (it ramps up over tens of milliseconds of load)
    // so the candidates timed first are not penalised
    {
        const int grid = std::max(ragged_blocks_per_cu(false), 1) * p->h->num_cus;
        float ms = 0.f, total = 0.f;
        for (int r = 0; r < 64 && total < 30.f && e == hipSuccess; ++r) {
            hipEventRecord(a, s);
            ragged_launch_pf(p, in, out, inv, false, 0, p->nunits, grid, s);
            hipEventRecord(b, s);
            if ((e = hipEventSynchronize(b)) == hipSuccess) hipEventElapsedTime(&ms, a, b);
            total += ms;
        }
    }
    for (int pf = 0; pf < 2 && e == hipSuccess; ++pf) {
        const int occ = std::max(ragged_blocks_per_cu(pf != 0), 1);
        for (int per_cu = occ; per_cu >= 1 && e == hipSuccess; --per_cu) {
            const int grid = per_cu * p->h->num_cus;
            ragged_launch_pf(p, in, out, inv, pf != 0, 0, p->nunits, grid, s);  // warm
            for (int r = 0; r < 3 && e == hipSuccess; ++r) {
                hipEventRecord(a, s);
                ragged_launch_pf(p, in, out, inv, pf != 0, 0, p->nunits, grid, s);
                hipEventRecord(b, s);
                float ms = 0.f;
                if ((e = hipEventSynchronize(b)) == hipSuccess) hipEventElapsedTime(&ms, a, b);
                if (trace && e == hipSuccess)
                    std::fprintf(stderr, "wgfft ragged: prefetch=%d workgroups/CU=%d %.4f ms\n", pf, per_cu, ms);
                if (e == hipSuccess && ms < best) {
                    best = ms;
                    p->pf = pf;
                    p->grid = grid;
                }
            }
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    return e;
}

static hipError_t ragged_launch(wgfft_ragged_t* p, const float* in, float* out, bool inv, uint32_t u0, uint32_t u1,
                                int grid, hipStream_t s) {
    return ragged_launch_pf(p, in, out, inv, p->pf > 0, u0, u1, grid, s);
}

extern "C" int wgfft_ragged_exec_device(wgfft_ragged_t* p, const float* d_in, float* d_out, int dir, int mode,
                                        void* stream) {
    if (!p) return fail(WGFFT_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (dir != WGFFT_FORWARD && dir != WGFFT_INVERSE)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "direction must be WGFFT_FORWARD (-1) or WGFFT_INVERSE (+1)");
    if (mode < 0 || mode > 2) return fail(WGFFT_ERR_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
    if (p->count == 0) return WGFFT_OK;
    if (!d_in || !d_out) return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffer is NULL");
    if ((((uintptr_t)d_in | (uintptr_t)d_out) & 15u) != 0)
        return fail(WGFFT_ERR_INVALID_ARGUMENT, "device buffers must be 16-byte aligned");
    wgfft_t* h = p->h;
    std::lock_guard<std::mutex> lk(h->mu);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;  // NULL: HIP default stream
    const bool inv = dir == WGFFT_INVERSE;
    if ((e = ragged_choose(p, d_in, d_out, inv, s)) != hipSuccess) return hip_fail(e, "ragged variant selection");
    if (mode == 0) {
        if ((e = ragged_launch(p, d_in, d_out, inv, 0, p->nunits, p->grid, s)) != hipSuccess)
            return hip_fail(e, "ragged kernel launch");
        return WGFFT_OK;
    }
    if (mode == 2) {
        for (auto& r : p->size_ranges) {
            const int grid = (int)std::min<uint64_t>((uint64_t)p->grid, r.second - r.first);
            if ((e = ragged_launch(p, d_in, d_out, inv, r.first, r.second, grid, s)) != hipSuccess)
                return hip_fail(e, "ragged kernel launch");
        }
        return WGFFT_OK;
    }
    // capture once on the handle's private stream (the legacy default stream
    // cannot be captured), replay on the caller's stream
    wgfft_ragged_plan::GraphKey key{d_in, d_out, dir, nullptr};
    auto it = p->graphs.find(key);
    if (it == p->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        if ((e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(e, "hipStreamBeginCapture");
        hipError_t le = ragged_launch(p, d_in, d_out, inv, 0, p->nunits, p->grid, h->stream);
        e = hipStreamEndCapture(h->stream, &g);
        if (le != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return hip_fail(le, "ragged kernel capture");
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        if ((e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0)) != hipSuccess) {
            hipGraphDestroy(g);
            return hip_fail(e, "hipGraphInstantiate");
        }
        it = p->graphs.emplace(key, std::make_pair(g, ge)).first;
    }
    if ((e = hipGraphLaunch(it->second.second, s)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    return WGFFT_OK;
}

extern "C" uint64_t wgfft_ragged_samples(const wgfft_ragged_t* p) { return p ? p->total_samples : 0; }

extern "C" int wgfft_ragged_describe(const wgfft_ragged_t* p, char* desc, uint64_t desc_len) {
    if (!p) return fail(WGFFT_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (desc && desc_len)
        std::snprintf(desc, desc_len, "mixed N=2^%d..2^%d units=%u tile-prefetch=%s workgroups=%d (%d per CU)",
                      kMinLog2, kMaxLog2, p->nunits, p->pf < 0 ? "undecided" : p->pf ? "yes" : "no", p->grid,
                      p->grid / std::max(p->h->num_cus, 1));
    return WGFFT_OK;
}
