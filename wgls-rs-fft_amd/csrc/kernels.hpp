// kernels.hpp -- kernel templates and the variant registry.
//
//  k_batched      one LDS-resident Stockham FFT per transform, G transforms per
//                 workgroup, grid-stride over the batch with the next group's
//                 input prefetched into registers (N = 32 .. 16384).
//  k_small        one thread per transform, register DFT (N = 1 .. 16).
//  k_global_pass  one pass of a multi-pass Stockham for N > 16384: a
//                 workgroup runs G radix-R butterflies (R up to 8192) as an
//                 LDS sub-FFT, twiddles W_{RL}^{mk} fused at load, tiles
//                 staged through LDS.
//  k_global_pass_direct  the same pass with each thread loading and storing
//                 its sub-FFT slots straight from the strided rows (no LDS
//                 staging of the I/O).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "fft_device.hpp"

namespace wgfft {

enum Mode : int { kForward = 0, kInverse = 1, kWindowed = 2 };

struct BatchedArgs {
    const float2* in;
    float2* out;
    const float2* tw;   // W_N^j, j < N
    const float* win;   // real window (kWindowed)
    uint64_t batch;
    float scale;        // applied at store (1/N for the inverse)
    int grid;           // workgroups to launch (0: one per group)
    hipStream_t stream;
    uint64_t idist = 0;  // complex elements between consecutive transforms of in (0: N)
    uint64_t odist = 0;  // ... of out (0: N)
};

// BatchedVariant::flags
constexpr int kFlagNtStore = 1;  // non-temporal output stores
constexpr int kFlagTwRegs = 2;   // twiddle base powers held in registers across transforms
constexpr int kFlagTwLds = 4;    // twiddle base powers from two short LDS tables
constexpr int kFlagNoPrefetch = 8;  // no register prefetch of the next transform (fewer registers)
constexpr int kFlagSplit2 = 256;    // k_split2: a radix-2 stage, then the two halves in turn through LDS
constexpr int kFlagPrio = 512;      // raised wave priority while issuing the prefetch loads and the stores
constexpr int kFlagHalfPrefetch = 2048;  // prefetch only the first E/2 slots of the next transform
constexpr int kFlagQuarterLoad = 4096;   // with kFlagHalfPrefetch: prefetch 3E/4 slots, load the last E/4 in turn
constexpr int kFlagCachedLoad = 8192;    // k_split2: input read with ordinary loads (else non-temporally)
// bits 4..7: minimum waves per SIMD the register allocation must allow (0: compiler's choice)
constexpr int flag_min_waves(int fl) { return (fl >> 4) & 15; }

struct BatchedVariant {
    int log2n;
    int E, G, flags, threads;
    int lds_bytes;
    const char* plan;  // radix split, e.g. "16x16x16"
    hipError_t (*launch)(Mode, const BatchedArgs&);
    int (*blocks_per_cu)(Mode);
};

struct GlobalPassArgs {
    const float2* in;
    float2* out;
    const float2* tw_r;   // W_R^j, j < R (sub-FFT table)
    const float2* tw_hi;  // W_N^{(j >> h) << h}
    const float2* tw_lo;  // W_N^{j & (2^h - 1)}
    const float* win;     // window for the first pass of the windowed forward
    uint32_t log2n;       // full transform length N = 2^log2n
    uint32_t h;           // split of the 2-level twiddle table
    uint64_t L;           // stride of this pass (product of earlier radices)
    uint64_t batch;
    float scale;
    int num_cus;
    hipStream_t stream;
};

struct GlobalPassVariant {
    int log2r;
    int E, G, prefetch, threads;
    int lds_bytes;
    const char* plan;
    hipError_t (*launch)(Mode, const GlobalPassArgs&);
};

std::vector<BatchedVariant>& batched_registry();
std::vector<GlobalPassVariant>& global_pass_registry();

struct Registrar {
    explicit Registrar(const BatchedVariant& v) { batched_registry().push_back(v); }
    explicit Registrar(const GlobalPassVariant& v) { global_pass_registry().push_back(v); }
};

hipError_t launch_small(Mode mode, int log2n, const BatchedArgs& a);

// ---------------------------------------------------------------------------
// Global I/O in the pair layout: d[2q + delta] = X[2t + delta + 2T q].
// 16-byte load, optionally non-temporal (streamed data read once).
template <bool NT>
__device__ __forceinline__ float4 ld16(const float4* p) {
    if constexpr (NT) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}
// 16-byte store, optionally non-temporal (streamed data not re-read soon).
template <bool NT>
__device__ __forceinline__ void st16(float4* p, float4 v) {
    if constexpr (NT) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v nv = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(nv, reinterpret_cast<f4v*>(p));
    } else {
        *p = v;
    }
}

template <int N, int E, int T, bool NT = false>
__device__ __forceinline__ void load_pairs(float2 (&d)[E], const float2* __restrict__ src, uint32_t t) {
    const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        float4 v = ld16<NT>(s4 + t + T * q);
        d[2 * q] = make_float2(v.x, v.y);
        d[2 * q + 1] = make_float2(v.z, v.w);
    }
}

template <int N, int E, int T>
__device__ __forceinline__ void window_pairs(float2 (&d)[E], const float* __restrict__ win, uint32_t t) {
    const float2* w2 = reinterpret_cast<const float2*>(win);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        float2 w = w2[t + T * q];
        d[2 * q] = cscale(d[2 * q], w.x);
        d[2 * q + 1] = cscale(d[2 * q + 1], w.y);
    }
}

// NT: non-temporal (streaming) stores -- the output is not re-read by the
// kernel, so it need not displace the input stream in L2.
template <int N, int E, int T, bool SCALE, bool NT = false>
__device__ __forceinline__ void store_pairs(float2* __restrict__ dst, const float2 (&d)[E], uint32_t t,
                                            float scale) {
    float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        float2 a = d[2 * q], b = d[2 * q + 1];
        if constexpr (SCALE) {
            a = cscale(a, scale);
            b = cscale(b, scale);
        }
        if constexpr (NT) {
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v v = {a.x, a.y, b.x, b.y};
            __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(d4) + t + T * q);
        } else {
            d4[t + T * q] = make_float4(a.x, a.y, b.x, b.y);
        }
    }
}

// Layout-generic global I/O of one transform (see IoKind): element index of
// register slot s of thread t, and complex / real loads and stores.
template <int N, int E, int... Rs>
struct Io {
    static constexpr int T = N / E;
    static constexpr bool PAIR = IoKind<E, Rs...>::PAIR;
    static constexpr int R = IoKind<E, Rs...>::RF;
    __device__ __forceinline__ static uint32_t elem(int s, uint32_t t) {
        if constexpr (PAIR) return 2 * t + (s & 1) + 2 * T * (s >> 1);
        else return t + T * (s / R) + (s % R) * (N / R);
    }
    template <bool NT = false>
    __device__ __forceinline__ static void load(float2 (&d)[E], const float2* __restrict__ src, uint32_t t) {
        if constexpr (PAIR) {
            load_pairs<N, E, T, NT>(d, src, t);
        } else {
#pragma unroll
            for (int s = 0; s < E; ++s) d[s] = src[elem(s, t)];
        }
    }
    // slots S0 <= s < S1 (S0, S1 even) into d[s - OFF]
    template <int S0, int S1, int OFF, int M>
    __device__ __forceinline__ static void load_slots(float2 (&d)[M], const float2* __restrict__ src, uint32_t t) {
        static_assert(S0 % 2 == 0 && S1 % 2 == 0 && S1 - OFF <= M && S0 >= OFF, "slot range");
        if constexpr (PAIR) {
            const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
            for (int q = S0 / 2; q < S1 / 2; ++q) {
                const float4 v = s4[t + T * q];
                d[2 * q - OFF] = make_float2(v.x, v.y);
                d[2 * q + 1 - OFF] = make_float2(v.z, v.w);
            }
        } else {
#pragma unroll
            for (int s = S0; s < S1; ++s) d[s - OFF] = src[elem(s, t)];
        }
    }
    template <bool SCALE, bool NT = false>
    __device__ __forceinline__ static void store(float2* __restrict__ dst, const float2 (&d)[E], uint32_t t,
                                                 float scale) {
        if constexpr (PAIR) {
            store_pairs<N, E, T, SCALE, NT>(dst, d, t, scale);
        } else {
#pragma unroll
            for (int s = 0; s < E; ++s) {
                float2 v = SCALE ? cscale(d[s], scale) : d[s];
                if constexpr (NT) {
                    typedef float f2v __attribute__((ext_vector_type(2)));
                    const f2v nv = {v.x, v.y};
                    __builtin_nontemporal_store(nv, reinterpret_cast<f2v*>(dst) + elem(s, t));
                } else {
                    dst[elem(s, t)] = v;
                }
            }
        }
    }
    __device__ __forceinline__ static void load_real(float (&r)[E], const float* __restrict__ src, uint32_t t) {
        if constexpr (PAIR) {
            const float2* s2 = reinterpret_cast<const float2*>(src);
#pragma unroll
            for (int q = 0; q < E / 2; ++q) {
                float2 v = s2[t + T * q];
                r[2 * q] = v.x;
                r[2 * q + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int s = 0; s < E; ++s) r[s] = src[elem(s, t)];
        }
    }
    __device__ __forceinline__ static void store_real(float* __restrict__ dst, const float (&r)[E], uint32_t t) {
        if constexpr (PAIR) {
            float2* d2 = reinterpret_cast<float2*>(dst);
#pragma unroll
            for (int q = 0; q < E / 2; ++q) d2[t + T * q] = make_float2(r[2 * q], r[2 * q + 1]);
        } else {
#pragma unroll
            for (int s = 0; s < E; ++s) dst[elem(s, t)] = r[s];
        }
    }
};

// ---------------------------------------------------------------------------
template <int N, int E, int G, int FL, int MODE, int... Rs>
__global__ void __launch_bounds__(G*(N / E)) __attribute__((amdgpu_waves_per_eu(flag_min_waves(FL) > 0 ? flag_min_waves(FL) : 1)))
k_batched(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
          const float* __restrict__ win, uint64_t batch, float scale, uint64_t idist, uint64_t odist) {
    constexpr int T = N / E;
    constexpr bool INV = MODE == kInverse;
    constexpr bool NT = (FL & kFlagNtStore) != 0;
    constexpr bool HOIST = (FL & kFlagTwRegs) != 0;
    constexpr bool TWLDS = (FL & kFlagTwLds) != 0;
    static_assert(!(HOIST && TWLDS), "one twiddle source per variant");
    constexpr int LEN = LdsLayout<N, E, Rs...>::len;
    using EN = Engine<N, E, T, INV>;
    constexpr int NTW = HOIST ? EN::hoisted_count(Radices<Rs...>{}) : 0;
    __shared__ __align__(16) float2 lds_all[G * LEN];
    __shared__ __align__(16) float2 lds_tw[TWLDS ? twlds_len(ilog2(N)) : 1];

    const uint32_t tid = threadIdx.x;
    const uint32_t g = tid / T, t = tid % T;
    float2* lds = lds_all + g * LEN;
    const uint64_t ngroups = (batch + G - 1) / G;

    using IO = Io<N, E, Rs...>;
    float2 twr[NTW > 0 ? NTW : 1];
    if constexpr (HOIST) EN::load_twiddles(twr, tw, t, Radices<Rs...>{});
    TwLds twl{};
    if constexpr (TWLDS) {
        twl = twlds_fill<N>(lds_tw, tw, tid, G * T);
        __syncthreads();
    }
    constexpr bool PREFETCH = (FL & kFlagNoPrefetch) == 0;
    // partial prefetch: slots [0, PS) of the next transform are held in
    // registers, slots [PS, E) are loaded when its turn comes
    constexpr bool HALF = PREFETCH && (FL & kFlagHalfPrefetch) != 0;
    constexpr int PS = (FL & kFlagQuarterLoad) ? 3 * E / 4 : E / 2;  // slots prefetched when HALF
    constexpr int NPF = PREFETCH ? (HALF ? PS : E) : 1;
    uint64_t grp = blockIdx.x;
    float2 nxt[NPF];
    if constexpr (PREFETCH) {
        if (grp < ngroups && grp * G + g < batch) {
            if constexpr (HALF) IO::template load_slots<0, PS, 0>(nxt, in + (grp * G + g) * idist, t);
            else IO::load(nxt, in + (grp * G + g) * idist, t);
        }
    }

    for (; grp < ngroups; grp += gridDim.x) {
        const uint64_t tr = grp * G + g;
        const bool active = tr < batch;
        float2 d[E];
        // Loads that this transform waits on are issued before the prefetch
        // (vmcnt retires in order).
        float w[E];
        if constexpr (MODE == kWindowed) IO::load_real(w, win, t);
        if constexpr (HALF) {
#pragma unroll
            for (int i = 0; i < PS; ++i) d[i] = nxt[i];
            const uint64_t ngrp = grp + gridDim.x;
            if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(2);
            if (active) IO::template load_slots<PS, E, 0>(d, in + tr * idist, t);
            if (ngrp < ngroups && ngrp * G + g < batch)
                IO::template load_slots<0, PS, 0>(nxt, in + (ngrp * G + g) * idist, t);
            if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(0);
        } else if constexpr (PREFETCH) {
#pragma unroll
            for (int i = 0; i < E; ++i) d[i] = nxt[i];
            const uint64_t ngrp = grp + gridDim.x;
            if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(2);
            if (ngrp < ngroups && ngrp * G + g < batch) IO::load(nxt, in + (ngrp * G + g) * idist, t);
            if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(0);
        } else {
            if (active) IO::load(d, in + tr * idist, t);
        }
        if constexpr (MODE == kWindowed) {
#pragma unroll
            for (int i = 0; i < E; ++i) d[i] = cscale(d[i], w[i]);
        }
        if constexpr (HOIST)
            EN::run_with(d, lds, TwHoisted{twr}, t, Radices<Rs...>{});
        else if constexpr (TWLDS)
            EN::run_with(d, lds, twl, t, Radices<Rs...>{});
        else
            EN::run(d, lds, tw, t, Radices<Rs...>{});
        if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(2);
        if (active) IO::template store<INV, NT>(out + tr * odist, d, t, scale);
        if constexpr ((FL & kFlagPrio) != 0) __builtin_amdgcn_s_setprio(0);
        WGFFT_LDS_SYNC();  // the next group's first pass overwrites the exchange buffer
    }
}

template <int N, int E, int G, int FL, int... Rs>
struct BatchedKernel {
    static constexpr int T = N / E;
    static constexpr int THREADS = G * T;
    static constexpr int LDS_BYTES =
        G * LdsLayout<N, E, Rs...>::len * 8 + ((FL & kFlagTwLds) ? twlds_len(ilog2(N)) * 8 : 0);

    template <int MODE>
    static hipError_t go(const BatchedArgs& a) {
        const uint64_t ngroups = (a.batch + G - 1) / G;
        uint64_t grid = a.grid > 0 ? (uint64_t)a.grid : ngroups;
        if (grid > ngroups) grid = ngroups;
        if (grid == 0) return hipSuccess;
        hipLaunchKernelGGL((k_batched<N, E, G, FL, MODE, Rs...>), dim3((unsigned)grid), dim3(THREADS), 0,
                           a.stream, a.in, a.out, a.tw, a.win, a.batch, a.scale, a.idist ? a.idist : (uint64_t)N,
                           a.odist ? a.odist : (uint64_t)N);
        return hipGetLastError();
    }
    static hipError_t launch(Mode m, const BatchedArgs& a) {
        switch (m) {
            case kForward: return go<kForward>(a);
            case kInverse: return go<kInverse>(a);
            default: return go<kWindowed>(a);
        }
    }
    template <int MODE>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)k_batched<N, E, G, FL, MODE, Rs...>,
                                                         THREADS, 0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    static int blocks_per_cu(Mode m) {
        switch (m) {
            case kForward: return occ<kForward>();
            case kInverse: return occ<kInverse>();
            default: return occ<kWindowed>();
        }
    }
};

// FL: kFlag* bits.
#define WGFFT_REGISTER_BATCHED_F(N, E, G, FL, PLAN, ...)                                                \
    static ::wgfft::Registrar _wgfft_reg_##N##_##E##_##G##_##FL##_##PLAN(::wgfft::BatchedVariant{      \
        ::wgfft::ilog2(N), E, G, FL, ::wgfft::BatchedKernel<N, E, G, FL, __VA_ARGS__>::THREADS,          \
        ::wgfft::BatchedKernel<N, E, G, FL, __VA_ARGS__>::LDS_BYTES, #PLAN,                             \
        &::wgfft::BatchedKernel<N, E, G, FL, __VA_ARGS__>::launch,                                      \
        &::wgfft::BatchedKernel<N, E, G, FL, __VA_ARGS__>::blocks_per_cu})
#define WGFFT_REGISTER_BATCHED(N, E, G, PLAN, ...) WGFFT_REGISTER_BATCHED_F(N, E, G, 0, PLAN, __VA_ARGS__)

// LDS slot (complex elements) per butterfly of a global pass: the larger of
// the sub-FFT's exchange footprint and the staging layout (R + R/32), kept
// 4 mod 32 so adjacent slots start on different banks.
template <int R, int E, int... Rs>
struct GlobalPassSlot {
    static constexpr int a = LdsLayout<R, E, Rs...>::len;
    static constexpr int b = R + (R >> 5);
    static constexpr int m = a > b ? a : b;
    static constexpr int len = m + ((4 - m % 32) + 32) % 32;
};

// Twiddle sources for W_N^j (forward sign).  raw(j) issues the table loads;
// value() combines them, so the loads of the next group can be issued ahead
// of its data prefetch and combined only when the group is transformed.
struct TwTwoLevel {
    const float2* hi;
    const float2* lo;
    uint32_t h;
    struct raw_t {
        float2 hi, lo;
    };
    __device__ __forceinline__ raw_t raw(uint64_t j) const { return {hi[j >> h], lo[j & ((1ull << h) - 1)]}; }
    __device__ __forceinline__ static float2 value(const raw_t& r) { return cmul(r.hi, r.lo); }
};
struct TwOneLevel {
    const float2* t;
    using raw_t = float2;
    __device__ __forceinline__ raw_t raw(uint64_t j) const { return t[j]; }
    __device__ __forceinline__ static float2 value(const raw_t& r) { return r; }
};

// One group of one multi-pass Stockham pass over a length-N transform with
// radix R done as an LDS sub-FFT: butterflies b0 .. b0+G-1 (G <= L, so both
// the strided loads and the strided stores are runs of G contiguous
// elements); when L == 1 each butterfly's R outputs are contiguous and are
// stored straight from registers.  src/dst point at the transform.
//
// Split in three so callers can software-pipeline: gp_load_side issues the
// loads the group's twiddles (or window) come from, gp_load issues the
// 16-byte loads of the R x G tile into E registers per thread, gp_compute
// applies the pass twiddles (or the window), stages the tile through LDS into
// the sub-FFT's pair layout, runs it and stores.  Lane i of step q holds rows
// m = m0 + 2T q (m0 = tid / (G/2)) of butterfly pair gp = tid mod (G/2), so
// its twiddles W_N^{m k s} (k = k0, k0 + 1) are W^{m0 k s} (W^{2T k s})^q: four
// table values per lane and group, advanced by one complex multiply per
// step.  The caller provides G * SLOT complex elements of LDS and the LDS
// barrier before the next gp_compute.
template <class TW, int E, int MODE>
struct GpSide {
    typename TW::raw_t a, da, b, db;  // W^{m0 k0 s}, W^{2T k0 s}, W^{m0 (k0+1) s}, W^{2T (k0+1) s}
    float2 w[MODE == kWindowed ? E / 2 : 1];  // window at the lane's element pairs (first pass)
};

template <int R, int E, int G, int MODE, class TW>
__device__ __forceinline__ void gp_load_side(GpSide<TW, E, MODE>& sd, const TW& twN, const float* __restrict__ win,
                                             uint64_t N, uint64_t L, uint64_t b0, uint32_t tid) {
    constexpr int T = R / E;
    const uint32_t gp = tid % (G / 2), m0 = tid / (G / 2);
    if (L > 1) {
        const uint64_t s = N / ((uint64_t)R * L);
        const uint64_t k0 = (b0 + 2 * gp) & (L - 1);
        sd.a = twN.raw((m0 * k0 * s) & (N - 1));
        sd.da = twN.raw((2 * T * k0 * s) & (N - 1));
        sd.b = twN.raw((m0 * (k0 + 1) * s) & (N - 1));
        sd.db = twN.raw((2 * T * (k0 + 1) * s) & (N - 1));
    } else if constexpr (MODE == kWindowed) {
        const uint64_t NR = N / R;
#pragma unroll
        for (int q = 0; q < E / 2; ++q)
            sd.w[q] = *reinterpret_cast<const float2*>(win + b0 + 2 * gp + (uint64_t)(m0 + 2 * T * q) * NR);
    }
}

template <int R, int E, int G, bool NT = false>
__device__ __forceinline__ void gp_load(float2 (&v)[E], const float2* __restrict__ src, uint64_t NR, uint64_t b0,
                                        uint32_t tid) {
    constexpr int THREADS = G * (R / E);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        const int it = tid + THREADS * q;
        const int gp = it % (G / 2);
        const int m = it / (G / 2);
        float4 x = ld16<NT>(reinterpret_cast<const float4*>(src + b0 + 2 * gp + (uint64_t)m * NR));
        v[2 * q] = make_float2(x.x, x.y);
        v[2 * q + 1] = make_float2(x.z, x.w);
    }
}

// NTS: the results are stored non-temporally.
template <int R, int E, int G, int MODE, class TW, class TS, bool NTS, int... Rs>
__device__ __forceinline__ void gp_compute(float2 (&v)[E], const GpSide<TW, E, MODE>& sd, float2* __restrict__ dst,
                                           const TS& ts, uint64_t N, uint64_t L, uint64_t b0, float sc,
                                           float2* __restrict__ lds_all, uint32_t tid) {
    constexpr int T = R / E;
    constexpr int THREADS = G * T;
    constexpr bool INV = MODE == kInverse;
    constexpr int SLOT = GlobalPassSlot<R, E, Rs...>::len;
    static_assert(IoKind<E, Rs...>::PAIR, "global-pass sub-FFTs use the pair layout");
    const bool first = (L == 1);

    float2 ta = make_float2(1.0f, 0.0f), tb = ta, da = ta, db = ta;
    if (!first) {
        ta = TW::value(sd.a);
        da = TW::value(sd.da);
        tb = TW::value(sd.b);
        db = TW::value(sd.db);
    }
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        const int it = tid + THREADS * q;
        const int gp = it % (G / 2);
        const int m = it / (G / 2);
        float2 a = v[2 * q], b = v[2 * q + 1];
        if constexpr (MODE == kWindowed) {
            if (first) {
                a = cscale(a, sd.w[q].x);
                b = cscale(b, sd.w[q].y);
            }
        }
        if (!first) {
            a = twmul<INV>(a, ta);
            b = twmul<INV>(b, tb);
            if (q + 1 < E / 2) {
                ta = cmul(ta, da);
                tb = cmul(tb, db);
            }
        }
        const uint32_t pm = pad_idx<5, 0>(m);
        lds_all[(2 * gp) * SLOT + pm] = a;
        lds_all[(2 * gp + 1) * SLOT + pm] = b;
    }
    WGFFT_LDS_SYNC();

    const uint32_t g = tid / T, t = tid % T;
    float2* lds = lds_all + g * SLOT;
    float2 d[E];
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        d[2 * q] = lds[pad_idx<5, 0>(2 * t + 2 * T * q)];
        d[2 * q + 1] = lds[pad_idx<5, 0>(2 * t + 1 + 2 * T * q)];
    }
    WGFFT_LDS_SYNC();
    Engine<R, E, T, INV>::run_with(d, lds, ts, t, Radices<Rs...>{});

    if (first) {
        // output b R + m: R contiguous elements per butterfly
        store_pairs<R, E, T, true, NTS>(dst + (b0 + g) * R, d, t, sc);
        return;
    }
    WGFFT_LDS_SYNC();
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        lds[pad_idx<5, 0>(2 * t + 2 * T * q)] = d[2 * q];
        lds[pad_idx<5, 0>(2 * t + 1 + 2 * T * q)] = d[2 * q + 1];
    }
    WGFFT_LDS_SYNC();
    const uint64_t bj = b0 / L, k0base = b0 & (L - 1);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        const int it = tid + THREADS * q;
        const int gp = it % (G / 2);
        const int m = it / (G / 2);
        const uint32_t pm = pad_idx<5, 0>(m);
        float2 a = cscale(lds_all[(2 * gp) * SLOT + pm], sc);
        float2 b = cscale(lds_all[(2 * gp + 1) * SLOT + pm], sc);
        const uint64_t o = bj * R * L + k0base + 2 * gp + (uint64_t)m * L;
        st16<NTS>(reinterpret_cast<float4*>(dst + o), make_float4(a.x, a.y, b.x, b.y));
    }
}

// ---------------------------------------------------------------------------
// Global pass with direct tile I/O: no LDS staging.  Thread (g, t) of the
// workgroup, g = tid mod G (column fastest), t = tid / G, holds slot
// i RF + m = row t + T i + m R/RF of column g in the single layout of its
// radix-R sub-FFT (RF its first and last radix).  A wavefront's 8-byte loads
// of one slot are 64/G rows x G adjacent columns, i.e. whole G*8-byte column
// runs; for L > 1 the results are stored the same way (rows m L apart), so a
// tile costs the sub-FFT's own exchanges only.  For L = 1 each column's R
// outputs are contiguous: they go through one LDS transposition and leave as
// contiguous runs.
template <int R, int E, int G, int... Rs>
struct GdMap {
    static constexpr int T = R / E;
    static constexpr int RF = RadixInfo<Rs...>::first;
    static constexpr int NG = E / RF;  // sub-FFT butterflies per thread in the first and last pass
    static_assert(RF == RadixInfo<Rs...>::last, "direct global passes need first radix = last radix");
    static_assert(E % RF == 0, "radix must divide E");
    __device__ __forceinline__ static uint32_t row(int s, uint32_t t) {
        return t + T * (s / RF) + (s % RF) * (R / RF);
    }
    // row(s, t) - t: the fixed part of the index of slot s
    static constexpr int offset(int s) { return T * (s / RF) + (s % RF) * (R / RF); }
};

// LDS slot (complex elements) per column of a direct global pass: the
// sub-FFT's exchange footprint or the transposition (R + R/32), made odd so
// the columns of a wavefront (column index fastest) start on different
// 8-byte bank slots.  Every access to it is 8 bytes wide.
template <int R, int E, int... Rs>
struct GdSlot {
    static constexpr int a = LdsLayoutSingle<R, E, Rs...>::len;
    static constexpr int b = R + (R >> 5);
    static constexpr int m = a > b ? a : b;
    static constexpr int len = m | 1;
};

// Twiddle bases of one tile: W_N^{row k s} for row = t + T i + (R/RF) m is
// u (W^{T k s})^i (W^{(R/RF) k s})^m with u = W^{t k s}; the windowed first
// pass needs the window at the thread's E elements.
template <class TW, int E, int MODE>
struct GdSide {
    typename TW::raw_t u, di, dm;
    float w[MODE == kWindowed ? E : 1];
};

template <int R, int E, int G, int MODE, class TW, int... Rs>
__device__ __forceinline__ void gd_load_side(GdSide<TW, E, MODE>& sd, const TW& twN, const float* __restrict__ win,
                                             uint64_t N, uint64_t L, uint64_t b0, uint32_t tid) {
    using MP = GdMap<R, E, G, Rs...>;
    const uint32_t g = tid % G, t = tid / G;
    if (L > 1) {
        const uint64_t s = N / ((uint64_t)R * L);
        const uint64_t k = (b0 + g) & (L - 1);
        sd.u = twN.raw(((uint64_t)t * k * s) & (N - 1));
        sd.di = twN.raw(((uint64_t)MP::T * k * s) & (N - 1));
        sd.dm = twN.raw(((uint64_t)(R / MP::RF) * k * s) & (N - 1));
    } else if constexpr (MODE == kWindowed) {
        const uint64_t NR = N / R;
#pragma unroll
        for (int e = 0; e < E; ++e) sd.w[e] = win[b0 + g + (uint64_t)MP::row(e, t) * NR];
    }
}

template <int R, int E, int G, bool NT, int... Rs>
__device__ __forceinline__ void gd_load(float2 (&v)[E], const float2* __restrict__ src, uint64_t NR, uint64_t b0,
                                        uint32_t tid) {
    using MP = GdMap<R, E, G, Rs...>;
    const uint32_t g = tid % G, t = tid / G;
    const float2* base = src + b0 + g;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const float2* p = base + (uint64_t)MP::row(e, t) * NR;
        if constexpr (NT) {
            typedef float f2v __attribute__((ext_vector_type(2)));
            const f2v x = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
            v[e] = make_float2(x.x, x.y);
        } else {
            v[e] = *p;
        }
    }
}

template <int R, int E, int G, int MODE, class TW, class TS, bool NTS, int... Rs>
__device__ __forceinline__ void gd_compute(float2 (&d)[E], const GdSide<TW, E, MODE>& sd, float2* __restrict__ dst,
                                           const TS& ts, uint64_t N, uint64_t L, uint64_t b0, float sc,
                                           float2* __restrict__ lds_all, uint32_t tid) {
    using MP = GdMap<R, E, G, Rs...>;
    constexpr int T = MP::T;
    constexpr int THREADS = G * T;
    constexpr int RF = MP::RF;
    constexpr bool INV = MODE == kInverse;
    constexpr int SLOT = GdSlot<R, E, Rs...>::len;
    const uint32_t g = tid % G, t = tid / G;
    const bool first = (L == 1);
    if (!first) {
        const float2 di = TW::value(sd.di), dm = TW::value(sd.dm);
        float2 ui = TW::value(sd.u);
#pragma unroll
        for (int i = 0; i < MP::NG; ++i) {
            float2 w = ui;
#pragma unroll
            for (int m = 0; m < RF; ++m) {
                d[i * RF + m] = twmul<INV>(d[i * RF + m], w);
                if (m + 1 < RF) w = cmul(w, dm);
            }
            if (i + 1 < MP::NG) ui = cmul(ui, di);
        }
    } else if constexpr (MODE == kWindowed) {
#pragma unroll
        for (int e = 0; e < E; ++e) d[e] = cscale(d[e], sd.w[e]);
    }
    float2* lds = lds_all + g * SLOT;
    Engine<R, E, T, INV>::run_single(d, lds, ts, t, Radices<Rs...>{});
    if (!first) {
        // output b + m L (b = b0 + g, rows of G contiguous butterflies)
        const uint64_t bj = b0 / L, k = (b0 & (L - 1)) + g;
        float2* o = dst + bj * R * L + k;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float2 v = cscale(d[e], sc);
            float2* p = o + (uint64_t)MP::row(e, t) * L;
            if constexpr (NTS) {
                typedef float f2v __attribute__((ext_vector_type(2)));
                const f2v nv = {v.x, v.y};
                __builtin_nontemporal_store(nv, reinterpret_cast<f2v*>(p));
            } else {
                *p = v;
            }
        }
        return;
    }
    // L = 1: output (b0 + g) R + q, contiguous per column: transpose through LDS
    WGFFT_LDS_SYNC();  // the sub-FFT's last exchange has been read
#pragma unroll
    for (int e = 0; e < E; ++e) lds[pad_idx<5, 0>(MP::row(e, t))] = d[e];
    WGFFT_LDS_SYNC();
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
        const uint32_t f = 2 * (tid + THREADS * e);  // element pair f, f + 1 of the tile's G x R outputs
        const uint32_t gg = f / R, q = f % R;
        const float2 a = cscale(lds_all[gg * SLOT + pad_idx<5, 0>(q)], sc);
        const float2 b = cscale(lds_all[gg * SLOT + pad_idx<5, 0>(q + 1)], sc);
        st16<NTS>(reinterpret_cast<float4*>(dst + (b0 + gg) * R + q), make_float4(a.x, a.y, b.x, b.y));
    }
}

// Persistent direct global pass (see gd_compute), tiles read and written
// non-temporally, next tile optionally prefetched into registers.
template <int R, int E, int G, bool PF, int MODE, int... Rs>
__global__ void __launch_bounds__(G*(R / E))
k_global_pass_direct(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw_r,
                     const float2* __restrict__ tw_hi, const float2* __restrict__ tw_lo,
                     const float* __restrict__ win, uint32_t log2n, uint32_t h, uint64_t L, uint64_t batch,
                     float scale) {
    constexpr int SLOT = GdSlot<R, E, Rs...>::len;
    __shared__ __align__(16) float2 lds_all[G * SLOT];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(R))];

    const uint64_t N = 1ull << log2n;
    const uint64_t NR = N / R;
    const uint64_t groups_per_tr = NR / G;
    const uint64_t total = groups_per_tr * batch;
    const uint64_t nb = gridDim.x;  // a multiple of 8
    const uint64_t pid = blockIdx.x;
    const uint64_t lid0 = (pid & 7) * (nb >> 3) + (pid >> 3);
    const bool last = (L * R == N);
    const TwTwoLevel twN{tw_hi, tw_lo, h};
    const TwLds twr = twlds_fill<R>(lds_tw, tw_r, threadIdx.x, G * (R / E));
    __syncthreads();
    using Side = GdSide<TwTwoLevel, E, MODE>;
    float2 cur[E];
    Side sd;
    if (PF && lid0 < total) {
        const uint64_t tr = lid0 / groups_per_tr, b0 = (lid0 % groups_per_tr) * G;
        gd_load_side<R, E, G, MODE, TwTwoLevel, Rs...>(sd, twN, win, N, L, b0, threadIdx.x);
        gd_load<R, E, G, true, Rs...>(cur, in + tr * N, NR, b0, threadIdx.x);
    }
    for (uint64_t lid = lid0; lid < total; lid += nb) {
        const uint64_t tr = lid / groups_per_tr, b0 = (lid % groups_per_tr) * G;
        const uint32_t tid = launder_lane(threadIdx.x);
        if constexpr (PF) {
            float2 nxt[E];
            Side nsd;
            const uint64_t nl = lid + nb;
            if (nl < total) {
                const uint64_t ntr = nl / groups_per_tr, nb0 = (nl % groups_per_tr) * G;
                gd_load_side<R, E, G, MODE, TwTwoLevel, Rs...>(nsd, twN, win, N, L, nb0, tid);
                __builtin_amdgcn_s_setprio(2);
                gd_load<R, E, G, true, Rs...>(nxt, in + ntr * N, NR, nb0, tid);
                __builtin_amdgcn_s_setprio(0);
            }
            gd_compute<R, E, G, MODE, TwTwoLevel, TwLds, true, Rs...>(cur, sd, out + tr * N, twr, N, L, b0,
                                                                      last ? scale : 1.0f, lds_all, tid);
#pragma unroll
            for (int i = 0; i < E; ++i) cur[i] = nxt[i];
            sd = nsd;
        } else {
            gd_load_side<R, E, G, MODE, TwTwoLevel, Rs...>(sd, twN, win, N, L, b0, tid);
            __builtin_amdgcn_s_setprio(2);
            gd_load<R, E, G, true, Rs...>(cur, in + tr * N, NR, b0, tid);
            __builtin_amdgcn_s_setprio(0);
            gd_compute<R, E, G, MODE, TwTwoLevel, TwLds, true, Rs...>(cur, sd, out + tr * N, twr, N, L, b0,
                                                                      last ? scale : 1.0f, lds_all, tid);
        }
        WGFFT_LDS_SYNC();
    }
}

template <int R, int E, int G, bool PF, int... Rs>
struct GlobalPassDirectKernel {
    static constexpr int T = R / E;
    static constexpr int THREADS = G * T;
    static constexpr int LDS_BYTES = G * GdSlot<R, E, Rs...>::len * 8 + twlds_len(ilog2(R)) * 8;

    template <int MODE>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, (const void*)k_global_pass_direct<R, E, G, PF, MODE, Rs...>, THREADS, 0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    template <int MODE>
    static hipError_t go(const GlobalPassArgs& a) {
        const uint64_t N = 1ull << a.log2n;
        const uint64_t groups = (N / R) / G * a.batch;
        if (groups == 0) return hipSuccess;
        static const int per_cu = occ<MODE>();
        uint64_t grid = std::min<uint64_t>(groups, (uint64_t)per_cu * (uint64_t)a.num_cus);
        if (grid % 8) grid += 8 - grid % 8;  // keep the XCD numbering a bijection; extra blocks exit
        hipLaunchKernelGGL((k_global_pass_direct<R, E, G, PF, MODE, Rs...>), dim3((unsigned)grid), dim3(THREADS), 0,
                           a.stream, a.in, a.out, a.tw_r, a.tw_hi, a.tw_lo, a.win, a.log2n, a.h, a.L, a.batch,
                           a.scale);
        return hipGetLastError();
    }
    static hipError_t launch(Mode m, const GlobalPassArgs& a) {
        switch (m) {
            case kForward: return go<kForward>(a);
            case kInverse: return go<kInverse>(a);
            default: return go<kWindowed>(a);
        }
    }
};

// prefetch = 3 (+ PF) in the registry marks the direct-I/O pipeline.
#define WGFFT_REGISTER_GLOBAL_PASS_DIRECT(R, E, G, PF, PLAN, ...)                                          \
    static ::wgfft::Registrar _wgfft_gpx_##R##_##E##_##G##_##PF##_##PLAN(::wgfft::GlobalPassVariant{    \
        ::wgfft::ilog2(R), E, G, 3 + PF, ::wgfft::GlobalPassDirectKernel<R, E, G, PF, __VA_ARGS__>::THREADS, \
        ::wgfft::GlobalPassDirectKernel<R, E, G, PF, __VA_ARGS__>::LDS_BYTES, #PLAN,                     \
        &::wgfft::GlobalPassDirectKernel<R, E, G, PF, __VA_ARGS__>::launch})

// ---------------------------------------------------------------------------
// Direct global pass over 16-column tiles, i.e. 128-byte column runs, with
// the LDS of C = 16/Q columns.  On MI355X a wavefront's 8-byte loads that
// cover 64-byte runs of many rows stream at well under the rate of 128-byte
// runs, and a tile of 16 radix-2048 columns (256 KB) does not fit in LDS.
// Each column's sub-FFT has T = R/E virtual threads; the workgroup has C T
// threads.  Real thread (g, t), g = tid mod 16, t = tid / 16 < T/Q, holds Q
// slot sets: set s is column g ^ (C s) at virtual thread t + (T/Q) s.  A
// wavefront's load or store of one slot of one set is then 4 rows x 16
// adjacent columns.  The sub-FFTs run in Q rounds of C columns (round q:
// columns C q .. C q + C - 1) through the same LDS; in round q thread (g, t)
// works on its set q ^ (g / C), which an XOR permutation of the sets by
// g / C (conditional swaps, one stage per bit) moves to index q.

// d[i] <- d[i ^ a] for i < Q (a < Q), one conditional-swap stage per bit of a
template <int Q, int E>
__device__ __forceinline__ void gw_permute(float2 (&d)[Q][E], uint32_t a) {
#pragma unroll
    for (int b = 1; b < Q; b <<= 1) {
        const bool c = (a & b) != 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (i & b) continue;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float2 x = c ? d[i + b][e] : d[i][e], y = c ? d[i][e] : d[i + b][e];
                d[i][e] = x;
                d[i + b][e] = y;
            }
        }
    }
}

// Pass twiddle bases of one column (u = W^{v k s}, di = W^{T k s},
// dm = W^{(R/RF) k s}) resolved to values.
struct GwTw {
    float2 u, di, dm;
};

// The sub-FFT of one round for the thread's column c (slot c mod C of the
// LDS) at virtual thread v, with the pass twiddles applied first (L > 1).
template <int R, int E, int Q, bool INV, class TS, int... Rs>
__device__ __forceinline__ void gw_round(float2 (&d)[E], const GwTw& tw, const TS& ts, uint64_t L, uint32_t cc,
                                         uint32_t v, float2* __restrict__ lds_all) {
    using MP = GdMap<R, E, 16 / Q, Rs...>;
    constexpr int RF = MP::RF;
    constexpr int SLOT = GdSlot<R, E, Rs...>::len;
    if (L > 1) {
        float2 ui = tw.u;
#pragma unroll
        for (int i = 0; i < MP::NG; ++i) {
            float2 w = ui;
#pragma unroll
            for (int m = 0; m < RF; ++m) {
                d[i * RF + m] = twmul<INV>(d[i * RF + m], w);
                if (m + 1 < RF) w = cmul(w, tw.dm);
            }
            if (i + 1 < MP::NG) ui = cmul(ui, tw.di);
        }
    }
    Engine<R, E, MP::T, INV>::run_single(d, lds_all + cc * SLOT, ts, v, Radices<Rs...>{});
}

// L = 1: the round's C columns (contiguous outputs, C R complex from
// dst + c0 R) leave through one LDS transposition.
template <int R, int E, int Q, int... Rs>
__device__ __forceinline__ void gw_store_first(const float2 (&d)[E], float2* __restrict__ dst, uint64_t c0, float sc,
                                               uint32_t cc, uint32_t v, float2* __restrict__ lds_all, uint32_t tid) {
    constexpr int C = 16 / Q;
    using MP = GdMap<R, E, C, Rs...>;
    constexpr int THREADS = C * MP::T;
    constexpr int SLOT = GdSlot<R, E, Rs...>::len;
    WGFFT_LDS_SYNC();  // the sub-FFT's last exchange has been read
    // opaque here, so no address of this step is computed early and held
    v = launder_lane(v);
    tid = launder_lane(tid);
    asm volatile("" : "+s"(c0));
    float2* lds = lds_all + cc * SLOT;
#pragma unroll
    for (int e = 0; e < E; ++e) lds[pad_idx<5, 0>(MP::row(e, v))] = d[e];
    WGFFT_LDS_SYNC();
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
        const uint32_t f = 2 * (tid + THREADS * e);  // element pair f, f + 1 of the round's C x R outputs
        const uint32_t gg = f / R, q = f % R;
        const float2 a = cscale(lds_all[gg * SLOT + pad_idx<5, 0>(q)], sc);
        const float2 b = cscale(lds_all[gg * SLOT + pad_idx<5, 0>(q + 1)], sc);
        st16<true>(reinterpret_cast<float4*>(dst + (c0 + gg) * R + q), make_float4(a.x, a.y, b.x, b.y));
    }
}

template <int R, int E, int Q, int MODE, int... Rs>
__global__ void __launch_bounds__((16 / Q) * (R / E))
k_global_pass_wide(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw_r,
                   const float2* __restrict__ tw_hi, const float2* __restrict__ tw_lo,
                   const float* __restrict__ win, uint32_t log2n, uint32_t h, uint64_t L, uint64_t batch,
                   float scale) {
    constexpr int G = 16;
    constexpr int C = G / Q;         // columns per round
    using MP = GdMap<R, E, C, Rs...>;
    constexpr int T = MP::T;         // virtual threads per column
    constexpr int TQ = T / Q;        // real threads per column group of a row run
    constexpr int THREADS = C * T;
    constexpr int SLOT = GdSlot<R, E, Rs...>::len;
    constexpr bool INV = MODE == kInverse;
    constexpr int TWMODE = MODE == kWindowed ? kForward : MODE;  // the window is applied at load
    static_assert(Q == 2 || Q == 4, "two or four rounds");
    __shared__ __align__(16) float2 lds_all[C * SLOT];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(R))];

    const uint64_t N = 1ull << log2n;
    const uint64_t NR = N / R;
    const uint64_t groups_per_tr = NR / G;
    const uint64_t total = groups_per_tr * batch;
    const uint64_t nb = gridDim.x;  // a multiple of 8
    const uint64_t pid = blockIdx.x;
    const uint64_t lid0 = (pid & 7) * (nb >> 3) + (pid >> 3);
    const bool last = (L * R == N);
    const float sc = last ? scale : 1.0f;
    const TwTwoLevel twN{tw_hi, tw_lo, h};
    const TwLds twr = twlds_fill<R>(lds_tw, tw_r, threadIdx.x, THREADS);
    __syncthreads();
    typedef float f2v __attribute__((ext_vector_type(2)));
    for (uint64_t lid = lid0; lid < total; lid += nb) {
        const uint64_t tr = lid / groups_per_tr, b0 = (lid % groups_per_tr) * G;
        const uint32_t tid = launder_lane(threadIdx.x);
        const uint32_t g = tid % G, t = tid / G, a = g / C, cc = g % C;
        // round q: column C q + cc at virtual thread t + TQ (q ^ a)
        float2 d[Q][E];
        const float2* src = in + tr * N + b0;
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int s = 0; s < Q; ++s)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const f2v x = __builtin_nontemporal_load(
                    reinterpret_cast<const f2v*>(src + (g ^ (C * s)) + (uint64_t)MP::row(e, t + TQ * s) * NR));
                d[s][e] = make_float2(x.x, x.y);
            }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (MODE == kWindowed) {
            if (L == 1) {
                const float* wsrc = win + b0;
#pragma unroll
                for (int s = 0; s < Q; ++s)
#pragma unroll
                    for (int e = 0; e < E; ++e)
                        d[s][e] = cscale(d[s][e], wsrc[(g ^ (C * s)) + (uint64_t)MP::row(e, t + TQ * s) * NR]);
            }
        }
        gw_permute<Q, E>(d, a);  // d[q]: round q's set
        float2* dst = out + tr * N;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            // opaque per round: nothing of a later round's sub-FFT is computed
            // early and held across an earlier one
            const uint32_t v = launder_lane(t + TQ * (q ^ a)), c = launder_lane(cc);
            // pass twiddle bases of the round, loaded when the round starts (held
            // for every round from before the tile loads, they cost spills)
            GwTw twq{};
            if (L > 1) {
                GdSide<TwTwoLevel, E, TWMODE> raw;
                gd_load_side<R, E, C, TWMODE, TwTwoLevel, Rs...>(raw, twN, nullptr, N, L, b0 + C * q, c + C * v);
                twq.u = TwTwoLevel::value(raw.u);
                twq.di = TwTwoLevel::value(raw.di);
                twq.dm = TwTwoLevel::value(raw.dm);
            }
            gw_round<R, E, Q, INV, TwLds, Rs...>(d[q], twq, twr, L, c, v, lds_all);
            if (L == 1) gw_store_first<R, E, Q, Rs...>(d[q], dst, b0 + C * q, sc, c, v, lds_all, tid);
            if (q + 1 < Q) WGFFT_LDS_SYNC();
        }
        if (L > 1) {
            // outputs b + m L of butterflies b0 .. b0 + 15, stored in the load pattern
            gw_permute<Q, E>(d, a);
            const uint64_t bj = b0 / L, k = b0 & (L - 1);
            uint64_t ob = bj * R * L + k;
            asm volatile("" : "+s"(ob));
            float2* o = dst + ob;
            const uint32_t tid2 = launder_lane(tid);
            const uint32_t gs = tid2 % G, ts = tid2 / G;
#pragma unroll
            for (int s = 0; s < Q; ++s)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const float2 v = cscale(d[s][e], sc);
                    const f2v nv = {v.x, v.y};
                    __builtin_nontemporal_store(
                        nv, reinterpret_cast<f2v*>(o + (gs ^ (C * s)) + (uint64_t)MP::row(e, ts + TQ * s) * L));
                }
        }
        WGFFT_LDS_SYNC();
    }
}

template <int R, int E, int Q, int... Rs>
struct GlobalPassWideKernel {
    static constexpr int T = R / E;
    static constexpr int THREADS = (16 / Q) * T;
    static constexpr int LDS_BYTES = (16 / Q) * GdSlot<R, E, Rs...>::len * 8 + twlds_len(ilog2(R)) * 8;

    template <int MODE>
    static hipError_t go(const GlobalPassArgs& a) {
        const uint64_t N = 1ull << a.log2n;
        const uint64_t groups = (N / R) / 16 * a.batch;
        if (groups == 0) return hipSuccess;
        static const int per_cu = [] {
            int blocks = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &blocks, (const void*)k_global_pass_wide<R, E, Q, MODE, Rs...>, THREADS, 0) != hipSuccess ||
                blocks < 1)
                blocks = 1;
            return blocks;
        }();
        uint64_t grid = std::min<uint64_t>(groups, (uint64_t)per_cu * (uint64_t)a.num_cus);
        if (grid % 8) grid += 8 - grid % 8;  // keep the XCD numbering a bijection; extra blocks exit
        hipLaunchKernelGGL((k_global_pass_wide<R, E, Q, MODE, Rs...>), dim3((unsigned)grid), dim3(THREADS), 0,
                           a.stream, a.in, a.out, a.tw_r, a.tw_hi, a.tw_lo, a.win, a.log2n, a.h, a.L, a.batch,
                           a.scale);
        return hipGetLastError();
    }
    static hipError_t launch(Mode m, const GlobalPassArgs& a) {
        switch (m) {
            case kForward: return go<kForward>(a);
            case kInverse: return go<kInverse>(a);
            default: return go<kWindowed>(a);
        }
    }
};

// prefetch = 5 (two rounds) or 6 (four rounds) in the registry marks the
// 16-column direct pipeline.
#define WGFFT_REGISTER_GLOBAL_PASS_WIDE(R, E, Q, PLAN, ...)                                            \
    static ::wgfft::Registrar _wgfft_gpw_##R##_##E##_##Q##_##PLAN(::wgfft::GlobalPassVariant{        \
        ::wgfft::ilog2(R), E, 16, Q == 2 ? 5 : 6, ::wgfft::GlobalPassWideKernel<R, E, Q, __VA_ARGS__>::THREADS, \
        ::wgfft::GlobalPassWideKernel<R, E, Q, __VA_ARGS__>::LDS_BYTES, #PLAN,                         \
        &::wgfft::GlobalPassWideKernel<R, E, Q, __VA_ARGS__>::launch})

// ---------------------------------------------------------------------------
// N = 2 H transforms with the LDS of an H-point one: a decimation-in-frequency
// radix-2 stage in registers, y0[n] = x[n] + x[n + H] and
// y1[n] = (x[n] - x[n + H]) W_N^n, then the two H-point sub-FFTs in turn
// through one LDS footprint, and X[2k] = Y0[k], X[2k + 1] = Y1[k] stored as
// 16-byte pairs.  At N = 16384 the footprint is 64 KB instead of 128 KB, so
// two workgroups share a CU and one's loads overlap the other's transform.
// Thread t holds the single-layout slots of its H-point sub-FFTs: slot
// i RF + m is index t + T i + m H/RF of both halves (RF the first radix).
// The wave priority is raised while it issues the global loads and stores
// (at 8192 x 1024, four workgroups per CU: 23-24 against 24-25 us).  The
// input is read non-temporally, or with ordinary loads under
// kFlagCachedLoad; which is faster depends on the working set and the
// 256 MB memory-side cache (same box, one buffer pair re-read each call):
// 8192 x 1024 (128 MiB) 21.2-22.3 us non-temporal against 23.5-23.7 us
// ordinary, 8192 x 2048 (256 MiB) 50.8 against 44.6 us, 16384 x 1024
// 52.1 against 48.7 us; over four rotating 256 MiB pairs non-temporal loads
// are ~1.5% faster.  Both forms are registered and the autotuner picks per
// shape.
template <int N, int E, int FL, int MODE, int... Rs>
__global__ void __launch_bounds__((N / 2) / E) __attribute__((amdgpu_waves_per_eu(4)))
k_split2(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw,
         const float* __restrict__ win, uint64_t batch, float scale, uint64_t idist, uint64_t odist) {
    constexpr int H = N / 2;
    constexpr int T = H / E;
    constexpr bool INV = MODE == kInverse;
    constexpr bool CACHED = (FL & kFlagCachedLoad) != 0;
    using MP = GdMap<H, E, 1, Rs...>;
    using EN = Engine<H, E, T, INV>;
    __shared__ __align__(16) float2 lds[LdsLayoutSingle<H, E, Rs...>::len];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(N))];
    const TwLds twn = twlds_fill<N>(lds_tw, tw, threadIdx.x, T);
    __syncthreads();
    const TwLds twh = twn.shifted(1);  // W_H for the halves
    typedef float f2v __attribute__((ext_vector_type(2)));
    for (uint64_t tr = blockIdx.x; tr < batch; tr += gridDim.x) {
        const uint32_t t = launder_lane(threadIdx.x);
        const float2* x = in + tr * idist;
        float2 y0[E], y1[E];
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t n = MP::row(e, t);
            f2v va, vb;
            if constexpr (CACHED) {
                va = *reinterpret_cast<const f2v*>(x + n);
                vb = *reinterpret_cast<const f2v*>(x + n + H);
            } else {
                va = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n));
                vb = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n + H));
            }
            y0[e] = make_float2(va.x, va.y);
            y1[e] = make_float2(vb.x, vb.y);
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (MODE == kWindowed) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t n = MP::row(e, t);
                y0[e] = cscale(y0[e], win[n]);
                y1[e] = cscale(y1[e], win[n + H]);
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t n = MP::row(e, t);  // < H = N/2, the range the tables hold
            const float2 w = cmul(twn.hi[n >> twn.h], twn.lo[n & ((1u << twn.h) - 1)]);
            const float2 a = y0[e], b = y1[e];
            y0[e] = cadd(a, b);
            y1[e] = twmul<INV>(csub(a, b), w);
        }
        EN::run_single(y0, lds, twh, t, Radices<Rs...>{});
        WGFFT_LDS_SYNC();  // the first half's last exchange has been read
        // opaque after the barrier: nothing of the second half is computed early
        const uint32_t t1 = launder_lane(t);
        EN::run_single(y1, lds, twh, t1, Radices<Rs...>{});
        float2* o = out + tr * odist;
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t k = MP::row(e, t1);
            const float2 a = cscale(y0[e], scale), b = cscale(y1[e], scale);
            st16<true>(reinterpret_cast<float4*>(o + 2 * k), make_float4(a.x, a.y, b.x, b.y));
        }
        __builtin_amdgcn_s_setprio(0);
        WGFFT_LDS_SYNC();  // the next transform's first pass overwrites the exchange buffer
    }
}

template <int N, int E, int FL, int... Rs>
struct Split2Kernel {
    static constexpr int T = (N / 2) / E;
    static constexpr int THREADS = T;
    static constexpr int LDS_BYTES = LdsLayoutSingle<N / 2, E, Rs...>::len * 8 + twlds_len(ilog2(N)) * 8;

    template <int MODE>
    static hipError_t go(const BatchedArgs& a) {
        uint64_t grid = a.grid > 0 ? (uint64_t)a.grid : a.batch;
        if (grid > a.batch) grid = a.batch;
        if (grid == 0) return hipSuccess;
        hipLaunchKernelGGL((k_split2<N, E, FL, MODE, Rs...>), dim3((unsigned)grid), dim3(THREADS), 0, a.stream, a.in,
                           a.out, a.tw, a.win, a.batch, a.scale, a.idist ? a.idist : (uint64_t)N,
                           a.odist ? a.odist : (uint64_t)N);
        return hipGetLastError();
    }
    static hipError_t launch(Mode m, const BatchedArgs& a) {
        switch (m) {
            case kForward: return go<kForward>(a);
            case kInverse: return go<kInverse>(a);
            default: return go<kWindowed>(a);
        }
    }
    template <int MODE>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)k_split2<N, E, FL, MODE, Rs...>, THREADS,
                                                         0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    static int blocks_per_cu(Mode m) {
        switch (m) {
            case kForward: return occ<kForward>();
            case kInverse: return occ<kInverse>();
            default: return occ<kWindowed>();
        }
    }
};

// A k_split2 variant of length N; PLAN is the radix split of N/2.
// FL: 0 or kFlagCachedLoad (8192).
#define WGFFT_REGISTER_SPLIT2_F(N, E, FL, PLAN, ...)                                                          \
    static ::wgfft::Registrar _wgfft_split2_##N##_##E##_##FL##_##PLAN(::wgfft::BatchedVariant{                \
        ::wgfft::ilog2(N), E, 1, ::wgfft::kFlagSplit2 | (FL), ::wgfft::Split2Kernel<N, E, FL, __VA_ARGS__>::THREADS, \
        ::wgfft::Split2Kernel<N, E, FL, __VA_ARGS__>::LDS_BYTES, "2x" #PLAN,                                  \
        &::wgfft::Split2Kernel<N, E, FL, __VA_ARGS__>::launch,                                              \
        &::wgfft::Split2Kernel<N, E, FL, __VA_ARGS__>::blocks_per_cu})
#define WGFFT_REGISTER_SPLIT2(N, E, PLAN, ...) WGFFT_REGISTER_SPLIT2_F(N, E, 0, PLAN, __VA_ARGS__)

// Persistent multi-pass kernel (tiles read and written non-temporally: each
// element is touched once per pass): workgroups loop over the groups of all
// transforms in the batch with the next group's tile prefetched into
// registers while the current one is transformed.  Consecutive workgroup ids
// are dealt round-robin over the 8 XCDs; the logical numbering makes the
// workgroups of one XCD take adjacent column groups at the same time, so the
// narrow column runs of neighbouring groups share that XCD's L2 lines.
template <int R, int E, int G, bool PF, int MODE, int... Rs>
__global__ void __launch_bounds__(G*(R / E))
k_global_pass(const float2* __restrict__ in, float2* __restrict__ out, const float2* __restrict__ tw_r,
              const float2* __restrict__ tw_hi, const float2* __restrict__ tw_lo,
              const float* __restrict__ win, uint32_t log2n, uint32_t h, uint64_t L, uint64_t batch,
              float scale) {
    constexpr int SLOT = GlobalPassSlot<R, E, Rs...>::len;
    __shared__ __align__(16) float2 lds_all[G * SLOT];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(R))];

    const uint64_t N = 1ull << log2n;
    const uint64_t NR = N / R;
    const uint64_t groups_per_tr = NR / G;
    const uint64_t total = groups_per_tr * batch;
    const uint64_t nb = gridDim.x;  // a multiple of 8
    const uint64_t pid = blockIdx.x;
    const uint64_t lid0 = (pid & 7) * (nb >> 3) + (pid >> 3);
    const bool last = (L * R == N);
    const TwTwoLevel twN{tw_hi, tw_lo, h};
    // the sub-FFT's twiddles come from LDS, so no wait on them stalls on the
    // prefetched tile (the tables are synchronised by gp_compute's first barrier)
    const TwLds twr = twlds_fill<R>(lds_tw, tw_r, threadIdx.x, G * (R / E));
    using Side = GpSide<TwTwoLevel, E, MODE>;
    float2 cur[E];
    Side sd;
    if (PF && lid0 < total) {
        const uint64_t tr = lid0 / groups_per_tr, b0 = (lid0 % groups_per_tr) * G;
        gp_load_side<R, E, G, MODE>(sd, twN, win, N, L, b0, threadIdx.x);
        gp_load<R, E, G, true>(cur, in + tr * N, NR, b0, threadIdx.x);
    }
    for (uint64_t lid = lid0; lid < total; lid += nb) {
        const uint64_t tr = lid / groups_per_tr, b0 = (lid % groups_per_tr) * G;
        const uint32_t tid = launder_lane(threadIdx.x);
        if constexpr (PF) {
            // next group's side values and tile in flight while this one is
            // transformed; they are issued after everything this group waits on
            float2 nxt[E];
            Side nsd;
            const uint64_t nl = lid + nb;
            if (nl < total) {
                const uint64_t ntr = nl / groups_per_tr, nb0 = (nl % groups_per_tr) * G;
                gp_load_side<R, E, G, MODE>(nsd, twN, win, N, L, nb0, tid);
                __builtin_amdgcn_s_setprio(2);
                gp_load<R, E, G, true>(nxt, in + ntr * N, NR, nb0, tid);
                __builtin_amdgcn_s_setprio(0);
            }
            gp_compute<R, E, G, MODE, TwTwoLevel, TwLds, true, Rs...>(cur, sd, out + tr * N, twr, N, L, b0,
                                                               last ? scale : 1.0f, lds_all, tid);
#pragma unroll
            for (int i = 0; i < E; ++i) cur[i] = nxt[i];
            sd = nsd;
        } else {
            gp_load_side<R, E, G, MODE>(sd, twN, win, N, L, b0, tid);
            __builtin_amdgcn_s_setprio(2);
            gp_load<R, E, G, true>(cur, in + tr * N, NR, b0, tid);
            __builtin_amdgcn_s_setprio(0);
            gp_compute<R, E, G, MODE, TwTwoLevel, TwLds, true, Rs...>(cur, sd, out + tr * N, twr, N, L, b0,
                                                               last ? scale : 1.0f, lds_all, tid);
        }
        WGFFT_LDS_SYNC();
    }
}

template <int R, int E, int G, bool PF, int... Rs>
struct GlobalPassKernel {
    static constexpr int T = R / E;
    static constexpr int THREADS = G * T;
    static constexpr int LDS_BYTES = G * GlobalPassSlot<R, E, Rs...>::len * 8 + twlds_len(ilog2(R)) * 8;

    template <int MODE>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)k_global_pass<R, E, G, PF, MODE, Rs...>,
                                                         THREADS, 0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    template <int MODE>
    static hipError_t go(const GlobalPassArgs& a) {
        const uint64_t N = 1ull << a.log2n;
        const uint64_t groups = (N / R) / G * a.batch;
        if (groups == 0) return hipSuccess;
        static const int per_cu = occ<MODE>();
        uint64_t grid = std::min<uint64_t>(groups, (uint64_t)per_cu * (uint64_t)a.num_cus);
        if (grid % 8) grid += 8 - grid % 8;  // keep the XCD numbering a bijection; extra blocks exit
        hipLaunchKernelGGL((k_global_pass<R, E, G, PF, MODE, Rs...>), dim3((unsigned)grid), dim3(THREADS), 0,
                           a.stream, a.in, a.out, a.tw_r, a.tw_hi, a.tw_lo, a.win, a.log2n, a.h, a.L,
                           a.batch, a.scale);
        return hipGetLastError();
    }
    static hipError_t launch(Mode m, const GlobalPassArgs& a) {
        switch (m) {
            case kForward: return go<kForward>(a);
            case kInverse: return go<kInverse>(a);
            default: return go<kWindowed>(a);
        }
    }
};

#define WGFFT_REGISTER_GLOBAL_PASS(R, E, G, PF, PLAN, ...)                                              \
    static ::wgfft::Registrar _wgfft_gp_##R##_##E##_##G##_##PF##_##PLAN(::wgfft::GlobalPassVariant{   \
        ::wgfft::ilog2(R), E, G, PF, ::wgfft::GlobalPassKernel<R, E, G, PF, __VA_ARGS__>::THREADS,       \
        ::wgfft::GlobalPassKernel<R, E, G, PF, __VA_ARGS__>::LDS_BYTES, #PLAN,                          \
        &::wgfft::GlobalPassKernel<R, E, G, PF, __VA_ARGS__>::launch})

}  // namespace wgfft
