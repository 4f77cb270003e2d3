// kernels_adjoint.hpp -- fused differentiation kernels for Y = FFT(w * x).
//
//  k_fwd_adjoint  L = sum m |Y|^2: load x once, forward LDS FFT of w*x,
//                 gY = 2 m Y in registers (and m|Y|^2 into the loss), the
//                 conjugate-twiddle adjoint FFT of gY in the same workgroup,
//                 gx = w * gu stored, Re(conj(gu) x) accumulated per element
//                 in registers across the transforms this workgroup handles.
//  k_vjp          given gY: adjoint FFT of gY, gx = w * gu, the same gw
//                 accumulation.
// Both write per-workgroup partial sums (gw: N floats, loss: 1 float) into
// one row of a partial matrix that a single column-reduction launch folds in
// a fixed order, so results are deterministic.
#pragma once

#include "kernels.hpp"

namespace wgfft {

struct AdjointArgs {
    const float2* x;
    const float* w;
    const float* mask;  // fwd_adjoint
    const float2* gy;   // vjp
    float2* y;          // fwd_adjoint, optional
    float2* gx;
    const float2* tw;
    float* part;        // [grid][ld]: gw partials in columns < N, loss partial in column N
    uint64_t ld;
    uint64_t batch;
    int grid;
    hipStream_t stream;
};

template <int N, int E, int T>
__device__ __forceinline__ void load_real_pairs(float (&r)[E], const float* __restrict__ src, uint32_t t) {
    const float2* s2 = reinterpret_cast<const float2*>(src);
#pragma unroll
    for (int q = 0; q < E / 2; ++q) {
        float2 v = s2[t + T * q];
        r[2 * q] = v.x;
        r[2 * q + 1] = v.y;
    }
}

// Options (OPT bits; kAdjSplit2 marks the split-form kernels below):
//  kAdjKeepX     hold x in registers from the load to the gw accumulation
//                (one HBM read, more registers) instead of re-reading it after
//                the adjoint transform (the workgroup read it microseconds
//                earlier, so it comes from cache).  The VJP always re-reads.
//  kAdjTwLds     twiddles of both transforms from two short LDS tables, so no
//                transform waits on a global twiddle load.
// (A register prefetch of the next transform's input on top of kAdjTwLds was
// measured slower at N=4096: the extra registers cost a workgroup per CU.)
// Y and gx are written once and not read back here, so they are stored
// non-temporally, as in the forward kernels (4096 x 4096, default fused
// variant on one box: 66.6-66.9 us against 71.5-72.5 us with ordinary
// stores, with the Y output 77.6-77.9 against 95.2-96.0 us; the split form's
// Y at 8192 x 1024: 53.7-54.0 against 55.6-56.7 us).
constexpr int kAdjKeepX = 1, kAdjTwLds = 2, kAdjSplit2 = 4;
// AdjointVariant::roles: the calls a variant is the untuned default for
// (the first registered variant of a length serves any role none claims).
constexpr int kRoleFused = 1, kRoleVjp = 2, kRoleFilter = 4;

template <int N, int E, bool FWD, int OPT, int... Rs>
__global__ void __launch_bounds__(N / E)
k_adjoint(AdjointArgs a) {
    constexpr int T = N / E;
    constexpr int LEN = LdsLayout<N, E, Rs...>::len;
    constexpr bool KEEPX = (OPT & kAdjKeepX) != 0;
    constexpr bool TWLDS = (OPT & kAdjTwLds) != 0;
    using IO = Io<N, E, Rs...>;
    __shared__ __align__(16) float2 lds[LEN];
    __shared__ __align__(16) float2 lds_tw[TWLDS ? twlds_len(ilog2(N)) : 1];
    __shared__ float red[T];
    TwLds twl{};
    if constexpr (TWLDS) {
        twl = twlds_fill<N>(lds_tw, a.tw, threadIdx.x, T);
        __syncthreads();
    }
    const float2* src = FWD ? a.x : a.gy;  // the input each transform starts from

    float acc[E];
#pragma unroll
    for (int i = 0; i < E; ++i) acc[i] = 0.f;
    float lacc = 0.f;

    // w and m are re-read per transform (L1/L2 hits) to keep registers for
    // x, the transform data and the gw accumulators.
    for (uint64_t tr = blockIdx.x; tr < a.batch; tr += gridDim.x) {
        const uint32_t t = launder_lane(threadIdx.x);  // keep per-thread addressing inside the loop
        float2 d[E];
        [[maybe_unused]] float2 xk[KEEPX && FWD ? E : 1];
        float2 in0[E];
        __builtin_amdgcn_s_setprio(2);  // raised while the wave issues its global loads
        IO::load(in0, src + tr * N, t);
        [[maybe_unused]] float wv0[FWD ? E : 1];
        if constexpr (FWD) IO::load_real(wv0, a.w, t);
        __builtin_amdgcn_s_setprio(0);
        if constexpr (FWD) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                if constexpr (KEEPX) xk[i] = in0[i];
                d[i] = cscale(in0[i], wv0[i]);
            }
            if constexpr (TWLDS)
                Engine<N, E, T, false>::run_with(d, lds, twl, t, Radices<Rs...>{});
            else
                Engine<N, E, T, false>::run(d, lds, a.tw, t, Radices<Rs...>{});
            // keep the epilogue's loads out of the transform's register peak
            __builtin_amdgcn_sched_barrier(0);
            if (a.y) IO::template store<false, true>(a.y + tr * N, d, t, 1.f);
            // gY = 2 m Y; loss += m |Y|^2
            float mv[E];
            IO::load_real(mv, a.mask, t);
#pragma unroll
            for (int i = 0; i < E; ++i) {
                lacc += mv[i] * (d[i].x * d[i].x + d[i].y * d[i].y);
                d[i] = cscale(d[i], 2.f * mv[i]);
            }
            WGFFT_LDS_SYNC();  // forward's last exchange read before the adjoint writes
            __builtin_amdgcn_sched_barrier(0);
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) d[i] = in0[i];
        }
        // the adjoint reads the same twiddle entries as the forward; an opaque
        // table pointer makes it reload them (cache hits) instead of keeping
        // every forward twiddle power live in registers across both transforms
        // (the same for the thread index: the adjoint recomputes its LDS and
        // table addresses instead of keeping the forward's live)
        const uint32_t ta = launder_lane(t);
        if constexpr (TWLDS)
            Engine<N, E, T, true>::run_with(d, lds, twl, ta, Radices<Rs...>{});
        else
            Engine<N, E, T, true>::run(d, lds, launder(a.tw), ta, Radices<Rs...>{});
        __builtin_amdgcn_sched_barrier(0);
        // gu = d; gx = w gu; gw += Re(conj(gu) x)
        // x and w are read again through opaque pointers: otherwise the
        // compiler reuses the forward's loads of the same addresses and keeps
        // x and w live in registers across both transforms
        float2 xs[E];
        if constexpr (KEEPX && FWD) {
#pragma unroll
            for (int i = 0; i < E; ++i) xs[i] = xk[i];
        } else {
            IO::load(xs, launder(a.x) + tr * N, ta);
        }
        float wv[E];
        IO::load_real(wv, launder(a.w), ta);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            acc[i] += d[i].x * xs[i].x + d[i].y * xs[i].y;
            d[i] = cscale(d[i], wv[i]);
        }
        __builtin_amdgcn_s_setprio(2);  // and its stores
        IO::template store<false, true>(a.gx + tr * N, d, ta, 1.f);
        __builtin_amdgcn_s_setprio(0);
        WGFFT_LDS_SYNC();
    }
    const uint32_t t = threadIdx.x;
    IO::store_real(a.part + (uint64_t)blockIdx.x * a.ld, acc, t);
    if constexpr (FWD) {
        red[t] = lacc;
        __syncthreads();
        if (t == 0) {
            float s = 0.f;
            for (int i = 0; i < T; ++i) s += red[i];
            a.part[(uint64_t)blockIdx.x * a.ld + N] = s;
        }
    }
}

// Spectral filter y = F^{-1}(s * F(x + noise)) (1/N scaled, s real): the
// reference pipeline's noise (optional), FFT, spectral window and IFFT + 1/N
// stages (examples/parallel_pipeline.rs:29-105, run at :586-642) in one pass
// over memory.  The forward output stays in registers, is multiplied by s at its
// own indices and goes back through the conjugate-twiddle transform with the
// same plan, as the adjoint half of k_adjoint does.
struct FilterArgs {
    const float2* x;
    const float* s;
    float2* y;
    const float2* tw;
    uint64_t batch;
    int grid;
    hipStream_t stream;
    uint32_t seed;  // the pipeline's noise stage on the loaded samples when sigma != 0
    float sigma;
};

template <int N, int E, int OPT, bool NOISE, int... Rs>
__global__ void __launch_bounds__(N / E)
k_filter(FilterArgs a) {
    constexpr int T = N / E;
    constexpr int LEN = LdsLayout<N, E, Rs...>::len;
    constexpr bool TWLDS = (OPT & kAdjTwLds) != 0;
    using IO = Io<N, E, Rs...>;
    __shared__ __align__(16) float2 lds[LEN];
    __shared__ __align__(16) float2 lds_tw[TWLDS ? twlds_len(ilog2(N)) : 1];
    TwLds twl{};
    if constexpr (TWLDS) {
        twl = twlds_fill<N>(lds_tw, a.tw, threadIdx.x, T);
        __syncthreads();
    }
    for (uint64_t tr = blockIdx.x; tr < a.batch; tr += gridDim.x) {
        const uint32_t t = launder_lane(threadIdx.x);
        float2 d[E];
        __builtin_amdgcn_s_setprio(2);  // raised while the wave issues its global loads
        IO::load(d, a.x + tr * N, t);
        __builtin_amdgcn_s_setprio(0);
        if constexpr (NOISE) {  // a separate instantiation: the hash's registers cost the plain filter occupancy
#pragma unroll
            for (int i = 0; i < E; ++i) d[i] = cadd(d[i], pipeline_noise(IO::elem(i, t), (uint32_t)tr, a.seed, a.sigma));
        }
        if constexpr (TWLDS)
            Engine<N, E, T, false>::run_with(d, lds, twl, t, Radices<Rs...>{});
        else
            Engine<N, E, T, false>::run(d, lds, a.tw, t, Radices<Rs...>{});
        __builtin_amdgcn_sched_barrier(0);
        float sv[E];
        IO::load_real(sv, a.s, t);
#pragma unroll
        for (int i = 0; i < E; ++i) d[i] = cscale(d[i], sv[i]);
        WGFFT_LDS_SYNC();  // the forward's last exchange read before the inverse writes
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ta = launder_lane(t);
        if constexpr (TWLDS)
            Engine<N, E, T, true>::run_with(d, lds, twl, ta, Radices<Rs...>{});
        else
            Engine<N, E, T, true>::run(d, lds, launder(a.tw), ta, Radices<Rs...>{});
        __builtin_amdgcn_s_setprio(2);  // and its stores
        IO::template store<true, true>(a.y + tr * N, d, ta, 1.0f / (float)N);
        __builtin_amdgcn_s_setprio(0);
        WGFFT_LDS_SYNC();
    }
}

// ---------------------------------------------------------------------------
// The same two kernels in split form for N = 2H (the structure of k_split2):
// the forward is a decimation-in-frequency radix-2 stage in registers,
// y0[n] = a + b, y1[n] = (a - b) W_N^n (a = w x[n], b = w x[n + H]), then
// the H-point transforms of y0 and y1 in turn through one H-point LDS
// footprint, giving Y[2k] and Y[2k + 1]; the adjoint is its transpose: the
// conjugate-twiddle H-point transforms of gY[2k] and gY[2k + 1] give g0, g1,
// and gu[n] = g0 + conj(W_N^n) g1, gu[n + H] = g0 - conj(W_N^n) g1.  Half
// the LDS of k_adjoint and H/E threads, so three or four workgroups share a
// CU.  The gw partials are held in 2E registers across the transforms a
// workgroup handles, which costs the fused kernel its occupancy (two
// workgroups per CU at 191 VGPRs); the VJP and the filter keep more.
// Thread t holds the single-layout slots of both halves: slot i RF + m is
// index t + T i + m H/RF (RF the first radix of the H-point plan).
// WAVES: the minimum waves per SIMD the register allocation must allow.
// CH: elements per step of the closing gw / gx loop.  The loop re-reads x
// and w for CH elements at a time, one step ahead of the arithmetic (CH = E:
// every load up front), so fewer values are in flight next to the live
// transform and the gw partials.
// STOREY: Y is stored (the fused kernel with a Y output; compiled apart, as
// holding Y for the store next to gY costs registers).
// FL (kS2* bits): kS2Prefetch loads the next transform's input (x, or gY
// for the VJP) into registers at the top of each transform, so its HBM
// latency overlaps the two transforms instead of stalling the workgroup
// between them; kS2CachedX reads x with ordinary loads rather than
// non-temporal ones, so the closing re-read of x finds it in the L2;
// kS2KeepX0 / kS2KeepX1 hold the thread's x at n / at n + H in registers
// from the load to the gw accumulation instead of re-reading them (the
// re-read misses the L2 about half the time: 64 workgroups of an XCD keep
// 64 KB of x each in flight, the size of that L2).
constexpr int kS2Prefetch = 1, kS2CachedX = 2, kS2KeepX0 = 4, kS2KeepX1 = 8;
template <int N, int E, bool FWD, int WAVES, int CH, bool STOREY, int FL, int... Rs>
__global__ void __launch_bounds__((N / 2) / E) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_adjoint_split2(AdjointArgs a) {
    constexpr int H = N / 2;
    constexpr int T = H / E;
    constexpr bool PF = (FL & kS2Prefetch) != 0;
    constexpr bool XNT = (FL & kS2CachedX) == 0;
    constexpr bool KX0 = FWD && (FL & kS2KeepX0) != 0, KX1 = FWD && (FL & kS2KeepX1) != 0;
    static_assert(!(PF && (KX0 || KX1)), "the prefetch and the kept x share the registers");
    using MP = GdMap<H, E, 1, Rs...>;
    __shared__ __align__(16) float2 lds[LdsLayoutSingle<H, E, Rs...>::len];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(N))];
    __shared__ float red[T];
    const TwLds twn = twlds_fill<N>(lds_tw, a.tw, threadIdx.x, T);
    __syncthreads();
    const TwLds twh = twn.shifted(1);  // W_H for the halves
    typedef float f2v __attribute__((ext_vector_type(2)));
    float acc0[E], acc1[E];  // gw at indices n and n + H
#pragma unroll
    for (int e = 0; e < E; ++e) acc0[e] = acc1[e] = 0.f;
    float lacc = 0.f;
    // the input of transform tr: x at n and n + H (forward), or gY as
    // 16-byte pairs (gY[2k], gY[2k + 1]) at k (VJP)
    auto load_input = [&](uint64_t tr, uint32_t t, float2 (&i0)[E], float2 (&i1)[E]) {
        __builtin_amdgcn_s_setprio(2);
        if constexpr (FWD) {
            const float2* x = a.x + tr * N;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t n = MP::row(e, t);
                if constexpr (XNT) {
                    const f2v va = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n));
                    const f2v vb = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n + H));
                    i0[e] = make_float2(va.x, va.y);
                    i1[e] = make_float2(vb.x, vb.y);
                } else {
                    i0[e] = x[n];
                    i1[e] = x[n + H];
                }
            }
        } else {
            const float4* g4 = reinterpret_cast<const float4*>(a.gy + tr * N);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float4 v = ld16<true>(g4 + MP::row(e, t));
                i0[e] = make_float2(v.x, v.y);
                i1[e] = make_float2(v.z, v.w);
            }
        }
        __builtin_amdgcn_s_setprio(0);
    };
    float2 p0[PF ? E : 1], p1[PF ? E : 1];  // the prefetched input (PF)
    if constexpr (PF) {
        if (blockIdx.x < a.batch) load_input(blockIdx.x, threadIdx.x, p0, p1);
    }
    for (uint64_t tr = blockIdx.x; tr < a.batch; tr += gridDim.x) {
        const uint32_t t = launder_lane(threadIdx.x);
        float2 y0[E], y1[E];
        if constexpr (PF) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                y0[e] = p0[e];
                y1[e] = p1[e];
            }
            if (tr + gridDim.x < a.batch) load_input(tr + gridDim.x, t, p0, p1);
        } else {
            load_input(tr, t, y0, y1);
        }
        float2 xk0[KX0 ? E : 1], xk1[KX1 ? E : 1];  // x kept for the gw accumulation
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if constexpr (KX0) xk0[e] = y0[e];
            if constexpr (KX1) xk1[e] = y1[e];
        }
        if constexpr (FWD) {
            // W_N^n with n = t + offset(e): W_N^t times a compile-time root
            const float2 wt = cmul(twn.hi[t >> twn.h], twn.lo[t & ((1u << twn.h) - 1)]);
            static_for<E>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                constexpr CplxF c = cx_root(N, MP::offset(e));
                const uint32_t n = MP::row(e, t);
                const float2 p = cscale(y0[e], a.w[n]), q = cscale(y1[e], a.w[n + H]);
                y0[e] = cadd(p, q);
                y1[e] = cmul(csub(p, q), cmul(wt, make_float2(c.x, c.y)));
            });
            Engine<H, E, T, false>::run_single(y0, lds, twh, t, Radices<Rs...>{});
            WGFFT_LDS_SYNC();
            const uint32_t t1 = launder_lane(t);
            Engine<H, E, T, false>::run_single(y1, lds, twh, t1, Radices<Rs...>{});
            __builtin_amdgcn_sched_barrier(0);
            // Y[2k] = y0, Y[2k + 1] = y1 at k = row(e); gY = 2 m Y, loss += m |Y|^2
            const float2* m2 = reinterpret_cast<const float2*>(launder(a.mask));
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t k = MP::row(e, t1);
                if constexpr (STOREY)
                    st16<true>(reinterpret_cast<float4*>(a.y + tr * N + 2 * k),
                                make_float4(y0[e].x, y0[e].y, y1[e].x, y1[e].y));
                const float2 mv = m2[k];
                lacc += mv.x * (y0[e].x * y0[e].x + y0[e].y * y0[e].y) + mv.y * (y1[e].x * y1[e].x + y1[e].y * y1[e].y);
                y0[e] = cscale(y0[e], 2.f * mv.x);
                y1[e] = cscale(y1[e], 2.f * mv.y);
            }
            WGFFT_LDS_SYNC();  // the second half's last exchange read before the adjoint writes
            __builtin_amdgcn_sched_barrier(0);
        }
        const uint32_t ta = launder_lane(t);
        Engine<H, E, T, true>::run_single(y0, lds, twh, ta, Radices<Rs...>{});
        WGFFT_LDS_SYNC();
        const uint32_t tb = launder_lane(ta);
        Engine<H, E, T, true>::run_single(y1, lds, twh, tb, Radices<Rs...>{});
        __builtin_amdgcn_sched_barrier(0);
        // gu[n] = g0 + conj(W^n) g1, gu[n + H] = g0 - conj(W^n) g1; gx = w gu;
        // gw += Re(conj(gu) x), x re-read (it was loaded microseconds ago)
        const float2* x = launder(a.x) + tr * N;
        const float* wv = launder(a.w);
        float2* gx = a.gx + tr * N;
        const float2 wtb = cmul(twn.hi[tb >> twn.h], twn.lo[tb & ((1u << twn.h) - 1)]);
        static_assert(E % CH == 0, "the closing loop steps through whole chunks");
        if constexpr (CH == E) {
            static_for<E>([&](auto ec) {
                constexpr int e = decltype(ec)::value;
                constexpr CplxF c = cx_root(N, MP::offset(e));
                const uint32_t n = MP::row(e, tb);
                const float2 g1 = twmul<true>(y1[e], cmul(wtb, make_float2(c.x, c.y)));
                const float2 ga = cadd(y0[e], g1), gb = csub(y0[e], g1);
                const float2 xa = KX0 ? xk0[KX0 ? e : 0] : x[n], xb = KX1 ? xk1[KX1 ? e : 0] : x[n + H];
                acc0[e] += ga.x * xa.x + ga.y * xa.y;
                acc1[e] += gb.x * xb.x + gb.y * xb.y;
                y0[e] = cscale(ga, wv[n]);
                y1[e] = cscale(gb, wv[n + H]);
            });
            __builtin_amdgcn_s_setprio(2);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t n = MP::row(e, tb);
                const f2v va = {y0[e].x, y0[e].y}, vb = {y1[e].x, y1[e].y};
                __builtin_nontemporal_store(va, reinterpret_cast<f2v*>(gx + n));
                __builtin_nontemporal_store(vb, reinterpret_cast<f2v*>(gx + n + H));
            }
            __builtin_amdgcn_s_setprio(0);
            WGFFT_LDS_SYNC();  // the next transform's first half overwrites the exchange buffer
            continue;
        }
        constexpr int NC = E / CH;
        struct Chunk {
            float2 xa[CH], xb[CH];
            float wa[CH], wb[CH];
        };
        auto fetch = [&](Chunk& k, auto cc) {
            constexpr int c0 = decltype(cc)::value * CH;
            static_for<CH>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const uint32_t n = MP::row(c0 + i, tb);
                k.xa[i] = KX0 ? xk0[KX0 ? c0 + i : 0] : x[n];
                k.xb[i] = KX1 ? xk1[KX1 ? c0 + i : 0] : x[n + H];
                k.wa[i] = wv[n];
                k.wb[i] = wv[n + H];
            });
        };
        Chunk cur;
        fetch(cur, std::integral_constant<int, 0>{});
        static_for<NC>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            Chunk nxt;
            if constexpr (c + 1 < NC) fetch(nxt, std::integral_constant<int, c + 1>{});
            if constexpr (CH < E) __builtin_amdgcn_sched_barrier(0);
            static_for<CH>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                constexpr int e = c * CH + i;
                constexpr CplxF r = cx_root(N, MP::offset(e));
                const uint32_t n = MP::row(e, tb);
                const float2 g1 = twmul<true>(y1[e], cmul(wtb, make_float2(r.x, r.y)));
                const float2 ga = cadd(y0[e], g1), gb = csub(y0[e], g1);
                acc0[e] += ga.x * cur.xa[i].x + ga.y * cur.xa[i].y;
                acc1[e] += gb.x * cur.xb[i].x + gb.y * cur.xb[i].y;
                const float2 oa = cscale(ga, cur.wa[i]), ob = cscale(gb, cur.wb[i]);
                const f2v va = {oa.x, oa.y}, vb = {ob.x, ob.y};
                __builtin_nontemporal_store(va, reinterpret_cast<f2v*>(gx + n));
                __builtin_nontemporal_store(vb, reinterpret_cast<f2v*>(gx + n + H));
            });
            if constexpr (c + 1 < NC) {
                cur = nxt;
                if constexpr (CH < E) __builtin_amdgcn_sched_barrier(0);
            }
        });
        WGFFT_LDS_SYNC();  // the next transform's first half overwrites the exchange buffer
    }
    const uint32_t t = threadIdx.x;
    float* row = a.part + (uint64_t)blockIdx.x * a.ld;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t n = MP::row(e, t);
        row[n] = acc0[e];
        row[n + H] = acc1[e];
    }
    if constexpr (FWD) {
        red[t] = lacc;
        __syncthreads();
        if (t == 0) {
            float s = 0.f;
            for (int i = 0; i < T; ++i) s += red[i];
            row[N] = s;
        }
    }
}

// Spectral filter in the same split form: noise (optional), the forward
// radix-2 stage and halves, s at Y[2k] and Y[2k + 1], the conjugate halves
// and the transposed radix-2 stage, 1/N at the store.
template <int N, int E, bool NOISE, int... Rs>
__global__ void __launch_bounds__((N / 2) / E) __attribute__((amdgpu_waves_per_eu(4)))
k_filter_split2(FilterArgs a) {
    constexpr int H = N / 2;
    constexpr int T = H / E;
    using MP = GdMap<H, E, 1, Rs...>;
    __shared__ __align__(16) float2 lds[LdsLayoutSingle<H, E, Rs...>::len];
    __shared__ __align__(16) float2 lds_tw[twlds_len(ilog2(N))];
    const TwLds twn = twlds_fill<N>(lds_tw, a.tw, threadIdx.x, T);
    __syncthreads();
    const TwLds twh = twn.shifted(1);
    typedef float f2v __attribute__((ext_vector_type(2)));
    const float sc = 1.0f / (float)N;
    for (uint64_t tr = blockIdx.x; tr < a.batch; tr += gridDim.x) {
        const uint32_t t = launder_lane(threadIdx.x);
        const float2* x = a.x + tr * N;
        float2 y0[E], y1[E];
        __builtin_amdgcn_s_setprio(2);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t n = MP::row(e, t);
            const f2v va = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n));
            const f2v vb = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(x + n + H));
            y0[e] = make_float2(va.x, va.y);
            y1[e] = make_float2(vb.x, vb.y);
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (NOISE) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t n = MP::row(e, t);
                y0[e] = cadd(y0[e], pipeline_noise(n, (uint32_t)tr, a.seed, a.sigma));
                y1[e] = cadd(y1[e], pipeline_noise(n + H, (uint32_t)tr, a.seed, a.sigma));
            }
        }
        const float2 wt = cmul(twn.hi[t >> twn.h], twn.lo[t & ((1u << twn.h) - 1)]);
        static_for<E>([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            constexpr CplxF c = cx_root(N, MP::offset(e));
            const float2 p = y0[e], q = y1[e];
            y0[e] = cadd(p, q);
            y1[e] = cmul(csub(p, q), cmul(wt, make_float2(c.x, c.y)));
        });
        Engine<H, E, T, false>::run_single(y0, lds, twh, t, Radices<Rs...>{});
        WGFFT_LDS_SYNC();
        const uint32_t t1 = launder_lane(t);
        Engine<H, E, T, false>::run_single(y1, lds, twh, t1, Radices<Rs...>{});
        __builtin_amdgcn_sched_barrier(0);
        const float2* s2 = reinterpret_cast<const float2*>(launder(a.s));
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float2 sv = s2[MP::row(e, t1)];
            y0[e] = cscale(y0[e], sv.x);
            y1[e] = cscale(y1[e], sv.y);
        }
        WGFFT_LDS_SYNC();
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t ta = launder_lane(t1);
        Engine<H, E, T, true>::run_single(y0, lds, twh, ta, Radices<Rs...>{});
        WGFFT_LDS_SYNC();
        const uint32_t tb = launder_lane(ta);
        Engine<H, E, T, true>::run_single(y1, lds, twh, tb, Radices<Rs...>{});
        float2* o = a.y + tr * N;
        __builtin_amdgcn_s_setprio(2);
        const float2 wtb = cmul(twn.hi[tb >> twn.h], twn.lo[tb & ((1u << twn.h) - 1)]);
        static_for<E>([&](auto ec) {
            constexpr int e = decltype(ec)::value;
            constexpr CplxF c = cx_root(N, MP::offset(e));
            const uint32_t n = MP::row(e, tb);
            const float2 g1 = twmul<true>(y1[e], cmul(wtb, make_float2(c.x, c.y)));
            const float2 ga = cscale(cadd(y0[e], g1), sc), gb = cscale(csub(y0[e], g1), sc);
            const f2v va = {ga.x, ga.y}, vb = {gb.x, gb.y};
            __builtin_nontemporal_store(va, reinterpret_cast<f2v*>(o + n));
            __builtin_nontemporal_store(vb, reinterpret_cast<f2v*>(o + n + H));
        });
        __builtin_amdgcn_s_setprio(0);
        WGFFT_LDS_SYNC();
    }
}

template <int N, int E, int WAVES, int CH, int FL, int... Rs>
struct AdjointSplit2Kernel {
    static constexpr int THREADS = (N / 2) / E;
    template <bool FWD, bool STOREY>
    static hipError_t go(const AdjointArgs& a) {
        hipLaunchKernelGGL((k_adjoint_split2<N, E, FWD, WAVES, CH, STOREY, FL, Rs...>), dim3((unsigned)a.grid),
                           dim3(THREADS), 0, a.stream, a);
        return hipGetLastError();
    }
    static hipError_t launch(bool fwd, const AdjointArgs& a) {
        return fwd ? (a.y ? go<true, true>(a) : go<true, false>(a)) : go<false, false>(a);
    }
    template <bool FWD, bool STOREY>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, (const void*)k_adjoint_split2<N, E, FWD, WAVES, CH, STOREY, FL, Rs...>, THREADS, 0) !=
            hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    static int blocks_per_cu(bool fwd, bool y) {
        return fwd ? (y ? occ<true, true>() : occ<true, false>()) : occ<false, false>();
    }
    static hipError_t filter(const FilterArgs& a) {
        if (a.sigma != 0.f)
            hipLaunchKernelGGL((k_filter_split2<N, E, true, Rs...>), dim3((unsigned)a.grid), dim3(THREADS), 0,
                               a.stream, a);
        else
            hipLaunchKernelGGL((k_filter_split2<N, E, false, Rs...>), dim3((unsigned)a.grid), dim3(THREADS), 0,
                               a.stream, a);
        return hipGetLastError();
    }
    static int filter_blocks_per_cu(bool noise) {
        int blocks = 0;
        const void* f = noise ? (const void*)k_filter_split2<N, E, true, Rs...> : (const void*)k_filter_split2<N, E, false, Rs...>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, THREADS, 0) != hipSuccess) return 1;
        return blocks > 0 ? blocks : 1;
    }
};

template <int N, int E, int OPT, int... Rs>
struct AdjointKernel {
    static constexpr int THREADS = N / E;
    template <bool FWD>
    static hipError_t go(const AdjointArgs& a) {
        hipLaunchKernelGGL((k_adjoint<N, E, FWD, OPT, Rs...>), dim3((unsigned)a.grid), dim3(THREADS), 0, a.stream,
                           a);
        return hipGetLastError();
    }
    static hipError_t launch(bool fwd, const AdjointArgs& a) { return fwd ? go<true>(a) : go<false>(a); }
    template <bool FWD>
    static int occ() {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)k_adjoint<N, E, FWD, OPT, Rs...>,
                                                         THREADS, 0) != hipSuccess)
            return 1;
        return blocks > 0 ? blocks : 1;
    }
    static int blocks_per_cu(bool fwd, bool) { return fwd ? occ<true>() : occ<false>(); }
    static hipError_t filter(const FilterArgs& a) {
        if (a.sigma != 0.f)
            hipLaunchKernelGGL((k_filter<N, E, OPT, true, Rs...>), dim3((unsigned)a.grid), dim3(THREADS), 0, a.stream,
                               a);
        else
            hipLaunchKernelGGL((k_filter<N, E, OPT, false, Rs...>), dim3((unsigned)a.grid), dim3(THREADS), 0,
                               a.stream, a);
        return hipGetLastError();
    }
    static int filter_blocks_per_cu(bool noise) {
        int blocks = 0;
        const void* f = noise ? (const void*)k_filter<N, E, OPT, true, Rs...> : (const void*)k_filter<N, E, OPT, false, Rs...>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, THREADS, 0) != hipSuccess) return 1;
        return blocks > 0 ? blocks : 1;
    }
};

struct AdjointVariant {
    int log2n;
    int E;
    int opt;  // kAdj* bits
    const char* plan;
    hipError_t (*launch)(bool, const AdjointArgs&);
    int (*blocks_per_cu)(bool fwd, bool y);  // y: the fused kernel stores Y
    hipError_t (*filter)(const FilterArgs&);  // the spectral filter with the same plan
    int (*filter_blocks_per_cu)(bool noise);
    int roles;  // kRole* bits: untuned default for these calls
};

std::vector<AdjointVariant>& adjoint_registry();

struct AdjointRegistrar {
    explicit AdjointRegistrar(const AdjointVariant& v) { adjoint_registry().push_back(v); }
};

#define WGFFT_REGISTER_ADJOINT(N, E, OPT, PLAN, ...)                                                 \
    static ::wgfft::AdjointRegistrar _wgfft_adj_##N##_##E##_##OPT##_##PLAN(::wgfft::AdjointVariant{ \
        ::wgfft::ilog2(N), E, OPT, #PLAN, &::wgfft::AdjointKernel<N, E, OPT, __VA_ARGS__>::launch,   \
        &::wgfft::AdjointKernel<N, E, OPT, __VA_ARGS__>::blocks_per_cu,                              \
        &::wgfft::AdjointKernel<N, E, OPT, __VA_ARGS__>::filter,                                     \
        &::wgfft::AdjointKernel<N, E, OPT, __VA_ARGS__>::filter_blocks_per_cu, 0})

// Split-form variant of length N; PLAN is the radix split of N/2 (first
// radix = last radix, the single layout).  opt carries WAVES in bits 8-11,
// CH in bits 12-17 and the kS2* flags FL in bits 18-19.
#define WGFFT_ADJ_SPLIT2_T(N, E, WAVES, CH, FL, ...) ::wgfft::AdjointSplit2Kernel<N, E, WAVES, CH, FL, __VA_ARGS__>
#define WGFFT_REGISTER_ADJOINT_SPLIT2(N, E, WAVES, CH, FL, ROLES, PLAN, ...)                                  \
    static ::wgfft::AdjointRegistrar _wgfft_adjs2_##N##_##E##_##WAVES##_##CH##_##FL##_##PLAN(                 \
        ::wgfft::AdjointVariant{                                                                              \
            ::wgfft::ilog2(N), E, ::wgfft::kAdjSplit2 | ((WAVES) << 8) | ((CH) << 12) | ((FL) << 18),         \
            "2x" #PLAN, &WGFFT_ADJ_SPLIT2_T(N, E, WAVES, CH, FL, __VA_ARGS__)::launch,                        \
            &WGFFT_ADJ_SPLIT2_T(N, E, WAVES, CH, FL, __VA_ARGS__)::blocks_per_cu,                             \
            &WGFFT_ADJ_SPLIT2_T(N, E, WAVES, CH, FL, __VA_ARGS__)::filter,                                    \
            &WGFFT_ADJ_SPLIT2_T(N, E, WAVES, CH, FL, __VA_ARGS__)::filter_blocks_per_cu, ROLES})

}  // namespace wgfft
