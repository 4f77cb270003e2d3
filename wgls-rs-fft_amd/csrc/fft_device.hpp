// fft_device.hpp -- gfx950 device building blocks for the batched complex64
// Stockham FFT: register DFTs (radix 2/4/8/16), twiddle generation and the
// LDS-resident multi-pass Stockham engine shared by every kernel.
//
// Algorithm (the same decimation as the reference, wgls-rs-fft
// src/shaders.rs:18-62, generalised from radix 4 to radix R): pass s with
// stride L (product of the earlier radices) takes butterfly b, k = b mod L,
// j = b / L, reads x_m = X[b + m N/R] (m < R), multiplies by W_{RL}^{m k},
// does a length-R DFT and writes Y[j R L + k + m L].  The first pass reads
// global memory and the last pass writes it; the passes between exchange
// data through LDS.  Each thread owns E elements.  In the pair layout the
// first and last pass give thread t the element pairs (2t + 2T q,
// 2t + 1 + 2T q), so global traffic is 16-byte vectors and a wavefront
// touches 1 KiB contiguous; in the single layout (radix = E) thread t holds
// t + T i + m N/R in slot i R + m, 8-byte lanes, a wavefront 512 B
// contiguous per slot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

namespace wgfft {

// LDS barrier that does not wait for outstanding global loads: only LDS
// traffic (lgkmcnt) is drained before s_barrier, so loads prefetched for the
// next transform stay in flight across the exchange.
#define WGFFT_LDS_SYNC() asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory")

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

// Returns p unchanged but opaque to the optimiser.  Called once per
// iteration of a persistent loop, it keeps the per-thread twiddle loads (which
// do not depend on the transform) inside the loop instead of being hoisted
// and held in registers for every pass of every length the loop handles.
// The pointer is offset by an opaque zero rather than passed through the asm
// itself, so it keeps its global address space (global, not flat, loads:
// a flat load also counts on lgkmcnt, and every LDS wait would wait for it).
template <class P>
__device__ __forceinline__ P* launder(P* p) {
    uint32_t z = 0;
    asm volatile("" : "+s"(z));
    return p + z;
}
// The same for a per-lane value (the thread index of a persistent loop whose
// body dispatches to many code paths): nothing derived from it is hoisted.
__device__ __forceinline__ uint32_t launder_lane(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Compile-time unrolled loop: f(std::integral_constant<int, I>{}) for I < N,
// so the body can use I in constant expressions.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// exp(-2 pi i j / M) rounded to f32, evaluated at compile time (Taylor
// series in double after reducing the angle to [-pi, pi]), so multiplies by
// fixed roots of unity take immediate operands instead of table loads.
struct CplxF {
    float x, y;
};
constexpr double kPiD = 3.14159265358979323846264338327950288;
constexpr double cx_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 26; ++k) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}
constexpr double cx_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 26; ++k) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}
constexpr CplxF cx_root(long long M, long long j) {
    const long long r = ((j % M) + M) % M;
    double a = -2.0 * kPiD * (double)r / (double)M;
    if (a < -kPiD) a += 2.0 * kPiD;
    return CplxF{(float)cx_cos(a), (float)cx_sin(a)};
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// The pipeline demo's Gaussian noise for sample t of signal b
// (examples/parallel_pipeline.rs:29-63): Box-Muller from two PCG-RXS-M-XS
// hashes of (t, b, seed) in wrapping 32-bit arithmetic.
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float2 pipeline_noise(uint32_t t, uint32_t b, uint32_t seed, float sigma) {
    const uint32_t h1 = pcg_hash(t * 1234567u + b * 7654321u + seed);
    const uint32_t h2 = pcg_hash(h1 + 999983u);
    const float u1 = fmaxf((float)h1 * 2.3283064365386963e-10f, 1e-7f);
    // the angle 2 pi u2 with u2 = h2 / 2^32 in [0, 1]: the hardware sine and
    // cosine take their argument in turns, and the hardware log2 needs no
    // denormal path since u1 >= 1e-7 (-2 ln u1 = -2 ln 2 log2 u1)
    const float u2 = (float)h2 * 2.3283064365386963e-10f;
    const float r = __builtin_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1)) * sigma;
    return make_float2(r * __builtin_amdgcn_cosf(u2), r * __builtin_amdgcn_sinf(u2));
}

// a * W_4 with W_4 = -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_w4(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
template <bool INV>
__device__ __forceinline__ float2 conj_if(float2 w) {
    return INV ? make_float2(w.x, -w.y) : w;
}

constexpr float kR2 = 0.70710678118654752440f;  // 1/sqrt(2)
constexpr float kC8 = 0.92387953251128675613f;  // cos(pi/8)
constexpr float kS8 = 0.38268343236508977173f;  // sin(pi/8)

// a * W_8^1, W_8 = exp(-+ 2 pi i / 8)
template <bool INV>
__device__ __forceinline__ float2 mul_w8_1(float2 a) {
    return INV ? make_float2((a.x - a.y) * kR2, (a.x + a.y) * kR2)
               : make_float2((a.x + a.y) * kR2, (a.y - a.x) * kR2);
}
// a * W_8^3
template <bool INV>
__device__ __forceinline__ float2 mul_w8_3(float2 a) {
    return INV ? make_float2(-(a.x + a.y) * kR2, (a.x - a.y) * kR2)
               : make_float2((a.y - a.x) * kR2, -(a.x + a.y) * kR2);
}
// a * (c - i s) forward, a * (c + i s) inverse, for constants c, s
template <bool INV>
__device__ __forceinline__ float2 mul_cs(float2 a, float c, float s) {
    return INV ? make_float2(a.x * c - a.y * s, a.y * c + a.x * s)
               : make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
    float2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
    float2 s13 = cadd(x1, x3), d13 = mul_w4<INV>(csub(x1, x3));
    x0 = cadd(s02, s13);
    x2 = csub(s02, s13);
    x1 = cadd(d02, d13);
    x3 = csub(d02, d13);
}

// In-place DFT of v[0..R), natural order in and out.
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
    __device__ __forceinline__ static void run(float2 (&v)[2]) { dft2<INV>(v[0], v[1]); }
};

template <bool INV>
struct Dft<4, INV> {
    __device__ __forceinline__ static void run(float2 (&v)[4]) { dft4<INV>(v[0], v[1], v[2], v[3]); }
};

template <bool INV>
struct Dft<8, INV> {
    // 8 = 2 x 4: A_r = DFT4(v[r], v[r+2], v[r+4], v[r+6]); y[q] = A0 + W8^q A1,
    // y[q+4] = A0 - W8^q A1.
    __device__ __forceinline__ static void run(float2 (&v)[8]) {
        dft4<INV>(v[0], v[2], v[4], v[6]);
        dft4<INV>(v[1], v[3], v[5], v[7]);
        float2 a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
        float2 b0 = v[1], b1 = mul_w8_1<INV>(v[3]), b2 = mul_w4<INV>(v[5]), b3 = mul_w8_3<INV>(v[7]);
        v[0] = cadd(a0, b0); v[4] = csub(a0, b0);
        v[1] = cadd(a1, b1); v[5] = csub(a1, b1);
        v[2] = cadd(a2, b2); v[6] = csub(a2, b2);
        v[3] = cadd(a3, b3); v[7] = csub(a3, b3);
    }
};

template <bool INV>
struct Dft<16, INV> {
    // 16 = 4 x 4: m = 4 m1 + m2, q = q1 + 4 q2.  DFT4 over m1 for each m2,
    // multiply by W16^{m2 q1}, DFT4 over m2 for each q1.
    __device__ __forceinline__ static void run(float2 (&v)[16]) {
        dft4<INV>(v[0], v[4], v[8], v[12]);
        dft4<INV>(v[1], v[5], v[9], v[13]);
        dft4<INV>(v[2], v[6], v[10], v[14]);
        dft4<INV>(v[3], v[7], v[11], v[15]);
        // A_{m2}[q1] now sits in v[4 q1 + m2]
        v[5] = mul_cs<INV>(v[5], kC8, kS8);    // q1=1 m2=1: W^1
        v[6] = mul_w8_1<INV>(v[6]);            // q1=1 m2=2: W^2
        v[7] = mul_cs<INV>(v[7], kS8, kC8);    // q1=1 m2=3: W^3
        v[9] = mul_w8_1<INV>(v[9]);            // q1=2 m2=1: W^2
        v[10] = mul_w4<INV>(v[10]);            // q1=2 m2=2: W^4
        v[11] = mul_w8_3<INV>(v[11]);          // q1=2 m2=3: W^6
        v[13] = mul_cs<INV>(v[13], kS8, kC8);  // q1=3 m2=1: W^3
        v[14] = mul_w8_3<INV>(v[14]);          // q1=3 m2=2: W^6
        v[15] = mul_cs<INV>(v[15], -kC8, -kS8);  // q1=3 m2=3: W^9 = -(c - i s)^... (cos 9pi/8, -sin 9pi/8)
        dft4<INV>(v[0], v[1], v[2], v[3]);
        dft4<INV>(v[4], v[5], v[6], v[7]);
        dft4<INV>(v[8], v[9], v[10], v[11]);
        dft4<INV>(v[12], v[13], v[14], v[15]);
        // y[q1 + 4 q2] sits in v[4 q1 + q2]: transpose the 4x4 index.
        float2 t[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
        for (int q1 = 0; q1 < 4; ++q1)
#pragma unroll
            for (int q2 = 0; q2 < 4; ++q2) v[q1 + 4 * q2] = t[4 * q1 + q2];
    }
};

// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// v * w (forward) or v * conj(w) (inverse), w in forward-sign form
template <bool INV>
__device__ __forceinline__ float2 twmul(float2 v, float2 w) {
    return INV ? cmulc(v, w) : cmul(v, w);
}

// Number of base powers w, w^2, w^4, w^8 a radix-R butterfly's twiddles are
// formed from.
constexpr int tw_bases(int R) { return R <= 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4; }

// v[m] *= w^m (w^-m for the inverse), given wb[i] = w^(2^i) in forward-sign
// form; the other powers are products of at most two of them.  The powers are
// conjugated inside the multiply, so forward and inverse transforms of the
// same length share them.
template <int R, bool INV>
__device__ __forceinline__ void apply_twiddle_powers(float2 (&v)[R], const float2* wb) {
    const float2 w1 = wb[0];
    if constexpr (R == 2) {
        v[1] = twmul<INV>(v[1], w1);
    } else {
        const float2 w2 = wb[1];
        const float2 w3 = cmul(w1, w2);
        v[1] = twmul<INV>(v[1], w1);
        v[2] = twmul<INV>(v[2], w2);
        v[3] = twmul<INV>(v[3], w3);
        if constexpr (R >= 8) {
            const float2 w4 = wb[2];
            const float2 w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
            v[4] = twmul<INV>(v[4], w4);
            v[5] = twmul<INV>(v[5], w5);
            v[6] = twmul<INV>(v[6], w6);
            v[7] = twmul<INV>(v[7], w7);
            if constexpr (R == 16) {
                const float2 w8 = wb[3];
                v[8] = twmul<INV>(v[8], w8);
                v[9] = twmul<INV>(v[9], cmul(w8, w1));
                v[10] = twmul<INV>(v[10], cmul(w8, w2));
                v[11] = twmul<INV>(v[11], cmul(w8, w3));
                v[12] = twmul<INV>(v[12], cmul(w8, w4));
                v[13] = twmul<INV>(v[13], cmul(w8, w5));
                v[14] = twmul<INV>(v[14], cmul(w8, w6));
                v[15] = twmul<INV>(v[15], cmul(w8, w7));
            }
        }
    }
}

// v[m] *= W_{RL}^{m k} = W_N^{m k s} (conjugated for the inverse), with
// tw[j] = W_N^j (forward sign) and base = k s: the base powers are read from
// the table at base, 2 base, 4 base, 8 base.
template <int R, bool INV>
__device__ __forceinline__ void apply_twiddles(float2 (&v)[R], const float2* __restrict__ tw,
                                               uint32_t base) {
    float2 wb[tw_bases(R)];
#pragma unroll
    for (int i = 0; i < tw_bases(R); ++i) wb[i] = tw[base << i];
    apply_twiddle_powers<R, INV>(v, wb);
}

// LDS index padding for one exchange: phys = i + ((i >> A) << B).
template <int A, int B>
__device__ __forceinline__ uint32_t pad_idx(uint32_t i) {
    return i + ((i >> A) << B);
}

// Padded LDS footprint (in complex elements) of one length-N transform for a
// given exchange padding.
constexpr int padded_len(int n, int a, int b) { return n + ((n >> a) << b); }

// Compile-time radix list.
template <int... Rs>
struct Radices {};

template <int... Rs>
struct RadixInfo;
template <int R>
struct RadixInfo<R> {
    static constexpr int product = R;
    static constexpr int first = R;
    static constexpr int last = R;
    static constexpr int count = 1;
};
template <int R, int... Rest>
struct RadixInfo<R, Rest...> {
    static constexpr int product = R * RadixInfo<Rest...>::product;
    static constexpr int first = R;
    static constexpr int last = RadixInfo<Rest...>::last;
    static constexpr int count = 1 + RadixInfo<Rest...>::count;
};

// Global I/O layout of a plan.  PAIR (first and last radix <= E/2): thread t
// owns element pairs, register slot 2q + delta = element 2t + delta + 2T q,
// so global traffic is 16-byte vectors.  SINGLE (first radix = last radix =
// R, E/R butterflies per thread): slot i R + m = element t + T i + m N/R,
// 8-byte lanes, one radix-R butterfly per slot group with no extra
// exchange; it lets E = 16 run radix-16 passes.
template <int E, int... Rs>
struct IoKind {
    static constexpr int RF = RadixInfo<Rs...>::first;
    static constexpr int RL = RadixInfo<Rs...>::last;
    static constexpr bool PAIR = (E % (2 * RF) == 0) && (E % (2 * RL) == 0);
    static_assert(PAIR || RF == RL, "single-layout plans need equal first and last radix");
};

// Exchange padding written after a pass of radix R at stride L.
//  first pass, pair layout (output b R + m, lanes 2R apart): one pad element
//  per 2R staggers the lanes over all banks; single layout (lanes R apart):
//  one per R;
//  later passes with L < 32 write runs of L contiguous elements spaced R L
//  apart, so a pad of L per R L staggers the runs;
//  L >= 32 writes are already contiguous.
template <int R, int L, bool FIRST, bool PAIR>
struct ExchangePad {
    static constexpr int A = FIRST ? ilog2(PAIR ? 2 * R : R) : (L < 32 ? ilog2(R * L) : 30);
    static constexpr int B = FIRST ? 0 : (L < 32 ? ilog2(L) : 0);
};

template <int N, int E, int... Rs>
struct LdsLayout;
// Max padded length over the exchanges of a plan (the LDS footprint).
template <int N, int E, bool PAIR>
struct LdsLayoutAcc {
    template <int L, bool FIRST, int R, int... Rest>
    static constexpr int max_len() {
        if constexpr (sizeof...(Rest) == 0) {
            return N;
        } else {
            using P = ExchangePad<R, L, FIRST, PAIR>;
            constexpr int here = padded_len(N, P::A, P::B);
            constexpr int rest = max_len<L * R, false, Rest...>();
            return here > rest ? here : rest;
        }
    }
};

template <int N, int E, int... Rs>
struct LdsLayout {
    static constexpr int raw =
        LdsLayoutAcc<N, E, IoKind<E, Rs...>::PAIR>::template max_len<1, true, Rs...>();
    // keep every transform's LDS slot 16-byte aligned
    static constexpr int len = (raw + 1) & ~1;
};

// The footprint of the same plan run in the single layout (Engine::run_single).
template <int N, int E, int... Rs>
struct LdsLayoutSingle {
    static constexpr int len = LdsLayoutAcc<N, E, false>::template max_len<1, true, Rs...>();
};

// Sources of the twiddle base powers w^(2^i), i < tw_bases(R), of butterfly
// slot j of a pass (its twiddle index base = k N / (R L); TO = the pass's
// first entry in a register-held set).
//  TwTable:   read from the W_N table in global memory (tw[base << i]).
//  TwHoisted: register copies loaded once per thread by load_twiddles.
//  TwLds:     W_N^j = hi[J >> h] * lo[J & (2^h - 1)], J = j << sh, from two
//             short tables of a length-(N << sh) table in LDS (2 sqrt entries);
//             LDS reads retire on lgkmcnt, so they never wait behind
//             prefetched global loads.
struct TwTable {
    const float2* __restrict__ tw;
    template <int R, int TO>
    __device__ __forceinline__ void bases(float2* wb, int, uint32_t base) const {
#pragma unroll
        for (int i = 0; i < tw_bases(R); ++i) wb[i] = tw[base << i];
    }
};
struct TwHoisted {
    const float2* twr;
    template <int R, int TO>
    __device__ __forceinline__ void bases(float2* wb, int j, uint32_t) const {
#pragma unroll
        for (int i = 0; i < tw_bases(R); ++i) wb[i] = twr[TO + j * tw_bases(R) + i];
    }
};
struct TwLds {
    const float2* lo;
    const float2* hi;
    uint32_t h;
    uint32_t sh;
    template <int R, int TO>
    __device__ __forceinline__ void bases(float2* wb, int, uint32_t base) const {
#pragma unroll
        for (int i = 0; i < tw_bases(R); ++i) {
            const uint32_t jj = base << (i + sh);
            wb[i] = cmul(hi[jj >> h], lo[jj & ((1u << h) - 1)]);
        }
    }
    // the same tables serving a length N >> s transform
    __device__ __forceinline__ TwLds shifted(uint32_t s) const { return TwLds{lo, hi, h, sh + s}; }
};

// Split of a length-N W_N table into TwLds's two LDS tables: lo has 2^h
// entries, hi has N >> h (only indices below N/2 are ever used).
constexpr int twlds_h(int log2n) { return (log2n + 1) / 2; }
constexpr int twlds_len(int log2n) { return (1 << twlds_h(log2n)) + ((1 << log2n) >> twlds_h(log2n)); }

// Fill the TwLds tables at lds_tw (twlds_len(log2 N) entries) from the W_N
// table; the caller synchronises before use.
template <int N>
__device__ __forceinline__ TwLds twlds_fill(float2* lds_tw, const float2* __restrict__ tw, uint32_t tid,
                                            uint32_t nthreads) {
    constexpr int H = twlds_h(ilog2(N));
    constexpr int NLO = 1 << H, NHI = N >> H;
    for (uint32_t i = tid; i < NLO + NHI; i += nthreads)
        lds_tw[i] = i < NLO ? tw[i] : tw[(i - NLO) << H];
    return TwLds{lds_tw, lds_tw + NLO, (uint32_t)H, 0u};
}

// Stockham pass engine.  d[] holds the thread's E elements; on entry (first
// pass) and on exit (last pass) in the plan's global I/O layout (IoKind).
template <int N, int E, int T, bool INV>
struct Engine {
    static_assert(N % E == 0, "E must divide N");
    static_assert(T * E == N, "T * E must equal N");

    template <int R>
    static constexpr int pair_slots() { return E / (2 * R); }

    // TS: twiddle base source (TwTable, TwHoisted, TwLds); TO = this pass's
    // first entry in a register-held set.
    template <bool PAIR, class TS, int L, int PA, int PB, int TO, bool FIRST, int R, int... Rest>
    __device__ __forceinline__ static void passes(float2 (&d)[E], float2* __restrict__ lds, const TS& ts,
                                                  uint32_t t) {
        constexpr bool LAST = sizeof...(Rest) == 0;
        constexpr bool IO = (FIRST || LAST) && PAIR;
        constexpr int NR = N / R;
        static_assert(E % R == 0, "radix must divide E");
        static_assert(!IO || (E % (2 * R) == 0), "first/last radix must satisfy E >= 2R");
        static_assert(!(FIRST && LAST), "a plan needs at least two passes");

        if constexpr (!FIRST) {
            // gather this pass's inputs from the exchange written by the previous pass
            if constexpr (IO) {
#pragma unroll
                for (int q = 0; q < E / 2; ++q) {
                    uint32_t i = 2 * t + 2 * T * q;
                    if constexpr (PB >= 1) {
                        float4 v = *reinterpret_cast<const float4*>(&lds[pad_idx<PA, PB>(i)]);
                        d[2 * q] = make_float2(v.x, v.y);
                        d[2 * q + 1] = make_float2(v.z, v.w);
                    } else {
                        d[2 * q] = lds[pad_idx<PA, PB>(i)];
                        d[2 * q + 1] = lds[pad_idx<PA, PB>(i + 1)];
                    }
                }
            } else {
#pragma unroll
                for (int i = 0; i < E / R; ++i) {
                    uint32_t b = t + T * i;
#pragma unroll
                    for (int m = 0; m < R; ++m) d[i * R + m] = lds[pad_idx<PA, PB>(b + m * NR)];
                }
            }
        }

        // butterflies
        if constexpr (IO) {
#pragma unroll
            for (int ip = 0; ip < pair_slots<R>(); ++ip) {
#pragma unroll
                for (int dl = 0; dl < 2; ++dl) {
                    float2 v[R];
#pragma unroll
                    for (int m = 0; m < R; ++m) v[m] = d[2 * (ip + m * pair_slots<R>()) + dl];
                    if constexpr (L > 1) {
                        const uint32_t b = 2 * t + dl + 2 * T * ip;
                        float2 wb[tw_bases(R)];
                        ts.template bases<R, TO>(wb, 2 * ip + dl, (b & (L - 1)) * (N / (R * L)));
                        apply_twiddle_powers<R, INV>(v, wb);
                    }
                    Dft<R, INV>::run(v);
#pragma unroll
                    for (int m = 0; m < R; ++m) d[2 * (ip + m * pair_slots<R>()) + dl] = v[m];
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < E / R; ++i) {
                float2 v[R];
#pragma unroll
                for (int m = 0; m < R; ++m) v[m] = d[i * R + m];
                if constexpr (L > 1) {
                    const uint32_t b = t + T * i;
                    float2 wb[tw_bases(R)];
                    ts.template bases<R, TO>(wb, i, (b & (L - 1)) * (N / (R * L)));
                    apply_twiddle_powers<R, INV>(v, wb);
                }
                Dft<R, INV>::run(v);
#pragma unroll
                for (int m = 0; m < R; ++m) d[i * R + m] = v[m];
            }
        }

        if constexpr (!LAST) {
            using P = ExchangePad<R, L, FIRST, PAIR>;
            if constexpr (!FIRST) WGFFT_LDS_SYNC();  // everyone has read the previous exchange
            if constexpr (IO) {
#pragma unroll
                for (int ip = 0; ip < pair_slots<R>(); ++ip)
#pragma unroll
                    for (int dl = 0; dl < 2; ++dl) {
                        uint32_t b = 2 * t + dl + 2 * T * ip;
                        uint32_t k = b & (L - 1), j = b / L;
#pragma unroll
                        for (int m = 0; m < R; ++m)
                            lds[pad_idx<P::A, P::B>(j * R * L + k + m * L)] =
                                d[2 * (ip + m * pair_slots<R>()) + dl];
                    }
            } else {
#pragma unroll
                for (int i = 0; i < E / R; ++i) {
                    uint32_t b = t + T * i;
                    uint32_t k = b & (L - 1), j = b / L;
#pragma unroll
                    for (int m = 0; m < R; ++m) lds[pad_idx<P::A, P::B>(j * R * L + k + m * L)] = d[i * R + m];
                }
            }
            WGFFT_LDS_SYNC();
            passes<PAIR, TS, L * R, P::A, P::B, FIRST ? TO : TO + (E / R) * tw_bases(R), false, Rest...>(d, lds,
                                                                                                         ts, t);
        }
    }

    template <bool PAIR, int L, int TO, bool FIRST, int R, int... Rest>
    __device__ __forceinline__ static void tw_load(float2* twr, const float2* __restrict__ tw, uint32_t t) {
        constexpr bool LAST = sizeof...(Rest) == 0;
        constexpr bool IO = (FIRST || LAST) && PAIR;
        if constexpr (!FIRST) {
#pragma unroll
            for (int j = 0; j < E / R; ++j) {
                const uint32_t b = IO ? 2 * t + (j & 1) + 2 * T * (j >> 1) : t + T * j;
                const uint32_t base = (b & (L - 1)) * (N / (R * L));
#pragma unroll
                for (int i = 0; i < tw_bases(R); ++i) twr[TO + j * tw_bases(R) + i] = tw[base << i];
            }
        }
        if constexpr (!LAST)
            tw_load<PAIR, L * R, FIRST ? TO : TO + (E / R) * tw_bases(R), false, Rest...>(twr, tw, t);
    }

    // Transform d[] (thread t's elements) reading twiddles from the table.
    template <int... Rs>
    __device__ __forceinline__ static void run(float2 (&d)[E], float2* __restrict__ lds,
                                               const float2* __restrict__ tw, uint32_t t,
                                               Radices<Rs...>) {
        static_assert(RadixInfo<Rs...>::product == N, "radices must multiply to N");
        passes<IoKind<E, Rs...>::PAIR, TwTable, 1, 30, 0, 0, true, Rs...>(d, lds, TwTable{tw}, t);
    }

    // Base twiddle powers thread t needs for a plan: tw_bases(R) for each of
    // the E/R butterflies of every pass after the first.  They depend only on
    // t, so a persistent kernel loads them once (load_twiddles) and keeps them
    // in registers; no per-pass table load then has to wait behind the
    // prefetched global loads of the next transform (vmcnt retires in order).
    template <int R0, int... Rest>
    static constexpr int hoisted_count(Radices<R0, Rest...>) {
        return (0 + ... + ((E / Rest) * tw_bases(Rest)));
    }
    template <int... Rs>
    __device__ __forceinline__ static void load_twiddles(float2* twr, const float2* __restrict__ tw, uint32_t t,
                                                         Radices<Rs...>) {
        tw_load<IoKind<E, Rs...>::PAIR, 1, 0, true, Rs...>(twr, tw, t);
    }
    // Transform with any twiddle source (TwHoisted, TwLds, TwTable).
    template <class TS, int... Rs>
    __device__ __forceinline__ static void run_with(float2 (&d)[E], float2* __restrict__ lds, const TS& ts,
                                                    uint32_t t, Radices<Rs...>) {
        static_assert(RadixInfo<Rs...>::product == N, "radices must multiply to N");
        passes<IoKind<E, Rs...>::PAIR, TS, 1, 30, 0, 0, true, Rs...>(d, lds, ts, t);
    }
    // The same in the single layout whatever E is: slot i R + m holds element
    // t + T i + m N/R on entry and on exit (R = first = last radix).
    template <class TS, int... Rs>
    __device__ __forceinline__ static void run_single(float2 (&d)[E], float2* __restrict__ lds, const TS& ts,
                                                      uint32_t t, Radices<Rs...>) {
        static_assert(RadixInfo<Rs...>::product == N, "radices must multiply to N");
        static_assert(RadixInfo<Rs...>::first == RadixInfo<Rs...>::last, "single layout needs first = last radix");
        passes<false, TS, 1, 30, 0, 0, true, Rs...>(d, lds, ts, t);
    }
};

}  // namespace wgfft
