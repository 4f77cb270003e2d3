// kernels_small.hip -- lengths 1..16: one thread per transform, the whole
// DFT in registers (N = 1 is the identity, as in the reference where
// log2 N = 0 gives no stage at all, src/lib.rs:726-732).
#include "kernels.hpp"

namespace wgfft {

template <int N, int MODE>
__global__ void __launch_bounds__(256)
k_small(const float2* __restrict__ in, float2* __restrict__ out, const float* __restrict__ win, uint64_t batch,
        float scale, uint64_t idist, uint64_t odist) {
    constexpr bool INV = MODE == kInverse;
    const uint64_t tr = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (tr >= batch) return;
    float2 v[N];
    if constexpr (N == 1) {
        v[0] = in[tr * idist];
    } else {
        const float4* s4 = reinterpret_cast<const float4*>(in + tr * idist);
#pragma unroll
        for (int q = 0; q < N / 2; ++q) {
            float4 a = s4[q];
            v[2 * q] = make_float2(a.x, a.y);
            v[2 * q + 1] = make_float2(a.z, a.w);
        }
    }
    if constexpr (MODE == kWindowed) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = cscale(v[i], win[i]);
    }
    if constexpr (N > 1) Dft<N, INV>::run(v);
    if constexpr (INV) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = cscale(v[i], scale);
    }
    if constexpr (N == 1) {
        out[tr * odist] = v[0];
    } else {
        float4* d4 = reinterpret_cast<float4*>(out + tr * odist);
#pragma unroll
        for (int q = 0; q < N / 2; ++q) d4[q] = make_float4(v[2 * q].x, v[2 * q].y, v[2 * q + 1].x, v[2 * q + 1].y);
    }
}

template <int N>
static hipError_t go_small(Mode m, const BatchedArgs& a) {
    const unsigned grid = (unsigned)((a.batch + 255) / 256);
    const uint64_t idist = a.idist ? a.idist : (uint64_t)N, odist = a.odist ? a.odist : (uint64_t)N;
    switch (m) {
        case kForward:
            hipLaunchKernelGGL((k_small<N, kForward>), dim3(grid), dim3(256), 0, a.stream, a.in, a.out, a.win, a.batch,
                               a.scale, idist, odist);
            break;
        case kInverse:
            hipLaunchKernelGGL((k_small<N, kInverse>), dim3(grid), dim3(256), 0, a.stream, a.in, a.out, a.win, a.batch,
                               a.scale, idist, odist);
            break;
        default:
            hipLaunchKernelGGL((k_small<N, kWindowed>), dim3(grid), dim3(256), 0, a.stream, a.in, a.out, a.win,
                               a.batch, a.scale, idist, odist);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_small(Mode mode, int log2n, const BatchedArgs& a) {
    switch (log2n) {
        case 0: return go_small<1>(mode, a);
        case 1: return go_small<2>(mode, a);
        case 2: return go_small<4>(mode, a);
        case 3: return go_small<8>(mode, a);
        case 4: return go_small<16>(mode, a);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace wgfft
