// variants_small_n.hip -- LDS-resident variants for N = 32 .. 512.
// The runtime's default per N is the measured variant; the others are
// autotuner candidates.
#include "kernels.hpp"

WGFFT_REGISTER_BATCHED(32, 32, 256, 16x2, 16, 2);
WGFFT_REGISTER_BATCHED(32, 16, 128, 8x4, 8, 4);
WGFFT_REGISTER_BATCHED(64, 32, 128, 16x4, 16, 4);
WGFFT_REGISTER_BATCHED(64, 16, 64, 8x8, 8, 8);
WGFFT_REGISTER_BATCHED(128, 32, 64, 16x8, 16, 8);
WGFFT_REGISTER_BATCHED(128, 16, 32, 4x4x8, 4, 4, 8);
WGFFT_REGISTER_BATCHED(256, 32, 32, 16x16, 16, 16);
WGFFT_REGISTER_BATCHED(256, 16, 16, 8x4x8, 8, 4, 8);
WGFFT_REGISTER_BATCHED(256, 16, 16, 16x16, 16, 16);
WGFFT_REGISTER_BATCHED(512, 32, 16, 16x16x2, 16, 16, 2);
WGFFT_REGISTER_BATCHED(512, 16, 8, 8x8x8, 8, 8, 8);

// Without the register prefetch (flag 8; 9 adds non-temporal stores): fewer
// registers, more resident waves to hide the loads instead.
WGFFT_REGISTER_BATCHED_F(256, 16, 16, 9, 16x16, 16, 16);
WGFFT_REGISTER_BATCHED_F(256, 32, 32, 9, 16x16, 16, 16);
WGFFT_REGISTER_BATCHED_F(512, 32, 16, 9, 16x16x2, 16, 16, 2);
WGFFT_REGISTER_BATCHED_F(512, 16, 8, 9, 8x8x8, 8, 8, 8);
// One transform per workgroup (G = 1, E = 8: one wavefront).
WGFFT_REGISTER_BATCHED(512, 8, 1, 8x8x8, 8, 8, 8);
