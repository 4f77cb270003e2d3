// variants_mid_n.hip -- LDS-resident variants for N = 1024, 2048.
#include "kernels.hpp"

WGFFT_REGISTER_BATCHED(1024, 32, 8, 16x16x4, 16, 16, 4);
WGFFT_REGISTER_BATCHED(1024, 32, 8, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_BATCHED(1024, 16, 4, 8x16x8, 8, 16, 8);
WGFFT_REGISTER_BATCHED(1024, 16, 4, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_BATCHED(2048, 32, 4, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_BATCHED(2048, 32, 4, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_BATCHED(2048, 16, 2, 8x4x8x8, 8, 4, 8, 8);
WGFFT_REGISTER_BATCHED(2048, 16, 2, 16x8x16, 16, 8, 16);

// Without the register prefetch (flag 8; 9 adds non-temporal stores).
WGFFT_REGISTER_BATCHED_F(1024, 32, 8, 9, 16x16x4, 16, 16, 4);
WGFFT_REGISTER_BATCHED_F(1024, 16, 4, 9, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_BATCHED_F(2048, 32, 4, 9, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_BATCHED_F(2048, 16, 2, 9, 16x8x16, 16, 8, 16);
// Flag 512: the wave priority raised while it issues global loads and stores
// (1024: 40.3-40.5 against 40.7-40.9 us, 2048: 40.8-40.9 against 41.3-41.5 us
// per 2^24 samples; no gain at 256 and 512).
WGFFT_REGISTER_BATCHED_F(1024, 16, 4, 512, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_BATCHED_F(2048, 16, 2, 512, 16x8x16, 16, 8, 16);
// One transform per workgroup (G = 1): a small batch spread over as many
// CUs as it has transforms.
WGFFT_REGISTER_BATCHED_F(1024, 16, 1, 512, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_BATCHED_F(2048, 16, 1, 512, 16x8x16, 16, 8, 16);
