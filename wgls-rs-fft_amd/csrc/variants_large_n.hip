// variants_large_n.hip -- LDS-resident variants for N = 8192, 16384.
// Flag 8 drops the register prefetch of the next transform: the 128 KB
// transforms of N = 16384 leave one workgroup per CU, where the prefetch
// registers otherwise spill.
#include "kernels.hpp"

WGFFT_REGISTER_BATCHED(8192, 32, 1, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_BATCHED(8192, 32, 1, 8x16x16x4, 8, 16, 16, 4);
WGFFT_REGISTER_BATCHED(8192, 16, 1, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_BATCHED_F(8192, 16, 1, 8, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_BATCHED_F(8192, 32, 1, 8, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_BATCHED(16384, 32, 1, 16x16x16x4, 16, 16, 16, 4);
WGFFT_REGISTER_BATCHED(16384, 32, 1, 16x16x4x16, 16, 16, 4, 16);
WGFFT_REGISTER_BATCHED(16384, 16, 1, 16x4x16x16, 16, 4, 16, 16);
WGFFT_REGISTER_BATCHED_F(16384, 16, 1, 8, 16x4x16x16, 16, 4, 16, 16);
WGFFT_REGISTER_BATCHED_F(16384, 16, 1, 9, 16x4x16x16, 16, 4, 16, 16);
WGFFT_REGISTER_BATCHED_F(16384, 32, 1, 8, 16x16x16x4, 16, 16, 16, 4);
WGFFT_REGISTER_BATCHED_F(16384, 32, 1, 9, 16x16x16x4, 16, 16, 16, 4);

// N = 16384 (8192) as a radix-2 stage and two 8192-point (4096-point)
// halves through one LDS footprint of the half length: 74 KB (36-38 KB), so
// two (four) workgroups share a CU; PLAN is the split of the half length.
// At 8192 x 1024 all transforms are then in flight at once (21.4 us against
// 26.5 us); at 8192 x 2048 the single-footprint kernel is 3% ahead.  (The
// same split of 4096 is slower.)
WGFFT_REGISTER_SPLIT2(16384, 16, 8x16x8x8, 8, 16, 8, 8);
WGFFT_REGISTER_SPLIT2(16384, 16, 8x8x16x8, 8, 8, 16, 8);
WGFFT_REGISTER_SPLIT2(16384, 16, 16x8x4x16, 16, 8, 4, 16);
WGFFT_REGISTER_SPLIT2(8192, 16, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_SPLIT2(8192, 16, 8x8x8x8, 8, 8, 8, 8);
// The same with the input read through the caches (kFlagCachedLoad = 8192):
// a re-read working set near the 256 MB memory-side cache's size is served
// from it (16384 x 1024: 48.7 against 52.1 us), a smaller one is not helped
// (8192 x 1024: 23.6 against 21.7 us); the autotuner picks per shape.
WGFFT_REGISTER_SPLIT2_F(16384, 16, 8192, 16x8x4x16, 16, 8, 4, 16);
WGFFT_REGISTER_SPLIT2_F(16384, 16, 8192, 8x8x16x8, 8, 8, 16, 8);
WGFFT_REGISTER_SPLIT2_F(8192, 16, 8192, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_SPLIT2_F(8192, 16, 8192, 8x8x8x8, 8, 8, 8, 8);
