// variants_global.hip -- sub-FFT variants for the passes of N > 16384:
// radix R as an LDS sub-FFT over G butterflies per workgroup with E
// elements per thread, with (PF = 1) or without the next group's tile
// prefetched into registers.  The runtime's default per length is the
// measured split and pass variants (first registration per R beyond 2^26);
// wgfft_autotune times the others.  E = 16 halves the registers of the
// sub-FFT (more workgroups per CU), E = 32 halves its LDS passes.
#include "kernels.hpp"

WGFFT_REGISTER_GLOBAL_PASS(128, 16, 16, 0, 4x4x8, 4, 4, 8);
WGFFT_REGISTER_GLOBAL_PASS(128, 32, 16, 0, 16x8, 16, 8);
WGFFT_REGISTER_GLOBAL_PASS(256, 16, 16, 0, 8x4x8, 8, 4, 8);
WGFFT_REGISTER_GLOBAL_PASS(256, 32, 16, 0, 16x16, 16, 16);
WGFFT_REGISTER_GLOBAL_PASS(512, 16, 8, 0, 8x8x8, 8, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS(512, 32, 16, 0, 16x16x2, 16, 16, 2);
WGFFT_REGISTER_GLOBAL_PASS(1024, 16, 4, 0, 8x16x8, 8, 16, 8);
WGFFT_REGISTER_GLOBAL_PASS(1024, 32, 8, 0, 16x16x4, 16, 16, 4);
WGFFT_REGISTER_GLOBAL_PASS(2048, 16, 2, 0, 8x4x8x8, 8, 4, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS(2048, 16, 4, 0, 8x4x8x8, 8, 4, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS(2048, 32, 8, 0, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_GLOBAL_PASS(2048, 32, 4, 1, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_GLOBAL_PASS(4096, 16, 2, 0, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS(4096, 32, 4, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_GLOBAL_PASS(8192, 16, 2, 0, 8x16x8x8, 8, 16, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS(8192, 32, 2, 0, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_GLOBAL_PASS(8192, 32, 2, 1, 16x16x16x2, 16, 16, 16, 2);

// Wider column groups for the small-radix splits (256-byte runs).
WGFFT_REGISTER_GLOBAL_PASS(256, 16, 32, 0, 8x4x8, 8, 4, 8);
WGFFT_REGISTER_GLOBAL_PASS(512, 16, 16, 0, 8x8x8, 8, 8, 8);

// Direct tile I/O (no LDS staging; single-layout sub-FFTs, first radix =
// last radix), without and with the register prefetch.
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(128, 16, 16, 0, 4x8x4, 4, 8, 4);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(128, 16, 32, 0, 4x8x4, 4, 8, 4);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(256, 16, 16, 0, 16x16, 16, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(256, 16, 16, 1, 16x16, 16, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(256, 16, 32, 0, 16x16, 16, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(512, 16, 16, 0, 8x8x8, 8, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(512, 16, 8, 1, 8x8x8, 8, 8, 8);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(1024, 16, 8, 0, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 16, 4, 0, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 16, 8, 0, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 16, 4, 1, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 16, 8, 1, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 32, 8, 0, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(2048, 32, 8, 1, 16x8x16, 16, 8, 16);
WGFFT_REGISTER_GLOBAL_PASS_DIRECT(4096, 16, 2, 0, 16x16x16, 16, 16, 16);

// 16-column direct tiles (128-byte runs) with the LDS of 8 columns: the
// sub-FFTs run in two rounds.  (Four rounds of 4 columns at E = 16 were
// measured slower at radix 2048, and so were two rounds at E = 16: 1024
// threads capped at 128 VGPRs spill 18, 573 against 532 us at 2^22 x 16.)
WGFFT_REGISTER_GLOBAL_PASS_WIDE(1024, 16, 2, 16x4x16, 16, 4, 16);
WGFFT_REGISTER_GLOBAL_PASS_WIDE(2048, 32, 2, 16x8x16, 16, 8, 16);
