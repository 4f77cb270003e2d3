// variants_4096.hip -- LDS-resident variants for N = 4096 (the headline length).
// 16x16x16 with E = 16 uses the single I/O layout (8-byte lanes, two LDS
// exchanges, ~half the registers of E = 32).  Flags (kFlag*): 1 stores
// non-temporally; 2 holds the twiddle base powers in registers across
// transforms; 4 reads them from two short LDS tables; 8 drops the register
// prefetch of the next transform; 512 raises the wave priority around the
// global loads and stores; 2048 prefetches half of the next transform, 2048 +
// 4096 three quarters of it.
#include "kernels.hpp"

WGFFT_REGISTER_BATCHED(4096, 32, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED(4096, 16, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED(4096, 32, 2, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED(4096, 16, 1, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_BATCHED_F(4096, 32, 1, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 1, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 2, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 3, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 4, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 5, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 32, 1, 5, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 5, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 9, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 32, 1, 9, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 32, 2, 9, 16x16x16, 16, 16, 16);

// 1 + 512: non-temporal stores, and the wave's priority raised while it issues
// the next transform's loads and this one's stores (41.1 against 41.5 us).
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 513, 16x16x16, 16, 16, 16);

// 1 + 512 + 2048 (+ 4096): as above with only the first half (three
// quarters) of the next transform prefetched into registers, the rest
// loaded when its turn comes: 113 (126) against 151 VGPRs, so four
// workgroups share a CU and 4096 transforms are exactly four rounds of 1024
// workgroups (three per CU left a last round of 256 transforms at one
// workgroup per CU).  Medians of six alternating timed regions on one box:
// 41.24 us (whole prefetch), 40.88 (half), 40.58 (three quarters; the
// default); 14 of 16 slots need 132 VGPRs (three workgroups per CU).  The
// half prefetch measured 1-3% slower at 256 ... 2048 (G = 2 ... 16
// transforms per workgroup), and with LDS twiddles or without the priority
// at 4096; the three-quarter prefetch with ordinary stores took 41.7 against
// 40.6 us on one buffer pair and 52.0 against 50.1 us over four rotating
// pairs.
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 2561, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_BATCHED_F(4096, 16, 1, 6657, 16x16x16, 16, 16, 16);
