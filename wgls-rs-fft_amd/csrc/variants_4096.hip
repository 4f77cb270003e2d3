// variants_4096.hip -- LDS-resident variants for N = 4096 (the headline length).
// 16x16x16 with E = 16 uses the single I/O layout (8-byte lanes, two LDS
// exchanges, ~half the registers of E = 32).  Flags (kFlag*): 1 stores
// non-temporally; 2 holds the twiddle base powers in registers across
// transforms; 4 reads them from two short LDS tables; 512 raises the wave
// priority around the global loads and stores.
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
