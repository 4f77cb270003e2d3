// variants_adjoint.hip -- fused forward+adjoint and VJP kernels per length.
// For each length the first registration is the default (at 4096 and 8192
// the fastest measured on MI355X); the third argument holds the kAdj*
// options (1 keep x in registers, 2 LDS twiddles).
#include "kernels_adjoint.hpp"

WGFFT_REGISTER_ADJOINT(512, 32, 0, 16x16x2, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(1024, 32, 0, 16x16x4, 16, 16, 4);
WGFFT_REGISTER_ADJOINT(2048, 32, 0, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 3, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 32, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 32, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 16, 0, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 16, 1, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 16, 2, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 3, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 2, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(8192, 16, 3, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_ADJOINT(8192, 32, 0, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(8192, 32, 1, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(8192, 16, 0, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_ADJOINT(8192, 16, 2, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_ADJOINT(8192, 32, 2, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(8192, 16, 2, 8x16x8x8, 8, 16, 8, 8);
WGFFT_REGISTER_ADJOINT(8192, 16, 3, 8x16x8x8, 8, 16, 8, 8);
WGFFT_REGISTER_ADJOINT(16384, 32, 0, 16x16x16x4, 16, 16, 16, 4);

// Split form (a radix-2 stage and two half-length transforms through a
// half-length LDS footprint), the untuned default of the spectral filter at
// both lengths (8192 x 1024: 31.9 against 47.5 us; 16384 x 512: 39.6
// against 48.7 us) and of the VJP at 8192 (kernel 35.6 us).  As the fused
// forward + adjoint it holds 188 VGPRs for the gw partials (two workgroups
// per CU) and measured no faster than the first registration; three or four
// workgroups per CU (capped registers, or per-transform partial rows
// instead of the gw registers) measured slower still (77-125 us).
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, ::wgfft::kRoleVjp | ::wgfft::kRoleFilter, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(16384, 16, 2, ::wgfft::kRoleFilter, 16x8x4x16, 16, 8, 4, 16);
