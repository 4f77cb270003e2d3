// variants_adjoint.hip -- fused forward+adjoint and VJP kernels per length.
// For each length and role (fused, VJP, filter) the untuned default is the
// first registration claiming the role, else the first registration (at
// 4096 and 8192 the fastest measured on MI355X); the third argument of
// WGFFT_REGISTER_ADJOINT holds the kAdj* options (1 keep x in registers,
// 2 LDS twiddles).
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
// half-length LDS footprint).  Arguments after E: minimum waves per SIMD,
// elements per step of the closing gw / gx loop, kS2* flags (1 prefetch the
// next transform's input, 2 read x through the caches, 4 and 8 keep the
// thread's x at n and at n + H in registers for the gw accumulation),
// default roles.  Measured at 8192 x 1024 on MI355X (gw reduction
// included): the spectral filter 28-33 us against 40-47 us for the
// one-footprint kernels; the fused forward + adjoint 47.2-47.4 us with x
// kept in registers and read through the caches, 50.0-52.0 us with x kept
// and read non-temporally, 52.7-52.9 us re-reading x through the caches
// and 61-63 us re-reading it after non-temporal reads (the re-read then
// misses the L2); the VJP 37.8-38.3 us with the input prefetch against
// 39.9-41 us.  Measured and not kept for the fused kernel: capping
// registers for three or four workgroups per CU (spills, 57-65 us), the
// input prefetch (58-63 us) and the gw partials in an LDS array instead of
// registers (53.0 us).  At 16384 x 512: fused 66.1 us (x kept) against
// 67.5 us for the one-footprint kernel, VJP 50.0 (prefetch) against 51.0.
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, 16, 0, ::wgfft::kRoleFilter, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, 4, 14, ::wgfft::kRoleFused, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, 16, 1, ::wgfft::kRoleVjp, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, 2, 14, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(8192, 16, 2, 16, 2, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(16384, 16, 2, 16, 0, ::wgfft::kRoleFilter, 16x8x4x16, 16, 8, 4, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(16384, 16, 2, 16, 2, 0, 16x8x4x16, 16, 8, 4, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(16384, 16, 2, 16, 1, ::wgfft::kRoleVjp, 16x8x4x16, 16, 8, 4, 16);
WGFFT_REGISTER_ADJOINT_SPLIT2(16384, 16, 2, 4, 14, ::wgfft::kRoleFused, 16x8x4x16, 16, 8, 4, 16);
// 4096 x 4096: the VJP in split form with the input prefetch takes 69.3 us
// against 81.6-84.8 us for the one-footprint kernels (384 MiB moved: 5.8
// TB/s); as the fused kernel or the filter the split form measured slower
// (fused 80.4-124.6 against 75.7 us, filter 62.3 against 56.9 us; with x
// kept in registers and read through the caches, as at 8192: 77.3-77.7 us
// against 70.5 us for the default, two-wave register budget, 213-248 VGPRs).
WGFFT_REGISTER_ADJOINT_SPLIT2(4096, 16, 2, 16, 1, ::wgfft::kRoleVjp, 16x8x16, 16, 8, 16);
