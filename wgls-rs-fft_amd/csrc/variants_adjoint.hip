// variants_adjoint.hip -- fused forward+adjoint and VJP kernels per length.
// For each length the first registration is the default; KEEPX = 1 keeps x
// in registers between the forward transform and the gw accumulation.
#include "kernels_adjoint.hpp"

WGFFT_REGISTER_ADJOINT(512, 32, 0, 16x16x2, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(1024, 32, 0, 16x16x4, 16, 16, 4);
WGFFT_REGISTER_ADJOINT(2048, 32, 0, 16x16x8, 16, 16, 8);
WGFFT_REGISTER_ADJOINT(4096, 32, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 32, 1, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(4096, 16, 0, 8x8x8x8, 8, 8, 8, 8);
WGFFT_REGISTER_ADJOINT(4096, 16, 0, 16x16x16, 16, 16, 16);
WGFFT_REGISTER_ADJOINT(8192, 32, 0, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(8192, 32, 1, 16x16x16x2, 16, 16, 16, 2);
WGFFT_REGISTER_ADJOINT(8192, 16, 0, 16x2x16x16, 16, 2, 16, 16);
WGFFT_REGISTER_ADJOINT(16384, 32, 0, 16x16x16x4, 16, 16, 16, 4);
