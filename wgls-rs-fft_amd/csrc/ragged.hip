// ragged.hip -- mixed-length batches (placeholder until the persistent
// work-queue kernel lands).
#include "../../include/wgfft.h"
#include "runtime_internal.hpp"

using namespace wgfft;

struct wgfft_ragged_plan {
    int unused;
};

extern "C" int wgfft_ragged_create(wgfft_t*, const uint64_t*, const uint32_t*, uint64_t, wgfft_ragged_t** out) {
    if (out) *out = nullptr;
    return fail(WGFFT_ERR_INVALID_ARGUMENT, "mixed-length batches are not available in this build");
}
extern "C" void wgfft_ragged_destroy(wgfft_ragged_t*) {}
extern "C" int wgfft_ragged_exec_device(wgfft_ragged_t*, const float*, float*, int, int, void*) {
    return fail(WGFFT_ERR_INVALID_ARGUMENT, "mixed-length batches are not available in this build");
}
