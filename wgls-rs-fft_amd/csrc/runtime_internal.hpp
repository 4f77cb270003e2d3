// runtime_internal.hpp -- types shared by the host-runtime translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wgfft.h"
#include "kernels.hpp"

namespace wgfft {

extern thread_local std::string g_last_error;
extern const char* kMsgLength;
extern const char* kMsgRagged;

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

struct DevBuf {
    void* p = nullptr;
    uint64_t bytes = 0;
};

// Per-length plan; the counterpart of the reference's SizeCache
// (src/lib.rs:231-249) without the fixed max_batch buffers.
struct Plan {
    enum Kind { kSmall, kBatched, kMultiPass };
    uint64_t n = 0;
    int log2n = 0;
    Kind kind = kSmall;
    const BatchedVariant* var = nullptr;  // kBatched
    // kBatched, untuned plans only: the one-transform-per-workgroup variant
    // used for batches of at most kSmallBatchPerCu transforms per CU
    const BatchedVariant* var_small = nullptr;
    // kBatched, tuned plans only: the windowed entry's variant when the
    // autotuner found another one faster for it than var (nullptr: var)
    const BatchedVariant* var_win = nullptr;
    float2* tw = nullptr;                 // kBatched: W_N^j, j < N
    std::vector<int> pass_log2r;          // kMultiPass
    std::vector<const GlobalPassVariant*> pass_var;
    std::map<int, float2*> tw_r;          // sub-FFT tables W_R^j per radix
    float2* tw_hi = nullptr;
    float2* tw_lo = nullptr;
    uint32_t h = 0;
};

struct Handle;
struct HostPipe;
// Free every mixed-length plan still open on h (called by wgfft_destroy with
// h->mu held).
void ragged_destroy_all(Handle* h);
struct AdjointVariant;
void load_tune_file(Handle* h);
const AdjointVariant* find_adjoint(int log2n, int role);
// The handle's fused kernel for `role` (kRole*): its tuned or pinned choice,
// else the registry default.
const AdjointVariant* pick_adjoint(Handle* h, int log2n, int role);
const std::vector<const AdjointVariant*> adjoint_variants(int log2n);
int fused_adjoint(Handle* h, const AdjointVariant* v, bool fwd, const float* d_x, const float* d_w,
                  const float* d_mask, const float* d_gy, float* d_y, float* d_gx, float* d_gw, float* d_loss,
                  uint64_t n, uint64_t batch, const float2* tw, hipStream_t s);
const char* adjoint_describe(const AdjointVariant* v, char* buf, uint64_t len);
void synth_uniform(float* out, uint64_t nfloats, uint64_t seed);
const BatchedVariant* default_variant(int log2n);
const BatchedVariant* small_batch_variant(int log2n);
constexpr uint64_t kSmallBatchPerCu = 4;
hipError_t launch_add_noise(const float2* src, float2* dst, uint64_t n, uint64_t batch, uint32_t seed,
                            float sigma, hipStream_t s);
std::vector<int> multipass_radices(int log2n);
bool multipass_needs_two_buffers(int passes);
std::vector<std::vector<int>> multipass_splits(int log2n);
const BatchedVariant* find_variant(int log2n, const char* plan, int E, int G, int flags);
std::vector<const GlobalPassVariant*> global_pass_candidates(int log2r, uint64_t L, uint64_t nr);

struct Handle {
    int device = 0;
    int num_cus = 1;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::map<uint64_t, Plan> plans;
    std::map<int, const BatchedVariant*> tuned;  // log2n -> autotuned variant
    std::map<int, const BatchedVariant*> tuned_win;  // log2n -> windowed-mode variant, where it differs
    std::map<int, const AdjointVariant*> tuned_adj;     // fused forward + adjoint
    std::map<int, const AdjointVariant*> tuned_vjp;     // VJP
    std::map<int, const AdjointVariant*> tuned_filter;  // spectral filter
    std::map<std::pair<const BatchedVariant*, int>, int> occ_cache;
    std::map<std::pair<const AdjointVariant*, int>, int> adj_occ_cache;
    std::vector<DevBuf> tables;
    DevBuf io_in, io_out, io_w, io_aux, scratch, scratch2, red, red2, lossbuf;
    uint64_t held = 0;
    uint64_t mp_chunk_bytes = 2ull << 30;  // batch chunk of the multipass path (bounds its scratch)
    // Host-buffer calls of at least host_pipe_min_bytes move through pinned
    // staging in chunks of whole transforms of a quarter of the call, between
    // host_chunk_min_bytes and host_chunk_bytes, pipelined over two copy
    // streams; smaller calls stage as one chunk on the handle's stream
    // (transforms larger than a chunk are copied directly).  Measured: 4 MB
    // calls through 1 MiB chunks 0.31 against 0.39 ms as one; 2 MB calls as
    // one staged chunk 0.16 against 0.21 ms copied directly from pageable
    // memory.
    uint64_t host_pipe_min_bytes = 4ull << 20;   // WGFFT_HOST_PIPE_MIN_BYTES
    uint64_t host_chunk_min_bytes = 1ull << 20;  // WGFFT_HOST_CHUNK_MIN_BYTES
    uint64_t host_chunk_bytes = 32ull << 20;  // the largest chunk
    bool host_chunk_fixed = false;            // WGFFT_HOST_CHUNK_BYTES given: every chunk that size
    HostPipe* pipe = nullptr;
    // mixed-length plans created on this handle and not yet destroyed; they
    // use its device and lock, so wgfft_destroy frees them with it
    std::set<wgfft_ragged_t*> ragged_plans;
    // Device calls enqueued on different streams share the scratch buffers
    // above, and two calls must never use them at the same time on the
    // device.  While every use is on one stream, that stream orders them and
    // nothing is recorded (an event record costs the GPU ~4.5 us of idle
    // time before the next launch).  The first stream-ordered use on a
    // second stream while another stream's use may still be pending waits
    // for the device once; from then on every use records scratch_ev, and a
    // use on another stream waits for it.  Host-blocking calls wait for the
    // device up front when needed (scratch_quiesce) and settle their own
    // stream when they have synchronized it.
    hipEvent_t scratch_ev = nullptr;
    hipStream_t scratch_stream = nullptr;  // last use; may be pending if scratch_pending
    bool scratch_pending = false;
    bool scratch_recorded = false;  // scratch_ev follows the last use
    bool scratch_events = false;    // record after every use
    hipError_t scratch_begin(hipStream_t s);
    void scratch_end(hipStream_t s);
    hipError_t scratch_quiesce();
    void scratch_settled(hipStream_t s);

    int ensure(DevBuf& b, uint64_t bytes);
    void release(DevBuf& b);
    int upload_table(const std::vector<float2>& host, float2** dev);
    int get_plan(uint64_t n, Plan** out);
    int set_split(Plan* p, const std::vector<int>& log2r);
    // workgroups per CU of a variant (kept in occ_cache when cache is set)
    int blocks_per_cu(const BatchedVariant* v, Mode m, bool cache = true);
    // workgroups per CU of a fused kernel, kept in adj_occ_cache (the
    // occupancy query costs microseconds per call): kind 0 the VJP, 1 the
    // fused forward + adjoint, 2 the filter, 3 the filter with noise, 4 the
    // fused forward + adjoint that also stores Y
    int adjoint_blocks_per_cu(const AdjointVariant* v, int kind);
    int run(Plan* p, Mode mode, const float2* in, float2* out, const float* win, uint64_t batch, float scale,
            hipStream_t s, const BatchedVariant* force = nullptr, float2* fixed = nullptr, uint64_t fixed_bytes = 0,
            uint64_t idist = 0, uint64_t odist = 0);
    int run_passes(Plan* p, Mode mode, const float2* in, float2* out, const float* win, uint64_t batch, float scale,
                   hipStream_t s, float2* buf_a, float2* buf_b, uint64_t chunk);
    // host calls of transforms up to a staging chunk go through pinned
    // staging (run_host); larger transforms are copied directly
    bool host_staged(uint64_t n) const { return n * sizeof(float2) <= host_chunk_bytes; }
    // Host in -> host out through the staging pipeline; run_chunk(d_in, d_out,
    // transforms) enqueues one chunk's transform on `stream`.  Calls below
    // host_pipe_min_bytes stage as one chunk on `stream`.  With in_rows /
    // out_rows the signals are separate buffers, gathered and scattered by the
    // staging copies (in / out unused).
    int run_host(const float* in, float* out, uint64_t n, uint64_t batch,
                 const std::function<int(const float2*, float2*, uint64_t)>& run_chunk,
                 const float* const* in_rows = nullptr, float* const* out_rows = nullptr);
    void free_pipe();
};

}  // namespace wgfft

namespace wgfft {
// Orders one call's use of the handle's scratch buffers after the previous
// call's (on whatever stream that ran) and notes it as the last use when the
// call returns, on success and on error paths alike (recording an event
// after it once more than one stream has used the scratch).
struct ScratchUse {
    Handle* h;
    hipStream_t s;
    hipError_t e;
    ScratchUse(Handle* h_, hipStream_t s_) : h(h_), s(s_), e(h_->scratch_begin(s_)) {}
    ~ScratchUse() { h->scratch_end(s); }
    ScratchUse(const ScratchUse&) = delete;
    ScratchUse& operator=(const ScratchUse&) = delete;
};
}  // namespace wgfft

struct wgfft_handle : wgfft::Handle {};
