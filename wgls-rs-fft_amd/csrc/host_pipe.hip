// host_pipe.hip -- the host-buffer entry points' copy pipeline.
//
// The reference's GpuFft::fft/ifft take host vectors and return host vectors
// (src/lib.rs:540-568: flatten, upload, run, read back).  Copying a caller's
// pageable buffer straight to the device runs at the HIP runtime's staged rate
// (measured on MI355X: 30 GB/s up, 18 GB/s down), and the transform then waits
// for the whole upload.  Here the batch moves in chunks of whole transforms
// through four pinned staging slots per direction: host threads copy chunk
// i into pinned memory and chunk i - 3 out of it, while the DMA engines move
// the chunks between (57 GB/s per direction from pinned memory) on two copy
// streams and the transforms run on the handle's stream.  4096 x 4096 host
// to host: 4.3-4.8 ms with 32 MiB chunks (two slots of 64 MiB: 5.4-7.0 ms;
// 16 MiB chunks: a similar median with outliers to 11 ms, the copy threads
// waking twice per chunk).  Calls below 4 MB stage as one chunk on the
// handle's stream.  The signals may also be separate buffers, one pointer per
// signal (wgfft_exec_host_rows, the reference's Vec<Vec<Complex<f32>>>): the
// copy threads then gather them into the staging slots and scatter the
// results, so the batch is never flattened.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wgfft.h"
#include "runtime_internal.hpp"

namespace wgfft {

// A fixed set of threads that split one copy between them; the calling
// thread takes a share too and returns when every share is done.  A copy is
// `bytes` of one contiguous region, or of rows of row_bytes each gathered
// from (src_rows) or scattered to (dst_rows) separate buffers: the byte range
// is split evenly whatever the row size.
class CopyPool {
public:
    explicit CopyPool(int helpers) {
        try {
            for (int i = 0; i < helpers; ++i) threads_.emplace_back([this, i] { work(i + 1); });
        } catch (...) {
            stop_all();  // the threads already started are joined before the exception leaves
            throw;
        }
    }
    ~CopyPool() { stop_all(); }
    void copy(void* dst, const void* src, uint64_t bytes) { run((char*)dst, (const char*)src, nullptr, nullptr, bytes, 0); }
    void gather(void* dst, const float* const* src_rows, uint64_t row_bytes, uint64_t rows) {
        run((char*)dst, nullptr, (const char* const*)src_rows, nullptr, row_bytes * rows, row_bytes);
    }
    void scatter(float* const* dst_rows, const void* src, uint64_t row_bytes, uint64_t rows) {
        run(nullptr, (const char*)src, nullptr, (char* const*)dst_rows, row_bytes * rows, row_bytes);
    }

private:
    void run(char* dst, const char* src, const char* const* src_rows, char* const* dst_rows, uint64_t bytes,
             uint64_t row_bytes) {
        // a small copy stays on the calling thread: waking the helpers costs
        // more than they save below ~256 KB
        const int parts = bytes < kInlineBytes ? 1 : (int)threads_.size() + 1;
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = dst;
            src_ = src;
            src_rows_ = src_rows;
            dst_rows_ = dst_rows;
            row_bytes_ = row_bytes;
            bytes_ = bytes;
            if (parts > 1) {
                pending_ = parts - 1;
                ++gen_;
            }
        }
        if (parts > 1) cv_.notify_all();
        share(0, parts);
        if (parts == 1) return;
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
    }
    void stop_all() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
        threads_.clear();
    }
    void share(int k, int parts) {
        // 4 KB-aligned shares of the byte range
        const uint64_t per = ((bytes_ / parts) + 4095) & ~4095ull;
        const uint64_t b = std::min<uint64_t>(bytes_, per * k), e = std::min<uint64_t>(bytes_, b + per);
        if (e <= b) return;
        if (!src_rows_ && !dst_rows_) {
            std::memcpy(dst_ + b, src_ + b, e - b);
            return;
        }
        // the rows overlapping [b, e), each from its own buffer
        for (uint64_t r = b / row_bytes_, pos = b; pos < e; ++r) {
            const uint64_t off = pos - r * row_bytes_;
            const uint64_t len = std::min(e, (r + 1) * row_bytes_) - pos;
            if (src_rows_)
                std::memcpy(dst_ + pos, src_rows_[r] + off, len);
            else
                std::memcpy(dst_rows_[r] + off, src_ + pos, len);
            pos += len;
        }
    }
    void work(int k) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            share(k, (int)threads_.size() + 1);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    static constexpr uint64_t kInlineBytes = 256ull << 10;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    bool stop_ = false;
    uint64_t gen_ = 0;
    int pending_ = 0;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    const char* const* src_rows_ = nullptr;
    char* const* dst_rows_ = nullptr;
    uint64_t row_bytes_ = 0;
    uint64_t bytes_ = 0;
};

struct HostPipe {
    static constexpr int kMaxSlots = 8;
    int slots = 4;  // staging slots per direction (WGFFT_HOST_SLOTS, 2 .. kMaxSlots)
    uint64_t slot_bytes = 0;
    void* pin_in[kMaxSlots] = {};
    void* pin_out[kMaxSlots] = {};
    hipStream_t s_in = nullptr, s_out = nullptr;
    hipEvent_t ev_in[kMaxSlots] = {}, ev_run[kMaxSlots] = {}, ev_out[kMaxSlots] = {};
    CopyPool* pool = nullptr;
};

void Handle::free_pipe() {
    if (!pipe) return;
    HostPipe* hp = pipe;
    for (hipStream_t s : {hp->s_in, hp->s_out})
        if (s) {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    for (int i = 0; i < HostPipe::kMaxSlots; ++i) {
        for (hipEvent_t ev : {hp->ev_in[i], hp->ev_run[i], hp->ev_out[i]})
            if (ev) hipEventDestroy(ev);
        if (hp->pin_in[i]) hipHostFree(hp->pin_in[i]);
        if (hp->pin_out[i]) hipHostFree(hp->pin_out[i]);
    }
    delete hp->pool;
    delete hp;
    pipe = nullptr;
}

// Streams, events, the copy pool and pinned slots of at least `bytes` each.
static int pipe_ready(Handle* h, uint64_t bytes) {
    hipError_t e = hipSuccess;
    if (!h->pipe) {
        HostPipe* hp = new (std::nothrow) HostPipe();
        if (!hp) return fail(WGFFT_ERR_OUT_OF_MEMORY, "host pipeline allocation failed");
        h->pipe = hp;
        if (const char* env = std::getenv("WGFFT_HOST_SLOTS"))
            hp->slots = std::max(2, std::min(HostPipe::kMaxSlots, std::atoi(env)));
        if ((e = hipStreamCreateWithFlags(&hp->s_in, hipStreamNonBlocking)) == hipSuccess)
            e = hipStreamCreateWithFlags(&hp->s_out, hipStreamNonBlocking);
        for (int i = 0; i < hp->slots && e == hipSuccess; ++i)
            for (hipEvent_t* ev : {&hp->ev_in[i], &hp->ev_run[i], &hp->ev_out[i]})
                if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            h->free_pipe();  // no half-built pipeline is kept
            return hip_fail(e, "host pipeline streams and events");
        }
        // copy threads (the caller and helpers): WGFFT_HOST_COPY_THREADS, else
        // 8, at most the CPUs this process may run on (8 threads copy at
        // ~89 GB/s on the MI355X host, 1 thread at 27 GB/s)
        int cpus = (int)std::thread::hardware_concurrency();
        cpu_set_t set;
        if (sched_getaffinity(0, sizeof set, &set) == 0) cpus = CPU_COUNT(&set);
        int want = 8;
        if (const char* env = std::getenv("WGFFT_HOST_COPY_THREADS")) want = std::max(1, std::atoi(env));
        const int threads = std::max(1, std::min(want, cpus));
        try {  // nothing may throw across the C-ABI
            hp->pool = new CopyPool(threads - 1);
        } catch (...) {
            hp->pool = nullptr;
        }
        if (!hp->pool) {
            h->free_pipe();
            return fail(WGFFT_ERR_OUT_OF_MEMORY, "copy pool threads could not be started");
        }
    }
    HostPipe* hp = h->pipe;
    if (hp->slot_bytes >= bytes) return WGFFT_OK;
    for (int i = 0; i < hp->slots; ++i) {
        if (hp->pin_in[i]) hipHostFree(hp->pin_in[i]);
        if (hp->pin_out[i]) hipHostFree(hp->pin_out[i]);
        hp->pin_in[i] = hp->pin_out[i] = nullptr;
    }
    hp->slot_bytes = 0;
    for (int i = 0; i < hp->slots; ++i)
        if ((e = hipHostMalloc(&hp->pin_in[i], bytes, hipHostMallocDefault)) != hipSuccess ||
            (e = hipHostMalloc(&hp->pin_out[i], bytes, hipHostMallocDefault)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(WGFFT_ERR_OUT_OF_MEMORY, "hipHostMalloc(" + std::to_string(bytes) + ") failed: " +
                                                          hipGetErrorString(e));
        }
    hp->slot_bytes = bytes;
    return WGFFT_OK;
}

int Handle::run_host(const float* in, float* out, uint64_t n, uint64_t batch,
                     const std::function<int(const float2*, float2*, uint64_t)>& run_chunk,
                     const float* const* in_rows, float* const* out_rows) {
    const uint64_t tb = n * sizeof(float2);
    int rc;
    if (tb * batch < host_pipe_min_bytes) {
        // one chunk on the handle's stream alone: below the pipelining
        // threshold the copy streams' event hand-offs cost more than the
        // overlap saves
        const uint64_t total = tb * batch;
        if ((rc = pipe_ready(this, total))) return rc;
        HostPipe* hp = pipe;
        if ((rc = ensure(io_in, total)) || (rc = ensure(io_out, total))) return rc;
        if (in_rows)
            hp->pool->gather(hp->pin_in[0], in_rows, tb, batch);
        else
            hp->pool->copy(hp->pin_in[0], in, total);
        hipError_t e = hipMemcpyAsync(io_in.p, hp->pin_in[0], total, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(H2D)");
        if ((rc = run_chunk((const float2*)io_in.p, (float2*)io_out.p, batch))) {
            hipStreamSynchronize(stream);
            return rc;
        }
        if ((e = hipMemcpyAsync(hp->pin_out[0], io_out.p, total, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess) {
            hipStreamSynchronize(stream);
            return hip_fail(e, "host staging");
        }
        if (out_rows)
            hp->pool->scatter(out_rows, hp->pin_out[0], tb, batch);
        else
            hp->pool->copy(out, hp->pin_out[0], total);
        return WGFFT_OK;
    }
    // a quarter of the call, 1 .. 32 MiB, so that a call of 4-128 MB has
    // chunks to overlap (32 MB host to host: 1.37 ms in 8 MiB chunks against
    // 2.06 ms as one; 128 MB: 4.4 ms in 32 MiB chunks against 5.8 in 8 MiB)
    uint64_t chunk_bytes = host_chunk_bytes;
    if (!host_chunk_fixed) chunk_bytes = std::min(host_chunk_bytes, std::max(host_chunk_min_bytes, tb * batch / 4));
    const uint64_t cb = std::max<uint64_t>(1, chunk_bytes / tb);  // transforms per chunk
    const uint64_t slot = cb * tb;
    if ((rc = pipe_ready(this, slot))) return rc;
    HostPipe* hp = pipe;
    const int K = hp->slots;
    if ((rc = ensure(io_in, K * slot)) || (rc = ensure(io_out, K * slot))) return rc;
    const uint64_t chunks = (batch + cb - 1) / cb;
    auto bytes_of = [&](uint64_t i) { return std::min(cb, batch - i * cb) * tb; };
    auto dev_in = [&](int s) { return (float2*)((char*)io_in.p + s * slot); };
    auto dev_out = [&](int s) { return (float2*)((char*)io_out.p + s * slot); };
    hipError_t e = hipSuccess;
    auto drain = [&]() {  // leave nothing in flight on an error path
        hipStreamSynchronize(hp->s_in);
        hipStreamSynchronize(stream);
        hipStreamSynchronize(hp->s_out);
    };
    // Chunk i goes through slot i % K; the host copies chunk i in, enqueues
    // its upload, transform and read-back, then copies chunk i - (K - 1)
    // out, so K - 1 chunks' DMA and transforms are in flight while it copies
    // (with one, each iteration waited ~0.57 ms for the previous chunk's
    // upload + transform + read-back of 16 MB).  Slot s is reused by chunk i
    // only after chunk i - K's upload, transform and read-back (events) and
    // its copy out (iteration i - 1 at the latest).
    const uint64_t lag = (uint64_t)K - 1;
    for (uint64_t i = 0; i < chunks + lag; ++i) {
        if (i < chunks) {
            const int s = (int)(i % K);
            const uint64_t bytes = bytes_of(i);
            // pinned slot s was last read by chunk i - K's upload
            if (i >= (uint64_t)K && (e = hipEventSynchronize(hp->ev_in[s])) != hipSuccess) break;
            if (in_rows)
                hp->pool->gather(hp->pin_in[s], in_rows + i * cb, tb, bytes / tb);
            else
                hp->pool->copy(hp->pin_in[s], (const char*)in + i * cb * tb, bytes);
            // device slot s: chunk i - K's transform has read it and its result has been read back
            if (i >= (uint64_t)K && (e = hipStreamWaitEvent(hp->s_in, hp->ev_run[s], 0)) != hipSuccess) break;
            if ((e = hipMemcpyAsync(dev_in(s), hp->pin_in[s], bytes, hipMemcpyHostToDevice, hp->s_in)) != hipSuccess ||
                (e = hipEventRecord(hp->ev_in[s], hp->s_in)) != hipSuccess ||
                (e = hipStreamWaitEvent(stream, hp->ev_in[s], 0)) != hipSuccess)
                break;
            if (i >= (uint64_t)K && (e = hipStreamWaitEvent(stream, hp->ev_out[s], 0)) != hipSuccess) break;
            if ((rc = run_chunk(dev_in(s), dev_out(s), bytes / tb))) {
                drain();
                return rc;
            }
            if ((e = hipEventRecord(hp->ev_run[s], stream)) != hipSuccess ||
                (e = hipStreamWaitEvent(hp->s_out, hp->ev_run[s], 0)) != hipSuccess ||
                (e = hipMemcpyAsync(hp->pin_out[s], dev_out(s), bytes, hipMemcpyDeviceToHost, hp->s_out)) != hipSuccess ||
                (e = hipEventRecord(hp->ev_out[s], hp->s_out)) != hipSuccess)
                break;
        }
        if (i >= lag) {  // chunk i - lag back to the caller's buffer
            const uint64_t j = i - lag;
            const int s = (int)(j % K);
            if ((e = hipEventSynchronize(hp->ev_out[s])) != hipSuccess) break;
            if (out_rows)
                hp->pool->scatter(out_rows + j * cb, hp->pin_out[s], tb, bytes_of(j) / tb);
            else
                hp->pool->copy((char*)out + j * cb * tb, hp->pin_out[s], bytes_of(j));
        }
    }
    if (e != hipSuccess) {
        drain();
        return hip_fail(e, "host pipeline");
    }
    return WGFFT_OK;
}

}  // namespace wgfft
