// pfa_hostio.cpp -- staged host <-> device copies for the host-buffer entry
// points.  A caller's numpy arrays are pageable, and pinning them in place is
// not available, so large copies go through a ring of pinned staging buffers:
// a small pool of host threads copies chunk c between the caller's buffer and
// a staging buffer while the DMA engine moves chunk c - 1 (H2D) or c + 1
// (D2H) on a copy stream of the context's device.  The compute stream waits
// on the copy stream with an event, so kernels still see the data in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "pfa_internal.hpp"

namespace {

constexpr size_t STAGE_BYTES = 16u << 20;   // one chunk
constexpr int STAGES = 4;                   // ring depth
constexpr size_t STAGED_MIN = 8u << 20;     // smaller copies go direct
constexpr unsigned MAX_THREADS = 16;        // host threads copying (caller included)

// Fixed pool of worker threads; copy() splits one memcpy over the workers and
// the calling thread and returns when every part is done.
class CopyPool {
  public:
    // Throws (std::system_error, std::bad_alloc) when a thread cannot be
    // started; the threads already running are stopped and joined first.
    explicit CopyPool(unsigned nthreads) {
        try {
            th_.reserve(nthreads);
            for (unsigned i = 1; i < nthreads; ++i) th_.emplace_back([this, i] { worker(i); });
        } catch (...) {
            shutdown();
            throw;
        }
    }
    ~CopyPool() { shutdown(); }
    void copy(void *dst, const void *src, size_t bytes) {
        const unsigned parts = (unsigned)th_.size() + 1;
        if (parts == 1 || bytes < (1u << 20)) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = (char *)dst;
            src_ = (const char *)src;
            bytes_ = bytes;
            parts_ = parts;
            remaining_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return remaining_ == 0; });
    }

  private:
    void shutdown() noexcept {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_)
            if (t.joinable()) t.join();
    }
    void part(unsigned i) {
        const size_t per = ((bytes_ + parts_ - 1) / parts_ + 4095) & ~(size_t)4095;
        const size_t lo = std::min(bytes_, (size_t)i * per), hi = std::min(bytes_, lo + per);
        if (hi > lo) std::memcpy(dst_ + lo, src_ + lo, hi - lo);
    }
    void worker(unsigned i) {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            part(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--remaining_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t bytes_ = 0;
    unsigned parts_ = 1, remaining_ = 0;
    unsigned long long gen_ = 0;
    bool stop_ = false;
};

}  // namespace

struct pfa_hostio {
    hipStream_t copy = nullptr;
    hipEvent_t ev[STAGES] = {}, fence = nullptr;
    void *stage[STAGES] = {};
    CopyPool *pool = nullptr;
};

void pfa_hostio_destroy(pfa_ctx *ctx) {
    pfa_hostio *io = ctx->io;
    if (!io) return;
    if (io->copy) hipStreamSynchronize(io->copy);
    for (int k = 0; k < STAGES; ++k) {
        if (io->stage[k]) hipHostFree(io->stage[k]);
        if (io->ev[k]) hipEventDestroy(io->ev[k]);
    }
    if (io->fence) hipEventDestroy(io->fence);
    if (io->copy) hipStreamDestroy(io->copy);
    delete io->pool;
    delete io;
    ctx->io = nullptr;
}

// The context's staging state, created on first use (pinned ring, copy stream
// at default priority, events, host threads); null when any of it cannot be
// had, and the copies then go direct (the runtime's own pageable path).
static pfa_hostio *hostio(pfa_ctx *ctx) {
    if (ctx->io) return ctx->io;
    if (ctx->io_failed) return nullptr;
    pfa_hostio *io = new (std::nothrow) pfa_hostio();
    if (!io) {
        ctx->io_failed = true;
        return nullptr;
    }
    ctx->io = io;
    bool ok = hipStreamCreate(&io->copy) == hipSuccess &&
              hipEventCreateWithFlags(&io->fence, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < STAGES; ++k)
        ok = hipHostMalloc(&io->stage[k], STAGE_BYTES, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&io->ev[k], hipEventDisableTiming) == hipSuccess;
    if (ok) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        try {
            io->pool = new CopyPool(std::min(MAX_THREADS, hw));
        } catch (...) {                        // no threads or no memory: copies go direct
            io->pool = nullptr;
        }
        ok = io->pool != nullptr;
    }
    if (!ok) {
        pfa_hostio_destroy(ctx);
        (void)hipGetLastError();
        ctx->io_failed = true;
        return nullptr;
    }
    return io;
}

int pfa_io_h2d(pfa_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes < STAGED_MIN) {
        PFA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return PFA_OK;
    }
    pfa_hostio *io = hostio(ctx);
    if (!io) {
        PFA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return PFA_OK;
    }
    // the copy stream starts after the compute stream's earlier work (the
    // destination may still be read by it)
    PFA_HIP(ctx, hipEventRecord(io->fence, ctx->stream));
    PFA_HIP(ctx, hipStreamWaitEvent(io->copy, io->fence, 0));
    size_t c = 0;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES, ++c) {
        const int k = (int)(c % STAGES);
        const size_t len = std::min(STAGE_BYTES, bytes - off);
        PFA_HIP(ctx, hipEventSynchronize(io->ev[k]));  // any DMA out of stage k (this or an earlier call) done
        io->pool->copy(io->stage[k], (const char *)src + off, len);
        PFA_HIP(ctx, hipMemcpyAsync((char *)dst + off, io->stage[k], len, hipMemcpyHostToDevice, io->copy));
        PFA_HIP(ctx, hipEventRecord(io->ev[k], io->copy));
    }
    PFA_HIP(ctx, hipEventRecord(io->fence, io->copy));
    PFA_HIP(ctx, hipStreamWaitEvent(ctx->stream, io->fence, 0));
    return PFA_OK;
}

int pfa_io_d2h(pfa_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes < STAGED_MIN) {
        PFA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PFA_OK;
    }
    pfa_hostio *io = hostio(ctx);
    if (!io) {
        PFA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PFA_OK;
    }
    PFA_HIP(ctx, hipEventRecord(io->fence, ctx->stream));       // the result is complete
    PFA_HIP(ctx, hipStreamWaitEvent(io->copy, io->fence, 0));
    const size_t nch = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t c) -> int {
        const int k = (int)(c % STAGES);
        const size_t off = c * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        PFA_HIP(ctx, hipMemcpyAsync(io->stage[k], (const char *)src + off, len, hipMemcpyDeviceToHost, io->copy));
        PFA_HIP(ctx, hipEventRecord(io->ev[k], io->copy));
        return PFA_OK;
    };
    for (size_t c = 0; c < std::min(nch, (size_t)STAGES); ++c) {
        const int s = issue(c);
        if (s != PFA_OK) return s;
    }
    for (size_t c = 0; c < nch; ++c) {
        const int k = (int)(c % STAGES);
        const size_t off = c * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        PFA_HIP(ctx, hipEventSynchronize(io->ev[k]));
        io->pool->copy((char *)dst + off, io->stage[k], len);
        if (c + STAGES < nch) {
            const int s = issue(c + STAGES);
            if (s != PFA_OK) return s;
        }
    }
    return PFA_OK;
}
