// Large copies between pageable host memory and the device go through a ring
// of pinned buffers, chunk by chunk: the host copies one chunk between the
// pageable array and a pinned buffer, split over a small pool of threads,
// while the DMA engine moves the neighbouring chunk between that pinned
// memory and HBM.  The DMA then reads and writes pinned memory at the link's
// rate, and the host copy and the transfer overlap.  Copies below kSmall
// bytes use hipMemcpyAsync directly.
//
// The ring and the pool are the library's and are not re-entrant, like the
// host-array entry points that use them.
#include "gf_xfer.h"

#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

constexpr size_t kChunk = size_t(16) << 20;   // bytes per pinned buffer
constexpr size_t kSmall = size_t(4) << 20;    // below this, one hipMemcpyAsync
constexpr int kBufs = 4;                      // pinned buffers in the ring
constexpr int kWorkers = 7;                   // pool threads beside the caller

// memcpy split into kWorkers + 1 parts, the caller taking the last.
class CopyPool {
public:
    CopyPool() {
        for (int i = 0; i < kWorkers; ++i) threads_.emplace_back([this, i] { run(i); });
        for (auto& t : threads_) t.detach();   // they live as long as the process
    }

    void copy(void* dst, const void* src, size_t n) {
        if (n < (size_t(1) << 20)) {
            std::memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            n_ = n;
            pending_ = kWorkers;
            ++gen_;
        }
        cv_.notify_all();
        part(kWorkers);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
    }

private:
    void part(int i) {
        const size_t per = (n_ + kWorkers) / (kWorkers + 1);
        const size_t b = std::min(n_, per * size_t(i)), e = std::min(n_, b + per);
        if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
    }

    void run(int i) {
        unsigned seen = 0;
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            part(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
    int pending_ = 0;
    unsigned gen_ = 0;
};

struct Ring {
    void* buf[kBufs] = {};
    hipEvent_t ev[kBufs] = {};
    bool used[kBufs] = {};
    CopyPool* pool = nullptr;
    bool ready = false;
};
Ring g_ring;

int ring_init() {
    if (g_ring.ready) return 0;
    for (int b = 0; b < kBufs; ++b) {
        GF_HIP(hipHostMalloc(&g_ring.buf[b], kChunk, hipHostMallocDefault));
        GF_HIP(hipEventCreateWithFlags(&g_ring.ev[b], hipEventDisableTiming));
    }
    g_ring.pool = new CopyPool();      // never freed: its threads outlive every call
    g_ring.ready = true;
    return 0;
}

}  // namespace

int gf_copy_h2d(void* dst_dev, const void* src_host, size_t bytes, hipStream_t s) {
    if (bytes == 0) return 0;
    if (bytes < kSmall) {
        GF_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, s));
        return 0;
    }
    if (ring_init()) return -1;
    const char* src = static_cast<const char*>(src_host);
    char* dst = static_cast<char*>(dst_dev);
    int b = 0;
    for (size_t off = 0; off < bytes; off += kChunk, b = (b + 1) % kBufs) {
        const size_t len = std::min(kChunk, bytes - off);
        if (g_ring.used[b]) GF_HIP(hipEventSynchronize(g_ring.ev[b]));   // its last DMA is done
        g_ring.pool->copy(g_ring.buf[b], src + off, len);
        GF_HIP(hipMemcpyAsync(dst + off, g_ring.buf[b], len, hipMemcpyHostToDevice, s));
        GF_HIP(hipEventRecord(g_ring.ev[b], s));
        g_ring.used[b] = true;
    }
    return 0;
}

int gf_copy_d2h(void* dst_host, const void* src_dev, size_t bytes, hipStream_t s) {
    if (bytes == 0) return 0;
    if (bytes < kSmall) {
        GF_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        return 0;
    }
    if (ring_init()) return -1;
    const char* src = static_cast<const char*>(src_dev);
    char* dst = static_cast<char*>(dst_host);
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    // keep up to kBufs chunks in flight; copy each out once its DMA is done
    for (size_t c = 0; c < nchunks + kBufs; ++c) {
        if (c >= size_t(kBufs)) {
            const size_t d = c - kBufs;              // the chunk that left buffer d % kBufs
            if (d < nchunks) {
                const int b = int(d % kBufs);
                GF_HIP(hipEventSynchronize(g_ring.ev[b]));
                const size_t off = d * kChunk, len = std::min(kChunk, bytes - off);
                g_ring.pool->copy(dst + off, g_ring.buf[b], len);
            }
        }
        if (c < nchunks) {
            const int b = int(c % kBufs);
            // a buffer not drained above may still feed an earlier call's copy
            if (c < size_t(kBufs) && g_ring.used[b]) GF_HIP(hipEventSynchronize(g_ring.ev[b]));
            const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
            GF_HIP(hipMemcpyAsync(g_ring.buf[b], src + off, len, hipMemcpyDeviceToHost, s));
            GF_HIP(hipEventRecord(g_ring.ev[b], s));
            g_ring.used[b] = true;
        }
    }
    return 0;
}
