// staging.cpp -- copies between pageable host buffers and device memory at
// close to link speed, for the host stage layer's burst rasters.  A copy from
// pageable memory goes through a single CPU copy into the runtime's own
// staging (about 13 GB/s measured here); instead, several host threads copy
// each 32 MiB chunk into one of two pinned buffers (hipHostMalloc) while the
// DMA of the previous chunk runs on a stream of the pool's own.
//
// Each call is ordered after the work already queued on the null stream (the
// stage layer's launches) and returns when its copies are complete, so it
// replaces a synchronous hipMemcpy.  The pool of each device (2 x 32 MiB
// pinned, one stream, two events) is made on first use and kept until the
// process ends; calls on one device are serialised by its lock.
#include "tops_internal.hpp"

#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

namespace tops {
namespace {

constexpr size_t kChunk = 32u << 20;
constexpr int kMaxDevices = 16;
constexpr int kMaxThreads = 8;

struct Pool {
    std::mutex mu;
    bool ready = false;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t s = nullptr;
};

Pool g_pools[kMaxDevices];

int pool_of(Pool** out)
{
    int dev = 0;
    TOPS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(TOPS_ERR_ARG, "staging: device %d beyond %d", dev, kMaxDevices);
    *out = &g_pools[dev];
    return TOPS_OK;
}

int pool_init(Pool& p)
{
    if (p.ready) return TOPS_OK;
    TOPS_HIP(hipStreamCreateWithFlags(&p.s, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) {
        TOPS_HIP(hipHostMalloc(&p.buf[i], kChunk, hipHostMallocDefault));
        TOPS_HIP(hipEventCreateWithFlags(&p.ev[i], hipEventDisableTiming));
    }
    p.ready = true;
    return TOPS_OK;
}

// dst <- src (n bytes, host memory) with up to kMaxThreads threads
void parallel_copy(void* dst, const void* src, size_t n)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int t = (int)std::min<size_t>(std::min<unsigned>(kMaxThreads, hw),
                                        std::max<size_t>(1, n / (1u << 20)));
    if (t <= 1) {
        memcpy(dst, src, n);
        return;
    }
    const size_t piece = (n / t + 63) / 64 * 64;
    std::vector<std::thread> th;
    th.reserve(t - 1);
    for (int i = 1; i < t; i++) {
        const size_t off = piece * i;
        if (off >= n) break;
        const size_t len = std::min(piece, n - off);
        th.emplace_back([=] { memcpy((char*)dst + off, (const char*)src + off, len); });
    }
    memcpy(dst, src, std::min(piece, n));
    for (auto& x : th) x.join();
}

}  // namespace

int staged_upload(void* d_dst, const void* h_src, size_t bytes)
{
    if (bytes == 0) return TOPS_OK;
    Pool* p = nullptr;
    int rc = pool_of(&p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(p->mu);
    if ((rc = pool_init(*p))) return rc;
    TOPS_HIP(hipStreamSynchronize(p->s));      // no copy of an earlier call is in flight
    TOPS_HIP(hipStreamSynchronize(nullptr));   // the destination is no longer in use
    bool used[2] = {false, false};
    int k = 0;
    for (size_t off = 0; off < bytes; off += kChunk, k ^= 1) {
        const size_t n = std::min(kChunk, bytes - off);
        if (used[k]) TOPS_HIP(hipEventSynchronize(p->ev[k]));
        parallel_copy(p->buf[k], (const char*)h_src + off, n);
        TOPS_HIP(hipMemcpyAsync((char*)d_dst + off, p->buf[k], n, hipMemcpyHostToDevice, p->s));
        TOPS_HIP(hipEventRecord(p->ev[k], p->s));
        used[k] = true;
    }
    TOPS_HIP(hipStreamSynchronize(p->s));
    return TOPS_OK;
}

int staged_download(void* h_dst, const void* d_src, size_t bytes)
{
    if (bytes == 0) return TOPS_OK;
    Pool* p = nullptr;
    int rc = pool_of(&p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(p->mu);
    if ((rc = pool_init(*p))) return rc;
    TOPS_HIP(hipStreamSynchronize(p->s));      // no copy of an earlier call is in flight
    TOPS_HIP(hipStreamSynchronize(nullptr));   // the source has been written
    const size_t nchunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * kChunk, n = std::min(kChunk, bytes - off);
        TOPS_HIP(hipMemcpyAsync(p->buf[i & 1], (const char*)d_src + off, n, hipMemcpyDeviceToHost,
                                p->s));
        TOPS_HIP(hipEventRecord(p->ev[i & 1], p->s));
        return TOPS_OK;
    };
    if ((rc = issue(0))) return rc;
    for (size_t i = 0; i < nchunks; i++) {
        // the next chunk's DMA goes into the other buffer, whose host copy
        // finished in the previous iteration
        if (i + 1 < nchunks && (rc = issue(i + 1))) return rc;
        TOPS_HIP(hipEventSynchronize(p->ev[i & 1]));
        const size_t off = i * kChunk, n = std::min(kChunk, bytes - off);
        parallel_copy((char*)h_dst + off, p->buf[i & 1], n);
    }
    return TOPS_OK;
}

}  // namespace tops
