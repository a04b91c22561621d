// Copies between device memory and pageable caller memory. A pageable hipMemcpy moves ~17 GB/s
// device to host and ~35 GB/s host to device on the MI355X host link, while DMA to or from pinned
// memory moves ~56 GB/s and a few host threads copy between pinned and pageable memory at
// ~100 GB/s. copy_to_host and copy_from_host therefore stream the data through two pinned
// chunks: the DMA of one chunk runs while host threads copy the other.
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "kq_solve.h"

namespace kq {

namespace {

constexpr size_t kChunk = (size_t)8 << 20;  // bytes per pinned chunk
constexpr size_t kMinPinned = (size_t)1 << 20;  // smaller copies go straight through hipMemcpy

// Helper threads that split one memcpy with the calling thread. The library is driven by one
// host thread at a time, so one job is in flight at most.
class CopyPool {
public:
    static CopyPool &get()
    {
        static CopyPool p;
        return p;
    }
    void copy(void *dst, const void *src, size_t n)
    {
        const int parts = (int)workers_.size() + 1;
        if (parts == 1 || n < ((size_t)256 << 10)) {
            memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = static_cast<char *>(dst);
            src_ = static_cast<const char *>(src);
            n_ = n;
            parts_ = parts;
            pending_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
    }

private:
    CopyPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int nw = hw >= 8 ? 3 : (hw >= 4 ? 1 : 0);  // 4 copying threads were fastest on the box
        for (int i = 0; i < nw; ++i)
            workers_.emplace_back([this, i] { loop(i + 1); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : workers_)
            t.join();
    }
    void part(int i)
    {
        const size_t per = (n_ + parts_ - 1) / parts_, o = per * (size_t)i;
        if (o < n_)
            memcpy(dst_ + o, src_ + o, std::min(per, n_ - o));
    }
    void loop(int id)
    {
        unsigned long seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
            if (quit_)
                return;
            seen = gen_;
            lk.unlock();
            part(id);
            lk.lock();
            if (--pending_ == 0)
                done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    char *dst_ = nullptr;
    const char *src_ = nullptr;
    size_t n_ = 0;
    int parts_ = 1, pending_ = 0;
    unsigned long gen_ = 0;
    bool quit_ = false;
};

struct PinnedChunks {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ready = false;
    bool init()
    {
        if (ready)
            return true;
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault) != hipSuccess ||
                hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
        }
        ready = true;
        return true;
    }
};

PinnedChunks &chunks()
{
    static PinnedChunks c;
    return c;
}

// true when p is memory the device cannot address directly (plain pageable host memory)
bool pageable(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return a.type == hipMemoryTypeUnregistered;
}

}  // namespace

void copy_to_host(void *h, const void *d, size_t bytes, hipStream_t st)
{
    PinnedChunks &pc = chunks();
    if (bytes < kMinPinned || !pageable(h) || !pc.init()) {
        KQ_HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDefault, st));
        KQ_HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    const char *src = static_cast<const char *>(d);
    char *dst = static_cast<char *>(h);
    const size_t nch = (bytes + kChunk - 1) / kChunk;
    for (size_t i = 0; i <= nch; ++i) {
        if (i < nch) {
            const size_t o = i * kChunk, m = std::min(kChunk, bytes - o);
            KQ_HIP_CHECK(hipMemcpyAsync(pc.buf[i & 1], src + o, m, hipMemcpyDeviceToHost, st));
            KQ_HIP_CHECK(hipEventRecord(pc.ev[i & 1], st));
        }
        if (i > 0) {
            const size_t j = i - 1, o = j * kChunk, m = std::min(kChunk, bytes - o);
            KQ_HIP_CHECK(hipEventSynchronize(pc.ev[j & 1]));
            CopyPool::get().copy(dst + o, pc.buf[j & 1], m);
        }
    }
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

void copy_from_host(void *d, const void *h, size_t bytes, hipStream_t st)
{
    PinnedChunks &pc = chunks();
    if (bytes < kMinPinned || !pageable(h) || !pc.init()) {
        KQ_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyDefault, st));
        KQ_HIP_CHECK(hipStreamSynchronize(st));
        return;
    }
    const char *src = static_cast<const char *>(h);
    char *dst = static_cast<char *>(d);
    const size_t nch = (bytes + kChunk - 1) / kChunk;
    for (size_t i = 0; i < nch; ++i) {
        const size_t o = i * kChunk, m = std::min(kChunk, bytes - o);
        if (i >= 2)  // the DMA that last read this chunk has finished
            KQ_HIP_CHECK(hipEventSynchronize(pc.ev[i & 1]));
        CopyPool::get().copy(pc.buf[i & 1], src + o, m);
        KQ_HIP_CHECK(hipMemcpyAsync(dst + o, pc.buf[i & 1], m, hipMemcpyHostToDevice, st));
        KQ_HIP_CHECK(hipEventRecord(pc.ev[i & 1], st));
    }
    // both chunks free again before the next copy in either direction uses them
    KQ_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace kq
