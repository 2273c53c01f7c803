// Host copy helpers shared by the whole-call pipeline and the server.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace bcrt {

// Fixed pool of host threads; copy() splits one memcpy over all of them and the
// calling thread (one copy at a time).
class CopyPool {
   public:
    explicit CopyPool(unsigned workers) {
        for (unsigned i = 0; i < workers; ++i) threads_.emplace_back([this, i] { loop(i); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }
    void copy(void* dst, const void* src, size_t bytes) {
        const size_t parts = threads_.size() + 1;
        if (bytes < (size_t(1) << 20) || parts == 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        const size_t step = (bytes / parts + 4095) / 4096 * 4096;
        auto part = [=](size_t k) {
            const size_t b0 = std::min(bytes, k * step), b1 = std::min(bytes, (k + 1) * step);
            if (b1 > b0) std::memcpy(static_cast<char*>(dst) + b0, static_cast<const char*>(src) + b0, b1 - b0);
        };
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = part;
            pending_ = threads_.size();
            ++gen_;
        }
        cv_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
    }

   private:
    void loop(unsigned i) {
        uint64_t seen = 0;
        for (;;) {
            std::function<void(size_t)> job;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = job_;
            }
            job(i + 1);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::function<void(size_t)> job_;
    size_t pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Transparent huge pages for the 2 MiB-aligned interior of a caller's output
// array, so first-touch faults of a fresh buffer come 2 MiB at a time.  Only
// the page size changes; failure (no THP) is harmless.
void advise_huge(void* p, size_t bytes);

// Workers for one copy pool of the whole call: half the hardware threads, less
// the caller, at most 7; BCRT_COPY_THREADS overrides the pool size.
unsigned pool_workers();

// Workers for one copy pool of the product server: all hardware threads but
// two, less the caller; BCRT_SERVER_COPY_THREADS overrides the pool size.
unsigned server_pool_workers();

}  // namespace bcrt
