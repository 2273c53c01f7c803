// fherma_init / fherma_run / fherma_free over the C ABI (include/bytecrt.h).
//
// Error behaviour follows the reference adapter: B outside 1..128 and wrong
// input sizes raise std::runtime_error (native-batched/solve.cu:161-162,
// 207-208); geometry and modulus problems raise std::invalid_argument
// (rns-plan.hpp:182-196); device failures raise std::runtime_error
// (core/solve.cu:46-49).  All messages start with "ByteCRT".
#include "../../include/fherma.h"

#include <malloc.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bytecrt.h"
#include "bytecrt_copy.hpp"
#include "bytecrt_trace.hpp"

#ifndef MADV_POPULATE_WRITE
#define MADV_POPULATE_WRITE 23
#endif

namespace {

struct Solver {
    bcrt_engine* engine = nullptr;
    uint32_t B = 0, N = 0, L = 0;
    std::unique_ptr<bcrt::CopyPool> out_pool;  // batched: copies finished pieces into the owned vector
    ~Solver() { bcrt_destroy(engine); }
};

// Owned outputs of at least this many bytes are filled in parallel (Grower);
// smaller ones by appends.
constexpr size_t kParallelOutputBytes = size_t(64) << 20;

[[noreturn]] void raise(int code, const std::string& msg) {
    if (code == BCRT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

// Faults in the pages of the output vector's reserved storage from a few
// threads, front to back, ahead of the single-threaded appends (a fresh
// allocation otherwise costs one page fault per 4 KiB inside the timed call).
// Only page tables change: the storage is not written.  Without kernel
// support for MADV_POPULATE_WRITE it does nothing.
class Prefault {
   public:
    Prefault(void* p, size_t bytes) {
        constexpr uintptr_t kHuge = uintptr_t(2) << 20;
        begin_ = (uintptr_t(p) + kHuge - 1) & ~(kHuge - 1);
        end_ = (uintptr_t(p) + bytes) & ~(kHuge - 1);
        if (end_ <= begin_ + kPiece) return;
        (void)madvise(reinterpret_cast<void*>(begin_), end_ - begin_, MADV_HUGEPAGE);
        const unsigned n = std::max(1u, std::min(4u, std::thread::hardware_concurrency() / 4));
        for (unsigned i = 0; i < n; ++i) threads_.emplace_back([this] { work(); });
    }
    ~Prefault() {
        stop_ = true;
        for (auto& t : threads_) t.join();
    }

   private:
    static constexpr uintptr_t kPiece = uintptr_t(32) << 20;
    void work() {
        for (;;) {
            const uintptr_t b = begin_ + next_.fetch_add(1) * kPiece;
            if (stop_ || b >= end_) return;
            if (madvise(reinterpret_cast<void*>(b), std::min(kPiece, end_ - b), MADV_POPULATE_WRITE) != 0 &&
                errno == EINVAL) {
                stop_ = true;
                return;
            }
        }
    }
    uintptr_t begin_ = 0, end_ = 0;
    std::atomic<uintptr_t> next_{0};
    std::atomic<bool> stop_{false};
    std::vector<std::thread> threads_;
};

void append(void* ctx, const uint32_t* words, size_t nwords) {
    auto* v = static_cast<std::vector<uint32_t>*>(ctx);
    v->insert(v->end(), words, words + nwords);
}

// Grows the owned output vector, whose storage is reserved for the whole
// result, in steps ahead of the finished pieces: a resize within the
// capacity value-initializes (zero-fills) the new elements and never moves
// the storage, and once a step is published the pieces below it are copied
// in by several threads at once, instead of one thread appending them (the
// append had been the longest lane of a B=128 call).  Only this thread
// touches the vector object; the copies write its elements through a
// pointer taken before the thread starts, and only below the published size.
class Grower {
   public:
    Grower(std::vector<uint32_t>& v, size_t total) : v_(v), total_(total) {
        th_ = std::thread([this] { run(); });
    }
    ~Grower() {
        th_.join();  // always reaches total: the vector then has its full size
    }
    // Waits until the first n elements exist.
    void wait_for(size_t n) {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return ready_ >= n; });
    }

   private:
    static constexpr size_t kStep = (size_t(16) << 20) / sizeof(uint32_t);  // 16 MiB per resize
    void run() {
        bcrt::TraceScope tr("bcrt:grow");
        size_t n = v_.size();
        while (n < total_) {
            n = std::min(total_, n + kStep);
            v_.resize(n);
            {
                std::lock_guard<std::mutex> g(m_);
                ready_ = n;
            }
            cv_.notify_all();
        }
    }
    std::vector<uint32_t>& v_;
    const size_t total_;
    std::mutex m_;
    std::condition_variable cv_;
    size_t ready_ = 0;
    std::thread th_;
};

// Sink of a batched call: each finished piece, once the vector has grown
// past it, is copied into place by the solver's copy pool.
struct ParallelOut {
    uint32_t* base;
    size_t off;
    Grower* grower;
    bcrt::CopyPool* pool;
};

void place(void* ctx, const uint32_t* words, size_t nwords) {
    auto* o = static_cast<ParallelOut*>(ctx);
    o->grower->wait_for(o->off + nwords);
    bcrt::TraceScope tr("bcrt:place");
    o->pool->copy(o->base + o->off, words, nwords * sizeof(uint32_t));
    o->off += nwords;
}

}  // namespace

void* fherma_init(const fherma::Point& p) {
    if (p.B == 0 || p.B > 128) throw std::runtime_error("ByteCRT batch coverage B=1..128");
    if (p.q.data.size() != p.L) throw std::invalid_argument("ByteCRT RNS modulus limb count");
    if (p.B > 1) {
        // A batch result is one large owned vector per call.  Keep large
        // blocks in the ordinary heap and never trim it, so each call's output
        // reuses pages that earlier calls already faulted in instead of a
        // fresh mapping (the reference's batched adapter makes the same two
        // settings at init: native-batched/owned-output-cpu.md).  This is a
        // process-wide allocator policy; memory is returned at exit.
        (void)mallopt(M_MMAP_THRESHOLD, INT_MAX);
        (void)mallopt(M_TRIM_THRESHOLD, -1);
    }
    auto s = std::make_unique<Solver>();
    char err[256] = {0};
    const int rc = bcrt_create(&s->engine, p.B, p.N, p.W, p.L, p.q.data.data(), err, sizeof(err));
    if (rc != BCRT_OK) raise(rc, err);
    s->B = p.B;
    s->N = p.N;
    s->L = p.L;
    if (size_t(p.B) * p.N * p.L * sizeof(uint32_t) >= kParallelOutputBytes)
        s->out_pool = std::make_unique<bcrt::CopyPool>(bcrt::pool_workers());
    return s.release();
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto& s = *static_cast<Solver*>(state);
    bcrt::TraceScope trace("bcrt:fherma_run");
    const size_t total = size_t(s.B) * s.N * s.L;
    if (in.a.data.size() != total || in.b.data.size() != total)
        throw std::runtime_error("ByteCRT input coefficient count");
    fherma::Outputs out;
    out.c.shape = {int64_t(s.B), int64_t(s.N), int64_t(s.L)};
    out.c.data.reserve(total);
    int rc;
    {
        Prefault pages(out.c.data.data(), total * sizeof(uint32_t));
        if (s.out_pool) {
            // the pointer is taken before the grower thread starts
            ParallelOut o{out.c.data.data(), 0, nullptr, s.out_pool.get()};
            Grower grow(out.c.data, total);
            o.grower = &grow;
            rc = bcrt_run_sink(s.engine, in.a.data.data(), in.b.data.data(), s.B, place, &o);
        } else {
            rc = bcrt_run_sink(s.engine, in.a.data.data(), in.b.data.data(), s.B, append, &out.c.data);
        }
    }
    if (rc != BCRT_OK) raise(rc, bcrt_last_error(s.engine));
    if (out.c.data.size() != total) throw std::runtime_error("ByteCRT output coefficient count");
    return out;
}

void fherma_free(void* state) { delete static_cast<Solver*>(state); }
