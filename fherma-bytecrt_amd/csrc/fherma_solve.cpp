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
#include <memory>
#include <thread>
#include <vector>

#include "../../include/bytecrt.h"
#include "bytecrt_trace.hpp"

#ifndef MADV_POPULATE_WRITE
#define MADV_POPULATE_WRITE 23
#endif

namespace {

struct Solver {
    bcrt_engine* engine = nullptr;
    uint32_t B = 0, N = 0, L = 0;
    ~Solver() { bcrt_destroy(engine); }
};

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
        rc = bcrt_run_sink(s.engine, in.a.data.data(), in.b.data.data(), s.B, append, &out.c.data);
    }
    if (rc != BCRT_OK) raise(rc, bcrt_last_error(s.engine));
    if (out.c.data.size() != total) throw std::runtime_error("ByteCRT output coefficient count");
    return out;
}

void fherma_free(void* state) { delete static_cast<Solver*>(state); }
