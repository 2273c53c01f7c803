// Whole-call host I/O pipeline (replaces the reference's pinned-slot staging,
// core/solve.cu:191-291 and native-batched/solve.cu:56-156).
#include "bytecrt_io.hpp"

#include "bytecrt_copy.hpp"
#include "bytecrt_trace.hpp"

#include <sched.h>
#include <sys/mman.h>

#include <algorithm>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace bcrt {

namespace {

constexpr uint32_t kSlots = 4;
constexpr size_t kInPiece = size_t(16) << 20;   // pack + host-to-device piece
constexpr size_t kOutPiece = size_t(4) << 20;   // smallest device-to-host + unpack piece
constexpr uint32_t kMaxOutPieces = 8;
constexpr size_t kChunkInputBytes = size_t(64) << 20;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("ByteCRT ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

void advise_huge(void* p, size_t bytes) {
    const uintptr_t huge = uintptr_t(2) << 20;
    const uintptr_t b = (uintptr_t(p) + huge - 1) & ~(huge - 1), e = (uintptr_t(p) + bytes) & ~(huge - 1);
    if (e > b) (void)madvise(reinterpret_cast<void*>(b), e - b, MADV_HUGEPAGE);
}

namespace {
// CPUs this process may run on (its affinity mask); hardware_concurrency()
// counts every online CPU of the host, which a container may not own.
unsigned usable_cpus() {
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof(set), &set) == 0) {
        const int n = CPU_COUNT(&set);
        if (n > 0) return unsigned(n);
    }
    return std::max(1u, std::thread::hardware_concurrency());
}

// Threads per pool (the caller included) from an environment variable
// (1..64), else the given default; returns the worker count (threads - 1).
unsigned workers_from(const char* env, unsigned threads) {
    if (const char* v = std::getenv(env)) {
        const long n = std::strtol(v, nullptr, 10);
        if (n >= 1 && n <= 64) return unsigned(n - 1);
    }
    return std::max(1u, threads) - 1;
}
}  // namespace

unsigned pool_workers() {
    // BCRT_COPY_THREADS; default min(8, hw/2): the whole call also runs a
    // drain thread and the output pre-faulting threads beside both pools.
    const unsigned hw = std::max(2u, usable_cpus());
    return workers_from("BCRT_COPY_THREADS", std::min(8u, hw / 2));
}

unsigned server_pool_workers() {
    // BCRT_SERVER_COPY_THREADS; default hw - 2 per pool: the server's packing
    // and unpacking run at once and bound its tail latency (p99 at 1,000
    // requests/s on 16 threads: 12-15 ms with 8 per pool, 3-4 ms with 14).
    const unsigned hw = std::max(4u, usable_cpus());
    return workers_from("BCRT_SERVER_COPY_THREADS", std::min(64u, hw - 2));
}

struct HostIo::Impl {
    struct Job {
        uint32_t slot;
        size_t bytes, piece;
        uint32_t pieces;
        uint32_t* dest;
        bcrt_sink sink;
        void* ctx;
    };

    size_t words = 0;
    uint32_t chunk = 0;
    hipStream_t compute = nullptr;
    uint32_t* dev = nullptr;      // [kSlots][3][chunk][words]
    uint32_t* pin_in = nullptr;   // [kSlots][2][chunk][words]
    uint32_t* pin_out = nullptr;  // [kSlots][chunk][words]
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t ev_in[kSlots] = {};
    hipEvent_t ev_done[kSlots] = {};
    hipEvent_t ev_piece[kSlots][kMaxOutPieces] = {};
    std::unique_ptr<CopyPool> pack, unpack;
    std::thread drainer;
    std::mutex m;
    std::condition_variable cv, done_cv;
    std::deque<Job> jobs;
    uint64_t posted = 0, drained = 0;
    bool stop = false;
    std::exception_ptr error;

    void drain_loop() {
        for (;;) {
            Job j;
            bool skip;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                j = jobs.front();
                jobs.pop_front();
                skip = bool(error);
            }
            try {
                const uint32_t* src = pin_out + size_t(j.slot) * chunk * words;
                for (uint32_t p = 0; p < j.pieces; ++p) {
                    {
                        TraceScope tr("bcrt:wait");
                        check(hipEventSynchronize(ev_piece[j.slot][p]), "wait output copy");
                    }
                    if (skip) continue;
                    TraceScope tr("bcrt:unpack");
                    const size_t off = p * j.piece, len = std::min(j.piece, j.bytes - off);
                    const uint32_t* s = src + off / sizeof(uint32_t);
                    if (j.dest)
                        unpack->copy(j.dest + off / sizeof(uint32_t), s, len);
                    else
                        j.sink(j.ctx, s, len / sizeof(uint32_t));
                }
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!error) error = std::current_exception();
            }
            {
                std::lock_guard<std::mutex> g(m);
                ++drained;
            }
            done_cv.notify_all();
        }
    }

    void release() {
        {
            std::lock_guard<std::mutex> g(m);
            stop = true;
        }
        cv.notify_all();
        if (drainer.joinable()) drainer.join();
        for (hipStream_t s : {h2d, d2h})
            if (s) (void)hipStreamSynchronize(s), (void)hipStreamDestroy(s);
        for (uint32_t k = 0; k < kSlots; ++k) {
            for (hipEvent_t ev : {ev_in[k], ev_done[k]})
                if (ev) (void)hipEventDestroy(ev);
            for (hipEvent_t ev : ev_piece[k])
                if (ev) (void)hipEventDestroy(ev);
        }
        if (dev) (void)hipFree(dev);
        if (pin_in) (void)hipHostFree(pin_in);
        if (pin_out) (void)hipHostFree(pin_out);
    }
};

HostIo::HostIo(size_t words, uint32_t max_chunk, hipStream_t compute) : impl_(new Impl) {
    Impl& s = *impl_;
    try {
        s.words = words;
        s.compute = compute;
        const size_t per_product = 2 * words * sizeof(uint32_t);
        s.chunk = uint32_t(std::max<size_t>(1, std::min<size_t>(max_chunk, kChunkInputBytes / per_product)));
        const size_t slot_words = size_t(s.chunk) * words;
        check(hipMalloc(reinterpret_cast<void**>(&s.dev), kSlots * 3 * slot_words * sizeof(uint32_t)),
              "allocate io buffers");
        check(hipHostMalloc(reinterpret_cast<void**>(&s.pin_in), kSlots * 2 * slot_words * sizeof(uint32_t),
                            hipHostMallocDefault),
              "allocate pinned inputs");
        check(hipHostMalloc(reinterpret_cast<void**>(&s.pin_out), kSlots * slot_words * sizeof(uint32_t),
                            hipHostMallocDefault),
              "allocate pinned outputs");
        check(hipStreamCreateWithFlags(&s.h2d, hipStreamNonBlocking), "create h2d stream");
        check(hipStreamCreateWithFlags(&s.d2h, hipStreamNonBlocking), "create d2h stream");
        for (uint32_t k = 0; k < kSlots; ++k) {
            check(hipEventCreateWithFlags(&s.ev_in[k], hipEventDisableTiming), "create io event");
            check(hipEventCreateWithFlags(&s.ev_done[k], hipEventDisableTiming), "create io event");
            for (auto& ev : s.ev_piece[k]) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "create io event");
        }
        s.pack = std::make_unique<CopyPool>(pool_workers());
        s.unpack = std::make_unique<CopyPool>(pool_workers());
        s.drainer = std::thread([&s] { s.drain_loop(); });
    } catch (...) {
        s.release();
        throw;
    }
}

HostIo::~HostIo() { impl_->release(); }

uint32_t HostIo::chunk() const { return impl_->chunk; }

void HostIo::run(const uint32_t* a, const uint32_t* b, uint32_t batch, uint32_t* c, bcrt_sink sink, void* ctx,
                 const Enqueue& enqueue) {
    Impl& s = *impl_;
    TraceScope whole("bcrt:whole_call");
    const size_t words = s.words, chunk = s.chunk, slot_words = chunk * words;
    const uint32_t nchunks = uint32_t((batch + chunk - 1) / chunk);
    if (c) advise_huge(c, size_t(batch) * words * sizeof(uint32_t));
    uint64_t base;
    {
        std::lock_guard<std::mutex> g(s.m);
        base = s.posted;
        s.error = nullptr;
    }
    auto wait_drained = [&](uint64_t target) {
        std::unique_lock<std::mutex> lk(s.m);
        s.done_cv.wait(lk, [&] { return s.drained >= target; });
        return !s.error;
    };
    try {
        for (uint32_t k = 0; k < nchunks; ++k) {
            const uint32_t slot = k % kSlots;
            const uint32_t n = uint32_t(std::min<size_t>(chunk, batch - size_t(k) * chunk));
            // Slot k % kSlots is free once chunk k - kSlots has been unpacked:
            // its output copies, hence its compute and input copies, are done.
            if (k >= kSlots) {
                TraceScope tr("bcrt:slot_wait");
                if (!wait_drained(base + k - kSlots + 1)) break;
            }
            const size_t bytes = size_t(n) * words * sizeof(uint32_t);
            uint32_t* hin = s.pin_in + size_t(slot) * 2 * slot_words;
            uint32_t* din = s.dev + size_t(slot) * 3 * slot_words;
            for (int op = 0; op < 2; ++op) {
                const uint32_t* src = (op ? b : a) + size_t(k) * slot_words;
                uint32_t* h = hin + op * slot_words;
                uint32_t* d = din + op * slot_words;
                for (size_t off = 0; off < bytes; off += kInPiece) {
                    const size_t len = std::min(kInPiece, bytes - off), w = off / sizeof(uint32_t);
                    {
                        TraceScope tr("bcrt:pack");
                        s.pack->copy(h + w, src + w, len);
                    }
                    check(hipMemcpyAsync(d + w, h + w, len, hipMemcpyHostToDevice, s.h2d), "copy operand");
                }
            }
            check(hipEventRecord(s.ev_in[slot], s.h2d), "record input");
            check(hipStreamWaitEvent(s.compute, s.ev_in[slot], 0), "compute waits input");
            uint32_t* dout = din + 2 * slot_words;
            enqueue(din, din + slot_words, dout, n, s.compute);
            check(hipEventRecord(s.ev_done[slot], s.compute), "record compute");
            check(hipStreamWaitEvent(s.d2h, s.ev_done[slot], 0), "copy waits compute");
            Impl::Job j{};
            j.slot = slot;
            j.bytes = bytes;
            j.pieces = uint32_t(std::min<size_t>(kMaxOutPieces, std::max<size_t>(1, bytes / kOutPiece)));
            j.piece = ((bytes + j.pieces - 1) / j.pieces + 4095) / 4096 * 4096;
            j.pieces = uint32_t((bytes + j.piece - 1) / j.piece);
            j.dest = c ? c + size_t(k) * slot_words : nullptr;
            j.sink = sink;
            j.ctx = ctx;
            uint32_t* hout = s.pin_out + size_t(slot) * slot_words;
            for (uint32_t p = 0; p < j.pieces; ++p) {
                const size_t off = p * j.piece, len = std::min(j.piece, bytes - off), w = off / sizeof(uint32_t);
                check(hipMemcpyAsync(hout + w, dout + w, len, hipMemcpyDeviceToHost, s.d2h), "copy result");
                check(hipEventRecord(s.ev_piece[slot][p], s.d2h), "record output");
            }
            {
                std::lock_guard<std::mutex> g(s.m);
                s.jobs.push_back(j);
                ++s.posted;
            }
            s.cv.notify_one();
        }
    } catch (...) {
        uint64_t posted;
        {
            std::lock_guard<std::mutex> g(s.m);
            posted = s.posted;
        }
        wait_drained(posted);
        for (hipStream_t st : {s.h2d, s.compute, s.d2h}) (void)hipStreamSynchronize(st);
        throw;
    }
    uint64_t posted;
    {
        std::lock_guard<std::mutex> g(s.m);
        posted = s.posted;
    }
    wait_drained(posted);
    std::lock_guard<std::mutex> g(s.m);
    if (s.error) {
        std::exception_ptr e = s.error;
        s.error = nullptr;
        std::rethrow_exception(e);
    }
}

}  // namespace bcrt
