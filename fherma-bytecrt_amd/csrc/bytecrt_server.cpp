// Continuous-batching product server: requests of mixed geometry arrive at any
// time, are admitted into "waves" while HBM budget and in-flight slots allow,
// and each wave runs as one mixed launch per stage (bcrt group kernels) on one
// of two compute streams, so a wave admitted while another is on the GPU
// overlaps with it.  Plan tables stay resident per (N, W, q); per-request
// device memory (operands, result, residue planes) comes from one arena.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bytecrt.h"
#include "bytecrt_kernels.hpp"

namespace bcrt {
// Tables and geometry of an engine (bytecrt_engine.cpp).
const DevTables& engine_tables(const bcrt_engine* e);
}

namespace {

using Clock = std::chrono::steady_clock;

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("ByteCRT server ") + what + ": " + hipGetErrorString(e));
}

constexpr uint32_t kInflight = 4;   // waves on the device at once
constexpr size_t kAlign = 256;

struct Plan {
    bcrt_engine* engine = nullptr;  // owns the resident tables
    bcrt::DevTables t{};
    size_t io_words = 0, plane_words = 0;
    ~Plan() { bcrt_destroy(engine); }
};

struct Request {
    uint64_t id = 0;
    Plan* plan = nullptr;
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    uint32_t* c = nullptr;
    Clock::time_point submitted;
    double latency = 0;
    bool done = false;
    std::string error;
};

// A region of each of the three FIFO rings: device arena, pinned inputs,
// pinned outputs.
enum Ring { kArena = 0, kPinIn = 1, kPinOut = 2 };

struct Wave {
    std::vector<Request*> reqs;
    std::vector<bcrt::GroupDesc> groups;  // host copy, alive until the upload completes
    size_t start[3] = {0, 0, 0}, end[3] = {0, 0, 0};
    uint32_t slot = 0;
    hipEvent_t inputs_ready = nullptr;
    hipEvent_t done = nullptr;
};

size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

}  // namespace

struct bcrt_server {
    size_t budget = 0;      // device arena bytes
    size_t cap[3] = {0, 0, 0};
    uint32_t max_wave = 0;
    uint8_t* arena = nullptr;
    uint8_t* pin_in = nullptr;   // pinned staging of operands (cap[kPinIn] bytes)
    uint8_t* pin_out = nullptr;  // pinned staging of results (cap[kPinOut] bytes)
    bcrt::GroupDesc* d_groups = nullptr;  // [kInflight][kMaxGroups]
    hipStream_t copy_in = nullptr;
    hipStream_t compute[2] = {nullptr, nullptr};
    uint32_t next_compute = 0;
    std::mutex m;
    std::condition_variable cv, done_cv;
    std::deque<Request*> pending;
    std::deque<Wave> inflight;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    std::map<std::string, std::unique_ptr<Plan>> plans;
    std::mutex plan_m;
    uint64_t next_id = 1;
    bool stop = false;
    bool slot_busy[kInflight] = {false, false, false, false};
    std::string fatal;
    std::thread worker;
    // statistics
    uint64_t completed = 0, waves = 0, max_wave_seen = 0;
    std::string last_error;

    void run();
    bool admit(std::unique_lock<std::mutex>& lk);
    void retire();
};

namespace {

// Start of a free region of `size` bytes in ring `k` (capacity cap), given the
// in-flight waves in admission order.
bool ring_alloc(const std::deque<Wave>& inflight, int k, size_t cap, size_t size, size_t* start) {
    if (size > cap) return false;
    if (inflight.empty()) {
        *start = 0;
        return true;
    }
    const size_t first = inflight.front().start[k], last = inflight.back().end[k];
    if (last > first) {  // not wrapped: free space after `last` and before `first`
        if (cap - last >= size) {
            *start = last;
            return true;
        }
        if (first >= size) {
            *start = 0;
            return true;
        }
        return false;
    }
    if (first - last >= size) {  // wrapped (last <= first): free space between them
        *start = last;
        return true;
    }
    return false;
}

size_t footprint(const Plan* p, int k) {
    switch (k) {
        case kArena: return round_up((p->io_words * 3 + p->plane_words) * sizeof(uint32_t));
        case kPinIn: return round_up(p->io_words * 2 * sizeof(uint32_t));
        default: return round_up(p->io_words * sizeof(uint32_t));
    }
}

}  // namespace

bool bcrt_server::admit(std::unique_lock<std::mutex>& lk) {
    if (pending.empty()) return false;
    uint32_t slot = kInflight;
    for (uint32_t i = 0; i < kInflight; ++i)
        if (!slot_busy[i]) {
            slot = i;
            break;
        }
    if (slot == kInflight) return false;
    // Take pending requests in arrival order while all three rings have room;
    // group them by plan.
    std::vector<Plan*> order;
    std::map<Plan*, std::vector<Request*>> by_plan;
    size_t need[3] = {0, 0, 0};
    uint32_t taken = 0;
    for (Request* r : pending) {
        if (taken == max_wave) break;
        if (!by_plan.count(r->plan) && order.size() == bcrt::kMaxGroups) break;
        bool fits = true;
        for (int k = 0; k < 3 && fits; ++k) {
            size_t probe = 0;
            fits = ring_alloc(inflight, k, cap[k], need[k] + footprint(r->plan, k), &probe);
        }
        if (!fits) break;
        for (int k = 0; k < 3; ++k) need[k] += footprint(r->plan, k);
        if (!by_plan.count(r->plan)) order.push_back(r->plan);
        by_plan[r->plan].push_back(r);
        ++taken;
    }
    if (taken == 0) return false;
    Wave w;
    for (int k = 0; k < 3; ++k) {
        ring_alloc(inflight, k, cap[k], need[k], &w.start[k]);
        w.end[k] = w.start[k] + need[k];
    }
    pending.erase(pending.begin(), pending.begin() + taken);
    w.slot = slot;
    size_t off = w.start[kArena];
    uint32_t begin[5] = {0, 0, 0, 0, 0};
    for (Plan* p : order) {
        const auto& rs = by_plan[p];
        const uint32_t n = uint32_t(rs.size());
        const size_t io_bytes = p->io_words * sizeof(uint32_t) * n;
        uint32_t* da = reinterpret_cast<uint32_t*>(arena + off);
        uint32_t* db = reinterpret_cast<uint32_t*>(arena + off + io_bytes);
        uint32_t* dc = reinterpret_cast<uint32_t*>(arena + off + 2 * io_bytes);
        uint32_t* dp = reinterpret_cast<uint32_t*>(arena + off + round_up(3 * io_bytes));
        off += footprint(p, kArena) * n;
        bcrt::GroupDesc d{};
        d.t = p->t;
        d.args = bcrt::BatchArgs{da, db, dc, dp, n};
        for (int st = 0; st < 5; ++st) {
            d.begin[st] = begin[st];
            begin[st] += bcrt::group_workgroups(d.t, n, st);
        }
        w.groups.push_back(d);
        for (uint32_t i = 0; i < n; ++i) w.reqs.push_back(rs[i]);
    }
    slot_busy[slot] = true;
    lk.unlock();
    // Device work for the wave, without the lock: pack operands into pinned
    // staging, copy in on the copy stream, one mixed launch per stage on a
    // compute stream, copy results to pinned staging.
    try {
        hipStream_t cs = compute[next_compute++ & 1];
        bcrt::GroupDesc* table = d_groups + size_t(slot) * bcrt::kMaxGroups;
        hcheck(hipMemcpyAsync(table, w.groups.data(), sizeof(bcrt::GroupDesc) * w.groups.size(),
                              hipMemcpyHostToDevice, copy_in),
               "upload group table");
        size_t in_off = w.start[kPinIn], ri = 0;
        for (const auto& g : w.groups) {
            const size_t bytes = size_t(g.t.N) * g.t.L * sizeof(uint32_t);
            for (uint32_t i = 0; i < g.args.batch; ++i, ++ri) {
                Request* r = w.reqs[ri];
                uint8_t* h = pin_in + in_off;
                std::memcpy(h, r->a, bytes);
                std::memcpy(h + bytes, r->b, bytes);
                hcheck(hipMemcpyAsync(const_cast<uint32_t*>(g.args.a) + size_t(i) * g.t.N * g.t.L, h, bytes,
                                      hipMemcpyHostToDevice, copy_in),
                       "copy a");
                hcheck(hipMemcpyAsync(const_cast<uint32_t*>(g.args.b) + size_t(i) * g.t.N * g.t.L, h + bytes,
                                      bytes, hipMemcpyHostToDevice, copy_in),
                       "copy b");
                in_off += round_up(2 * bytes);
            }
        }
        hcheck(hipEventCreateWithFlags(&w.inputs_ready, hipEventDisableTiming), "wave event");
        hcheck(hipEventRecord(w.inputs_ready, copy_in), "record inputs");
        hcheck(hipStreamWaitEvent(cs, w.inputs_ready, 0), "compute waits inputs");
        hcheck(bcrt::launch_product_groups(table, w.groups.data(), uint32_t(w.groups.size()), cs), "launch wave");
        size_t out_off = w.start[kPinOut];
        for (const auto& g : w.groups) {
            const size_t bytes = size_t(g.t.N) * g.t.L * sizeof(uint32_t);
            for (uint32_t i = 0; i < g.args.batch; ++i) {
                hcheck(hipMemcpyAsync(pin_out + out_off, g.args.out + size_t(i) * g.t.N * g.t.L, bytes,
                                      hipMemcpyDeviceToHost, cs),
                       "copy c");
                out_off += round_up(bytes);
            }
        }
        hcheck(hipEventCreateWithFlags(&w.done, hipEventDisableTiming), "wave event");
        hcheck(hipEventRecord(w.done, cs), "record wave");
    } catch (const std::exception& x) {
        lk.lock();
        for (Request* r : w.reqs) r->error = x.what(), r->done = true;
        if (w.inputs_ready) (void)hipEventDestroy(w.inputs_ready);
        slot_busy[slot] = false;
        done_cv.notify_all();
        throw;
    }
    lk.lock();
    ++waves;
    max_wave_seen = std::max<uint64_t>(max_wave_seen, w.reqs.size());
    inflight.push_back(std::move(w));
    return true;
}

void bcrt_server::retire() {
    // Waves retire in admission order (the arena is a FIFO ring).
    while (!inflight.empty()) {
        Wave& w = inflight.front();
        const hipError_t q = hipEventQuery(w.done);
        if (q == hipErrorNotReady) break;
        size_t out_off = w.start[kPinOut], ri = 0;
        for (const auto& g : w.groups) {
            const size_t bytes = size_t(g.t.N) * g.t.L * sizeof(uint32_t);
            for (uint32_t i = 0; i < g.args.batch; ++i, ++ri) {
                if (q == hipSuccess) std::memcpy(w.reqs[ri]->c, pin_out + out_off, bytes);
                out_off += round_up(bytes);
            }
        }
        const auto now = Clock::now();
        for (Request* r : w.reqs) {
            r->latency = std::chrono::duration<double>(now - r->submitted).count();
            if (q != hipSuccess) r->error = std::string("ByteCRT server wave failed: ") + hipGetErrorString(q);
            r->done = true;
            ++completed;
        }
        (void)hipEventDestroy(w.done);
        (void)hipEventDestroy(w.inputs_ready);
        slot_busy[w.slot] = false;
        inflight.pop_front();
        done_cv.notify_all();
    }
}

void bcrt_server::run() {
    std::unique_lock<std::mutex> lk(m);
    while (!stop || !pending.empty() || !inflight.empty()) {
        try {
            retire();
            bool progressed = admit(lk);
            retire();
            if (!progressed) cv.wait_for(lk, std::chrono::microseconds(50));
        } catch (const std::exception& x) {
            if (!lk.owns_lock()) lk.lock();
            fatal = x.what();
            for (Request* r : pending) r->error = fatal, r->done = true;
            pending.clear();
            done_cv.notify_all();
            if (stop) break;
        }
    }
}

extern "C" {

int bcrt_server_create(bcrt_server** out, size_t hbm_budget_bytes, uint32_t max_wave, char* err, size_t errlen) {
    if (!out) return BCRT_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<bcrt_server>();
    try {
        if (hbm_budget_bytes < (size_t(1) << 20) || max_wave == 0)
            throw std::invalid_argument("ByteCRT server needs a budget >= 1 MiB and max_wave >= 1");
        s->budget = hbm_budget_bytes / kAlign * kAlign;
        s->cap[kArena] = s->budget;
        // Operands are at most half and results a quarter of a product's
        // device footprint (its residue planes outweigh the operands).
        s->cap[kPinIn] = s->budget / 2 / kAlign * kAlign;
        s->cap[kPinOut] = s->budget / 4 / kAlign * kAlign;
        s->max_wave = max_wave;
        hcheck(hipMalloc(reinterpret_cast<void**>(&s->arena), s->budget), "allocate arena");
        hcheck(hipHostMalloc(reinterpret_cast<void**>(&s->pin_in), s->cap[kPinIn], hipHostMallocDefault),
               "allocate pinned inputs");
        hcheck(hipHostMalloc(reinterpret_cast<void**>(&s->pin_out), s->cap[kPinOut], hipHostMallocDefault),
               "allocate pinned outputs");
        hcheck(hipMalloc(reinterpret_cast<void**>(&s->d_groups),
                         sizeof(bcrt::GroupDesc) * bcrt::kMaxGroups * kInflight),
               "allocate group tables");
        hcheck(hipStreamCreateWithFlags(&s->copy_in, hipStreamNonBlocking), "create stream");
        for (auto& c : s->compute) hcheck(hipStreamCreateWithFlags(&c, hipStreamNonBlocking), "create stream");
    } catch (const std::exception& x) {
        if (err && errlen) std::snprintf(err, errlen, "%s", x.what());
        if (s->arena) (void)hipFree(s->arena);
        if (s->pin_in) (void)hipHostFree(s->pin_in);
        if (s->pin_out) (void)hipHostFree(s->pin_out);
        if (s->d_groups) (void)hipFree(s->d_groups);
        for (auto c : s->compute)
            if (c) (void)hipStreamDestroy(c);
        if (s->copy_in) (void)hipStreamDestroy(s->copy_in);
        return dynamic_cast<const std::invalid_argument*>(&x) ? BCRT_EINVAL : BCRT_ERUNTIME;
    }
    bcrt_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return BCRT_OK;
}

int bcrt_server_submit(bcrt_server* s, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, const uint32_t* a,
                       const uint32_t* b, uint32_t* c, uint64_t* ticket) {
    if (!s || !q || !a || !b || !c || !ticket) return BCRT_EINVAL;
    Plan* plan = nullptr;
    {
        std::lock_guard<std::mutex> pg(s->plan_m);
        std::string key(reinterpret_cast<const char*>(q), size_t(L) * 4);
        key += "|" + std::to_string(N) + "|" + std::to_string(W);
        auto it = s->plans.find(key);
        if (it == s->plans.end()) {
            auto p = std::make_unique<Plan>();
            char msg[256] = {0};
            const int rc = bcrt_create(&p->engine, 1, N, W, L, q, msg, sizeof(msg));
            if (rc != BCRT_OK) {
                std::lock_guard<std::mutex> g(s->m);
                s->last_error = msg;
                return rc;
            }
            p->t = bcrt::engine_tables(p->engine);
            p->io_words = size_t(N) * L;
            p->plane_words = size_t(2) * p->t.P * N;
            it = s->plans.emplace(key, std::move(p)).first;
        }
        plan = it->second.get();
    }
    const size_t need = footprint(plan, kArena);
    std::lock_guard<std::mutex> g(s->m);
    if (need > s->budget || footprint(plan, kPinIn) > s->cap[kPinIn] || footprint(plan, kPinOut) > s->cap[kPinOut]) {
        s->last_error = "ByteCRT server: one product needs " + std::to_string(need) + " bytes, above the budget";
        return BCRT_EINVAL;
    }
    if (!s->fatal.empty()) {
        s->last_error = s->fatal;
        return BCRT_ERUNTIME;
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->plan = plan;
    r->a = a;
    r->b = b;
    r->c = c;
    r->submitted = Clock::now();
    s->pending.push_back(r.get());
    *ticket = r->id;
    s->requests.emplace(r->id, std::move(r));
    s->cv.notify_one();
    return BCRT_OK;
}

int bcrt_server_wait(bcrt_server* s, uint64_t ticket, double* latency_s) {
    if (!s) return BCRT_EINVAL;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) {
        s->last_error = "ByteCRT server: unknown ticket";
        return BCRT_EINVAL;
    }
    Request* r = it->second.get();
    s->done_cv.wait(lk, [r] { return r->done; });
    if (latency_s) *latency_s = r->latency;
    const bool ok = r->error.empty();
    if (!ok) s->last_error = r->error;
    s->requests.erase(it);
    return ok ? BCRT_OK : BCRT_ERUNTIME;
}

int bcrt_server_stats(bcrt_server* s, uint64_t* out) {
    if (!s || !out) return BCRT_EINVAL;
    std::lock_guard<std::mutex> g(s->m);
    out[0] = s->completed;
    out[1] = s->waves;
    out[2] = s->max_wave_seen;
    out[3] = s->plans.size();
    return BCRT_OK;
}

const char* bcrt_server_last_error(const bcrt_server* s) { return s ? s->last_error.c_str() : "null server"; }

void bcrt_server_destroy(bcrt_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto c : s->compute)
        if (c) (void)hipStreamSynchronize(c), (void)hipStreamDestroy(c);
    if (s->copy_in) (void)hipStreamDestroy(s->copy_in);
    if (s->arena) (void)hipFree(s->arena);
    if (s->pin_in) (void)hipHostFree(s->pin_in);
    if (s->pin_out) (void)hipHostFree(s->pin_out);
    if (s->d_groups) (void)hipFree(s->d_groups);
    delete s;
}

}  // extern "C"
