// Continuous-batching product server: requests of mixed geometry arrive at any
// time, are admitted into "waves" while the HBM budget, pinned staging and
// in-flight slots allow, and each wave runs as one mixed launch per stage
// (group kernels) on one of two compute streams, so a wave admitted while
// another is on the GPU overlaps with it.  Plan tables (twiddles and MFMA
// weights, no workspace) stay resident per (N, W, q) within a quarter of the
// HBM budget, least recently used idle plans evicted first; per-request device
// memory (operands, result, residue planes) comes from one arena, the other
// three quarters, used as a FIFO ring.
//
// Two host threads: the admit thread forms waves, packs operands into pinned
// staging (copy pool) and enqueues copies and kernels; the retire thread
// waits for the oldest wave, unpacks results (its own copy pool) and releases
// the wave's ring space.
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
#include "bytecrt_copy.hpp"
#include "bytecrt_kernels.hpp"
#include "bytecrt_sched.hpp"
#include "bytecrt_trace.hpp"

namespace bcrt {
// Tables and geometry of an engine (bytecrt_engine.cpp).
const DevTables& engine_tables(const bcrt_engine* e);
// An engine with the plan's device tables and no residue-plane workspace.
int create_tables_only(bcrt_engine** out, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, char* err,
                       size_t errlen);
size_t engine_table_bytes(const bcrt_engine* e);
// Primes of the plan for (N, W) (bytecrt_plan.cpp).
uint32_t plan_prime_count(uint32_t N, uint32_t W);
}

namespace {

using Clock = std::chrono::steady_clock;

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("ByteCRT server ") + what + ": " + hipGetErrorString(e));
}

constexpr uint32_t kInflight = 4;  // waves on the device at once
constexpr size_t kAlign = 256;

struct Plan {
    bcrt_engine* engine = nullptr;  // owns the resident tables (no workspace)
    bcrt::DevTables t{};
    size_t io_words = 0, plane_words = 0, table_bytes = 0;
    uint32_t refs = 0;      // requests submitted and not yet retired
    uint64_t last_use = 0;  // submit tick of the latest request
    ~Plan() { bcrt_destroy(engine); }
};

// Device bytes of a plan's tables before it is built: the two P x N twiddle
// tables plus an upper bound on the small ones (primes, weights, digit tiles).
size_t table_estimate(uint32_t N, uint32_t L, uint32_t P) {
    const size_t Kenc = (4 * L + 31) / 32 * 32, Kcrt = (4 * (P + 1) + 31) / 32 * 32;
    const size_t Mcrt = (4 * L + 1 + 31) / 32 * 32;
    return size_t(2) * P * N * 4 + size_t(P) * 96 + (P + 7) / 8 * 32 * Kenc + Mcrt * Kcrt + Mcrt * 2 + L * 4 +
           size_t(16) * 256;
}

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

// The three FIFO rings a wave takes space from (WaveScheduler's rings).
enum Ring { kArena = 0, kPinIn = 1, kPinOut = 2 };

struct Wave {
    std::vector<Request*> reqs;
    std::vector<bcrt::GroupDesc> groups;
    size_t start[3] = {0, 0, 0}, end[3] = {0, 0, 0};
    hipEvent_t inputs_ready = nullptr;
    hipEvent_t done = nullptr;
    bool enqueued = false;  // set under the lock once enqueue() returned
    uint64_t trace = 0;     // range from admission to retirement
    std::string error;      // enqueue() failed
};

size_t round_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

size_t footprint(size_t io_words, size_t plane_words, int k) {
    switch (k) {
        case kArena: return round_up((io_words * 3 + plane_words) * sizeof(uint32_t));
        case kPinIn: return round_up(io_words * 2 * sizeof(uint32_t));
        default: return round_up(io_words * sizeof(uint32_t));
    }
}

size_t footprint(const Plan* p, int k) { return footprint(p->io_words, p->plane_words, k); }

}  // namespace

struct bcrt_server {
    size_t cap[bcrt::kRings] = {0, 0, 0};  // arena (device), pinned inputs, pinned outputs
    uint32_t max_wave = 0;
    uint8_t* arena = nullptr;
    uint8_t* pin_in = nullptr;
    uint8_t* pin_out = nullptr;
    hipStream_t copy_in = nullptr;
    hipStream_t compute[2] = {nullptr, nullptr};
    uint32_t next_compute = 0;
    std::unique_ptr<bcrt::CopyPool> pack, unpack;
    std::mutex m;
    std::condition_variable admit_cv, retire_cv, done_cv;
    std::unique_ptr<bcrt::WaveScheduler> sched;  // pending queue and ring space
    std::deque<std::unique_ptr<Wave>> inflight;      // same waves, oldest first
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    std::map<std::string, std::unique_ptr<Plan>> plans;
    std::mutex plan_m;  // plans, table accounting; taken before m, never inside it
    std::condition_variable plan_cv;  // a plan's last request retired
    size_t table_cap = 0, table_used = 0;
    uint64_t plan_tick = 0, evictions = 0;
    uint64_t next_id = 1;
    bool stop = false, admit_done = false;
    std::string fatal;
    std::thread admitter, retirer;
    uint64_t completed = 0, waves = 0, max_wave_seen = 0;
    std::string last_error;

    void admit_loop();
    void retire_loop();
    std::unique_ptr<Wave> form_wave();
    void enqueue(Wave& w);
};

// Next wave from the scheduler (called with the lock held), with each plan's
// requests as one group in the device arena.
std::unique_ptr<Wave> bcrt_server::form_wave() {
    bcrt::SchedWave sw;
    if (!sched->form(&sw)) return nullptr;
    auto w = std::make_unique<Wave>();
    for (int k = 0; k < bcrt::kRings; ++k) w->start[k] = sw.start[k], w->end[k] = sw.end[k];
    size_t off = w->start[kArena];
    uint32_t begin[bcrt::kGroupStages] = {};
    for (size_t i = 0; i < sw.plans.size(); ++i) {
        const Plan* p = static_cast<const Plan*>(sw.plans[i]);
        const uint32_t n = uint32_t(sw.items[i].size());
        const size_t io_bytes = p->io_words * sizeof(uint32_t) * n;
        uint32_t* da = reinterpret_cast<uint32_t*>(arena + off);
        uint32_t* db = reinterpret_cast<uint32_t*>(arena + off + io_bytes);
        uint32_t* dc = reinterpret_cast<uint32_t*>(arena + off + 2 * io_bytes);
        uint32_t* dp = reinterpret_cast<uint32_t*>(arena + off + round_up(3 * io_bytes));
        off += footprint(p, kArena) * n;
        bcrt::GroupDesc d{};
        d.t = p->t;
        d.args = bcrt::BatchArgs{da, db, dc, dp, n};
        for (int st = 0; st < bcrt::kGroupStages; ++st) {
            d.begin[st] = begin[st];
            begin[st] += bcrt::group_workgroups(d.t, n, st);
        }
        w->groups.push_back(d);
        for (const auto& it : sw.items[i]) w->reqs.push_back(static_cast<Request*>(it.user));
    }
    return w;
}

// Packs the wave's operands into pinned staging and enqueues its copies and
// kernels (called without the lock).
void bcrt_server::enqueue(Wave& w) {
    hipStream_t cs = compute[next_compute++ & 1];
    size_t in_off = w.start[kPinIn], ri = 0;
    for (const auto& g : w.groups) {
        const size_t words = size_t(g.t.N) * g.t.L, bytes = words * sizeof(uint32_t);
        for (uint32_t i = 0; i < g.args.batch; ++i, ++ri) {
            const Request* r = w.reqs[ri];
            uint8_t* h = pin_in + in_off;
            pack->copy(h, r->a, bytes);
            pack->copy(h + bytes, r->b, bytes);
            hcheck(hipMemcpyAsync(const_cast<uint32_t*>(g.args.a) + i * words, h, bytes, hipMemcpyHostToDevice,
                                  copy_in),
                   "copy a");
            hcheck(hipMemcpyAsync(const_cast<uint32_t*>(g.args.b) + i * words, h + bytes, bytes,
                                  hipMemcpyHostToDevice, copy_in),
                   "copy b");
            in_off += round_up(2 * bytes);
        }
    }
    hcheck(hipEventCreateWithFlags(&w.inputs_ready, hipEventDisableTiming), "wave event");
    hcheck(hipEventRecord(w.inputs_ready, copy_in), "record inputs");
    hcheck(hipStreamWaitEvent(cs, w.inputs_ready, 0), "compute waits inputs");
    hcheck(bcrt::launch_product_groups(w.groups.data(), uint32_t(w.groups.size()), cs), "launch wave");
    size_t out_off = w.start[kPinOut];
    for (const auto& g : w.groups) {
        const size_t words = size_t(g.t.N) * g.t.L, bytes = words * sizeof(uint32_t);
        for (uint32_t i = 0; i < g.args.batch; ++i) {
            hcheck(hipMemcpyAsync(pin_out + out_off, g.args.out + i * words, bytes, hipMemcpyDeviceToHost, cs),
                   "copy c");
            out_off += round_up(bytes);
        }
    }
    hcheck(hipEventCreateWithFlags(&w.done, hipEventDisableTiming), "wave event");
    hcheck(hipEventRecord(w.done, cs), "record wave");
}

void bcrt_server::admit_loop() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        std::unique_ptr<Wave> w;
        admit_cv.wait(lk, [&] { return (w = form_wave()) != nullptr || (stop && sched->pending() == 0); });
        if (!w) break;
        Wave* raw = w.get();
        raw->trace = bcrt::trace_start("bcrt:wave");
        // Published before the device work is enqueued, so the next wave's ring
        // space is computed after this one's; the retire thread waits for
        // `enqueued` before touching it.
        inflight.push_back(std::move(w));
        ++waves;
        max_wave_seen = std::max<uint64_t>(max_wave_seen, raw->reqs.size());
        lk.unlock();
        std::string err;
        try {
            bcrt::TraceScope tr("bcrt:enqueue_wave");
            enqueue(*raw);
        } catch (const std::exception& x) {
            err = x.what();
        }
        lk.lock();
        raw->error = err;
        raw->enqueued = true;
        retire_cv.notify_one();
    }
    admit_done = true;
    retire_cv.notify_one();
}

void bcrt_server::retire_loop() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
        retire_cv.wait(lk, [&] { return (!inflight.empty() && inflight.front()->enqueued) ||
                                        (admit_done && inflight.empty()); });
        if (inflight.empty()) break;
        Wave& w = *inflight.front();
        std::string err = w.error;
        lk.unlock();
        if (err.empty()) {
            bcrt::TraceScope tr("bcrt:wait");
            const hipError_t q = hipEventSynchronize(w.done);
            if (q != hipSuccess) err = std::string("ByteCRT server wave failed: ") + hipGetErrorString(q);
        }
        if (err.empty()) {
            bcrt::TraceScope tr("bcrt:unpack");
            size_t out_off = w.start[kPinOut], ri = 0;
            for (const auto& g : w.groups) {
                const size_t bytes = size_t(g.t.N) * g.t.L * sizeof(uint32_t);
                for (uint32_t i = 0; i < g.args.batch; ++i, ++ri) {
                    unpack->copy(w.reqs[ri]->c, pin_out + out_off, bytes);
                    out_off += round_up(bytes);
                }
            }
        }
        const auto now = Clock::now();
        if (w.done) (void)hipEventDestroy(w.done);
        if (w.inputs_ready) (void)hipEventDestroy(w.inputs_ready);
        bcrt::trace_stop(w.trace);
        lk.lock();
        for (Request* r : w.reqs) {
            r->latency = std::chrono::duration<double>(now - r->submitted).count();
            r->error = err;
            r->done = true;
            ++completed;
        }
        if (!err.empty() && fatal.empty()) fatal = err;
        std::vector<Plan*> used;
        for (Request* r : w.reqs) used.push_back(r->plan);
        inflight.pop_front();
        sched->retire_oldest();
        done_cv.notify_all();
        admit_cv.notify_one();
        lk.unlock();
        {
            // The wave's plans may be evicted once no request holds them.
            std::lock_guard<std::mutex> pg(plan_m);
            for (Plan* p : used) --p->refs;
        }
        plan_cv.notify_all();
        lk.lock();
    }
}

// The plan for (q, N, W) with one more request holding it, built (and room
// made for its tables by evicting idle plans, least recently used first) if
// it is not resident.  Blocks while every plan that could make room is in
// use.  Returns a status code; *out is the plan on success.
static int acquire_plan(bcrt_server* s, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, uint32_t P,
                        Plan** out) {
    std::unique_lock<std::mutex> pg(s->plan_m);
    std::string key(reinterpret_cast<const char*>(q), size_t(L) * 4);
    key += "|" + std::to_string(N) + "|" + std::to_string(W);
    auto it = s->plans.find(key);
    if (it == s->plans.end()) {
        const size_t need = table_estimate(N, L, P);
        if (need > s->table_cap) {
            std::lock_guard<std::mutex> g(s->m);
            s->last_error = "ByteCRT server: the plan tables need " + std::to_string(need) +
                            " bytes of HBM, above the budget's table share of " + std::to_string(s->table_cap);
            return BCRT_EINVAL;
        }
        for (;;) {
            if (s->table_used + need <= s->table_cap) break;
            Plan* lru = nullptr;
            std::string lru_key;
            for (auto& kv : s->plans)
                if (kv.second->refs == 0 && (!lru || kv.second->last_use < lru->last_use)) {
                    lru = kv.second.get();
                    lru_key = kv.first;
                }
            if (lru) {
                s->table_used -= lru->table_bytes;
                s->plans.erase(lru_key);  // frees the tables (no request uses them)
                ++s->evictions;
                continue;
            }
            s->plan_cv.wait(pg);  // every resident plan is in use: wait for one to go idle
            it = s->plans.find(key);  // another submitter may have built it meanwhile
            if (it != s->plans.end()) break;
        }
        if (it == s->plans.end()) {
            auto p = std::make_unique<Plan>();
            char msg[256] = {0};
            const int rc = bcrt::create_tables_only(&p->engine, N, W, L, q, msg, sizeof(msg));
            if (rc != BCRT_OK) {
                std::lock_guard<std::mutex> g(s->m);
                s->last_error = msg;
                return rc;
            }
            p->t = bcrt::engine_tables(p->engine);
            p->io_words = size_t(N) * L;
            p->plane_words = size_t(2) * p->t.P * N;
            p->table_bytes = bcrt::engine_table_bytes(p->engine);
            s->table_used += p->table_bytes;
            it = s->plans.emplace(key, std::move(p)).first;
        }
    }
    Plan* plan = it->second.get();
    ++plan->refs;
    plan->last_use = ++s->plan_tick;
    *out = plan;
    return BCRT_OK;
}

extern "C" {

int bcrt_server_create(bcrt_server** out, size_t hbm_budget_bytes, uint32_t max_wave, char* err, size_t errlen) {
    if (!out) return BCRT_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<bcrt_server>();
    try {
        if (hbm_budget_bytes < (size_t(1) << 20) || max_wave == 0)
            throw std::invalid_argument("ByteCRT server needs a budget >= 1 MiB and max_wave >= 1");
        // Three quarters of the budget are the request arena, one quarter holds
        // resident plan tables.
        s->table_cap = hbm_budget_bytes / 4;
        s->cap[kArena] = (hbm_budget_bytes - s->table_cap) / kAlign * kAlign;
        // Operands are at most half and results a quarter of a product's
        // device footprint (its residue planes outweigh the operands).
        s->cap[kPinIn] = s->cap[kArena] / 2 / kAlign * kAlign;
        s->cap[kPinOut] = s->cap[kArena] / 4 / kAlign * kAlign;
        s->max_wave = max_wave;
        hcheck(hipMalloc(reinterpret_cast<void**>(&s->arena), s->cap[kArena]), "allocate arena");
        hcheck(hipHostMalloc(reinterpret_cast<void**>(&s->pin_in), s->cap[kPinIn], hipHostMallocDefault),
               "allocate pinned inputs");
        hcheck(hipHostMalloc(reinterpret_cast<void**>(&s->pin_out), s->cap[kPinOut], hipHostMallocDefault),
               "allocate pinned outputs");
        hcheck(hipStreamCreateWithFlags(&s->copy_in, hipStreamNonBlocking), "create stream");
        for (auto& c : s->compute) hcheck(hipStreamCreateWithFlags(&c, hipStreamNonBlocking), "create stream");
        s->pack = std::make_unique<bcrt::CopyPool>(bcrt::server_pool_workers());
        s->unpack = std::make_unique<bcrt::CopyPool>(bcrt::server_pool_workers());
        s->sched = std::make_unique<bcrt::WaveScheduler>(s->cap, max_wave, bcrt::kMaxGroups, kInflight);
    } catch (const std::exception& x) {
        if (err && errlen) std::snprintf(err, errlen, "%s", x.what());
        if (s->arena) (void)hipFree(s->arena);
        if (s->pin_in) (void)hipHostFree(s->pin_in);
        if (s->pin_out) (void)hipHostFree(s->pin_out);
        for (auto c : s->compute)
            if (c) (void)hipStreamDestroy(c);
        if (s->copy_in) (void)hipStreamDestroy(s->copy_in);
        return dynamic_cast<const std::invalid_argument*>(&x) ? BCRT_EINVAL : BCRT_ERUNTIME;
    }
    bcrt_server* raw = s.release();
    raw->admitter = std::thread([raw] { raw->admit_loop(); });
    raw->retirer = std::thread([raw] { raw->retire_loop(); });
    *out = raw;
    return BCRT_OK;
}

int bcrt_server_submit(bcrt_server* s, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, const uint32_t* a,
                       const uint32_t* b, uint32_t* c, uint64_t* ticket) {
    if (!s || !q || !a || !b || !c || !ticket) return BCRT_EINVAL;
    // Geometry and the request's footprint first: nothing is built for a
    // request that could never be admitted.
    uint32_t P = 0;
    try {
        if (N < 4096 || N > 65536 || (N & (N - 1)) || W < 64 || L != (W + 31) / 32 || L > 64)
            throw std::invalid_argument("ByteCRT plan: need N = 2^k in [4096, 65536], 64 <= W <= 2048, L = ceil(W/32)");
        P = bcrt::plan_prime_count(N, W);
    } catch (const std::exception& x) {
        std::lock_guard<std::mutex> g(s->m);
        s->last_error = x.what();
        return dynamic_cast<const std::invalid_argument*>(&x) ? BCRT_EINVAL : BCRT_ERUNTIME;
    }
    bcrt::SchedItem item;
    for (int k = 0; k < bcrt::kRings; ++k) item.need[k] = footprint(size_t(N) * L, size_t(2) * P * N, k);
    {
        std::lock_guard<std::mutex> g(s->m);
        if (!s->sched->fits_alone(item)) {
            s->last_error = "ByteCRT server: one product needs " + std::to_string(item.need[kArena]) +
                            " bytes of HBM, above the budget";
            return BCRT_EINVAL;
        }
        if (!s->fatal.empty()) {
            s->last_error = s->fatal;
            return BCRT_ERUNTIME;
        }
    }
    Plan* plan = nullptr;
    const int rc = acquire_plan(s, N, W, L, q, P, &plan);
    if (rc != BCRT_OK) return rc;
    item.plan = plan;
    std::lock_guard<std::mutex> g(s->m);
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->plan = plan;
    r->a = a;
    r->b = b;
    r->c = c;
    r->submitted = Clock::now();
    item.user = r.get();
    s->sched->push(item);
    *ticket = r->id;
    s->requests.emplace(r->id, std::move(r));
    s->admit_cv.notify_one();
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
    size_t nplans, table_used, table_cap;
    uint64_t evictions;
    {
        std::lock_guard<std::mutex> pg(s->plan_m);
        nplans = s->plans.size();
        table_used = s->table_used;
        table_cap = s->table_cap;
        evictions = s->evictions;
    }
    std::lock_guard<std::mutex> g(s->m);
    out[0] = s->completed;
    out[1] = s->waves;
    out[2] = s->max_wave_seen;
    out[3] = nplans;
    out[4] = evictions;
    out[5] = table_used;
    out[6] = table_cap;
    out[7] = s->cap[kArena];
    return BCRT_OK;
}

const char* bcrt_server_last_error(const bcrt_server* s) { return s ? s->last_error.c_str() : "null server"; }

void bcrt_server_destroy(bcrt_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
    }
    s->admit_cv.notify_all();
    if (s->admitter.joinable()) s->admitter.join();
    if (s->retirer.joinable()) s->retirer.join();
    for (auto c : s->compute)
        if (c) (void)hipStreamSynchronize(c), (void)hipStreamDestroy(c);
    if (s->copy_in) (void)hipStreamDestroy(s->copy_in);
    if (s->arena) (void)hipFree(s->arena);
    if (s->pin_in) (void)hipHostFree(s->pin_in);
    if (s->pin_out) (void)hipHostFree(s->pin_out);
    delete s;
}

}  // extern "C"
