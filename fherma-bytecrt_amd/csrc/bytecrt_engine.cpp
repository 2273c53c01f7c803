// C ABI of the ByteCRT engine: plan upload, workspace, whole-call and
// device-resident runs, graph replay.
#include "../../include/bytecrt.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <condition_variable>
#include <functional>
#include <thread>

#include "bytecrt_io.hpp"
#include "bytecrt_kernels.hpp"
#include "bytecrt_plan.hpp"

namespace {

}  // namespace

namespace bcrt {
hipError_t configure_kernels();
}

// Fork/join events of one pipelined launch sequence.  Eager runs and each
// captured graph use different sets: an event object is never recorded into a
// capture while an eager record of it may still be pending.
struct EventSet {
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
};

struct bcrt_engine {
    bcrt::HostPlan plan;
    bcrt::DevTables tables{};
    uint32_t max_batch = 0;  // 0: tables only, no residue-plane workspace
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    size_t table_bytes = 0;      // device bytes of the plan tables
    uint32_t* planes = nullptr;  // [max_batch][2][P][N]
    // Whole-call host I/O (pinned staging, copy pools, drain thread), created
    // on the first whole call.
    std::unique_ptr<bcrt::HostIo> hostio;
    // Captured graph (one key) with the events its capture recorded, and an
    // event recorded behind every launch of it so it is destroyed only once
    // its last launch has finished.
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    EventSet graph_events;
    hipEvent_t graph_done = nullptr;
    bool graph_launched = false;
    const void* graph_key[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t graph_batch = 0;
    // Batch pipelining: chunks of `pipe_chunk` products round-robin over worker
    // streams so that stages of different chunks overlap on the device.
    std::vector<hipStream_t> workers;
    EventSet eager_events;
    uint32_t pipe_chunk = 32;  // near the best of every B=128 sweep; tune.py records per geometry
    uint32_t pipe_streams = 2;
    // Mixed batches that hold this engine's tables and workspace; while any
    // does, bcrt_destroy only marks the engine and the last mix frees it.
    uint32_t mix_refs = 0;
    bool destroy_pending = false;
    std::string last_error;
};

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string("ByteCRT ") + what + ": " + hipGetErrorString(e));
}

template <class T>
T* upload(bcrt_engine& s, const std::vector<T>& v) {
    void* p = nullptr;
    check(hipMalloc(&p, v.size() * sizeof(T)), "allocate table");
    s.allocations.push_back(p);
    s.table_bytes += v.size() * sizeof(T);
    check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload table");
    return static_cast<T*>(p);
}

int fail(bcrt_engine* e, int code, const std::string& msg) {
    if (e) e->last_error = msg;
    return code;
}

// Runs f, mapping exceptions to status codes.
template <class F>
int guarded(bcrt_engine* e, F&& f) {
    try {
        f();
        if (e) e->last_error.clear();
        return BCRT_OK;
    } catch (const std::invalid_argument& x) {
        return fail(e, BCRT_EINVAL, x.what());
    } catch (const std::exception& x) {
        return fail(e, BCRT_ERUNTIME, x.what());
    }
}

hipStream_t pick(bcrt_engine* e, void* stream) {
    return stream ? static_cast<hipStream_t>(stream) : e->stream;
}

void require_batch(const bcrt_engine* e, uint32_t batch) {
    if (batch == 0 || batch > e->max_batch)
        throw std::invalid_argument("ByteCRT batch must be in [1, max_batch=" +
                                    std::to_string(e->max_batch) + "]");
}

void destroy_events(EventSet& ev) {
    for (auto x : ev.joins) (void)hipEventDestroy(x);
    ev.joins.clear();
    if (ev.fork) (void)hipEventDestroy(ev.fork);
    ev.fork = nullptr;
}

// Grows an event set to `n` join events (and its fork event).
void ensure_events(EventSet& ev, size_t n) {
    if (!ev.fork) check(hipEventCreateWithFlags(&ev.fork, hipEventDisableTiming), "create fork event");
    while (ev.joins.size() < n) {
        hipEvent_t x = nullptr;
        check(hipEventCreateWithFlags(&x, hipEventDisableTiming), "create join event");
        ev.joins.push_back(x);
    }
}

// Waits for the last launch of the captured graph, then destroys it with the
// events its capture recorded.
void drop_graph(bcrt_engine* e) {
    if (e->graph_launched) (void)hipEventSynchronize(e->graph_done);
    e->graph_launched = false;
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    destroy_events(e->graph_events);
    e->graph_exec = nullptr;
    e->graph = nullptr;
    e->graph_batch = 0;
}

// Waits for everything the engine enqueued on its own streams.
void quiesce(bcrt_engine* e) {
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto st : e->workers) (void)hipStreamSynchronize(st);
}

void release(bcrt_engine* e) {
    drop_graph(e);
    e->hostio.reset();
    quiesce(e);
    destroy_events(e->eager_events);
    for (auto st : e->workers) (void)hipStreamDestroy(st);
    e->workers.clear();
    if (e->graph_done) (void)hipEventDestroy(e->graph_done);
    e->graph_done = nullptr;
    for (void* p : e->allocations) (void)hipFree(p);
    e->allocations.clear();
    if (e->stream) (void)hipStreamDestroy(e->stream);
    e->stream = nullptr;
}

void ensure_workers(bcrt_engine* e) {
    while (e->workers.size() < e->pipe_streams) {
        hipStream_t st = nullptr;
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "create worker stream");
        e->workers.push_back(st);
    }
}

bool pipelined(const bcrt_engine* e, uint32_t batch) { return batch > e->pipe_chunk && e->pipe_streams > 1; }

// Enqueue the product for a batch on `origin`: one launch sequence when the
// batch fits a chunk, else chunks spread over the worker streams, forked from
// and joined back into `origin` with the events of `ev` (also valid inside a
// stream capture).
void enqueue_batch(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                   uint32_t batch, hipStream_t origin, EventSet& ev) {
    const auto& g = e->plan.g;
    if (!e->planes) throw std::invalid_argument("ByteCRT engine has no residue-plane workspace");
    if (!pipelined(e, batch)) {
        const bcrt::BatchArgs args{d_a, d_b, d_c, e->planes, batch};
        check(bcrt::launch_product(e->tables, args, origin), "launch product");
        return;
    }
    ensure_workers(e);
    const size_t io = size_t(g.N) * g.L, pl = size_t(2) * g.P * g.N;
    const uint32_t used = std::min<uint32_t>(e->pipe_streams, (batch + e->pipe_chunk - 1) / e->pipe_chunk);
    ensure_events(ev, used);
    check(hipEventRecord(ev.fork, origin), "fork");
    for (uint32_t k = 0; k < used; ++k) check(hipStreamWaitEvent(e->workers[k], ev.fork, 0), "fork wait");
    uint32_t i = 0;
    for (uint32_t c0 = 0; c0 < batch; c0 += e->pipe_chunk, ++i) {
        const uint32_t n = std::min(e->pipe_chunk, batch - c0);
        const bcrt::BatchArgs args{d_a + c0 * io, d_b + c0 * io, d_c + c0 * io, e->planes + c0 * pl, n};
        check(bcrt::launch_product(e->tables, args, e->workers[i % used]), "launch chunk");
    }
    for (uint32_t k = 0; k < used; ++k) {
        check(hipEventRecord(ev.joins[k], e->workers[k]), "join record");
        check(hipStreamWaitEvent(origin, ev.joins[k], 0), "join wait");
    }
}

}  // namespace

namespace {

int run_whole(bcrt_engine* e, const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t batch, bcrt_sink sink,
              void* ctx) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        require_batch(e, batch);
        if (!a || !b || (!c && !sink)) throw std::invalid_argument("ByteCRT null host pointer");
        const auto& g = e->plan.g;
        if (!e->hostio) e->hostio = std::make_unique<bcrt::HostIo>(size_t(g.N) * g.L, e->max_batch, e->stream);
        e->hostio->run(a, b, batch, c, sink, ctx,
                       [e](const uint32_t* da, const uint32_t* db, uint32_t* dc, uint32_t n, hipStream_t st) {
                           enqueue_batch(e, da, db, dc, n, st, e->eager_events);
                       });
    });
}

}  // namespace

namespace {

// Global lock for the engine <-> mix reference counts.
std::mutex g_refs;

// Builds an engine: plan, device tables and, when max_batch > 0, the residue-
// plane workspace for max_batch products (max_batch 0: tables only, for the
// server, whose requests bring their planes from its arena).
int create_engine(bcrt_engine** out, uint32_t max_batch, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q,
                  char* err, size_t errlen) {
    if (!out) return BCRT_EINVAL;
    *out = nullptr;
    auto* e = new bcrt_engine();
    const int rc = guarded(e, [&] {
        e->plan = bcrt::make_plan(N, W, L, q);
        const auto& g = e->plan.g;
        e->max_batch = max_batch;
        check(hipGetDevice(&e->device), "get device");
        check(bcrt::configure_kernels(), "configure kernels");
        // A blocking stream: work enqueued here is ordered with the legacy
        // default stream, which is where callers' allocations and copies land.
        check(hipStreamCreate(&e->stream), "create stream");
        check(hipEventCreateWithFlags(&e->graph_done, hipEventDisableTiming), "create graph event");
        bcrt::DevTables& t = e->tables;
        t.N = g.N;
        t.W = g.W;
        t.L = g.L;
        t.P = g.P;
        t.N1 = g.N1;
        t.n1 = g.n1;
        t.N2 = g.N2;
        t.Kenc = g.Kenc;
        t.PTenc = g.PTenc;
        t.Kcrt = g.Kcrt;
        t.Mcrt = g.Mcrt;
        t.qshift = e->plan.qshift;
        t.enc_ts_log = std::min(5u, g.n1);
        t.fuse_min_tiles = bcrt::kFusedMinTiles;
        t.kflags = e->plan.tw5_late ? bcrt::kPlanTw5Late : 0;
        t.gap = e->plan.gap;
        t.qt1 = e->plan.qt1;
        t.inv_qt1 = e->plan.inv_qt1;
        t.pd = upload(*e, e->plan.pd);
        t.zf = upload(*e, e->plan.zf);
        t.zi = upload(*e, e->plan.zi);
        t.scale = upload(*e, e->plan.scale);
        t.scale_z1 = upload(*e, e->plan.scale_z1);
        t.zf16 = upload(*e, e->plan.zf16);
        t.enc_prime = upload(*e, e->plan.enc_prime);
        t.wenc = upload(*e, e->plan.wenc);
        t.wcrt = upload(*e, e->plan.wcrt);
        t.corr = upload(*e, e->plan.corr);
        t.invp = upload(*e, e->plan.invp);
        t.q = upload(*e, e->plan.q);
        // The host copies of the big tables are no longer needed.
        std::vector<uint32_t>().swap(e->plan.zf);
        std::vector<uint32_t>().swap(e->plan.zi);
        if (max_batch > 0) {
            void* p = nullptr;
            check(hipMalloc(&p, size_t(max_batch) * 2 * g.P * g.N * sizeof(uint32_t)), "allocate residue planes");
            e->allocations.push_back(p);
            e->planes = static_cast<uint32_t*>(p);
        }
    });
    if (rc != BCRT_OK) {
        if (err && errlen) {
            std::snprintf(err, errlen, "%s", e->last_error.c_str());
        }
        release(e);
        delete e;
        return rc;
    }
    *out = e;
    return BCRT_OK;
}

// Captures the batch's launch sequence into the engine's graph, keyed by
// (pointers, batch).  Nothing the engine enqueued earlier is still pending
// when the capture starts, and the capture records its own fork/join events.
void capture_graph(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch) {
    drop_graph(e);
    if (pipelined(e, batch)) ensure_workers(e);
    quiesce(e);
    check(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    std::string failure;
    try {
        enqueue_batch(e, d_a, d_b, d_c, batch, e->stream, e->graph_events);
    } catch (const std::exception& x) {
        failure = x.what();
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (!failure.empty()) {
        if (g) (void)hipGraphDestroy(g);
        throw std::runtime_error(failure + " (during capture)");
    }
    check(ce, "end capture");
    e->graph = g;
    check(hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0), "instantiate graph");
    e->graph_key[0] = d_a;
    e->graph_key[1] = d_b;
    e->graph_key[2] = d_c;
    e->graph_batch = batch;
}

}  // namespace

namespace bcrt {

int create_tables_only(bcrt_engine** out, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, char* err,
                       size_t errlen) {
    return create_engine(out, 0, N, W, L, q, err, errlen);
}

size_t engine_table_bytes(const bcrt_engine* e) { return e->table_bytes; }

}  // namespace bcrt

extern "C" {

int bcrt_create(bcrt_engine** out, uint32_t max_batch, uint32_t N, uint32_t W, uint32_t L,
                const uint32_t* q, char* err, size_t errlen) {
    if (out) *out = nullptr;
    if (out && max_batch == 0) {
        if (err && errlen) std::snprintf(err, errlen, "%s", "ByteCRT max_batch must be positive");
        return BCRT_EINVAL;
    }
    return create_engine(out, max_batch, N, W, L, q, err, errlen);
}

void bcrt_destroy(bcrt_engine* e) {
    if (!e) return;
    {
        std::lock_guard<std::mutex> g(g_refs);
        if (e->mix_refs > 0) {
            e->destroy_pending = true;  // the last mix that holds it frees it
            return;
        }
    }
    release(e);
    delete e;
}

int bcrt_run_device(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                    uint32_t batch, void* stream) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        require_batch(e, batch);
        if (!d_a || !d_b || !d_c) throw std::invalid_argument("ByteCRT null device pointer");
        enqueue_batch(e, d_a, d_b, d_c, batch, pick(e, stream), e->eager_events);
    });
}

int bcrt_run_device_graph(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                          uint32_t batch, void* stream) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        require_batch(e, batch);
        if (!d_a || !d_b || !d_c) throw std::invalid_argument("ByteCRT null device pointer");
        hipStream_t s = pick(e, stream);
        const bool hit = e->graph_exec && e->graph_batch == batch && e->graph_key[0] == d_a &&
                         e->graph_key[1] == d_b && e->graph_key[2] == d_c;
        // Capture on the engine's own stream, then launch on the caller's.
        if (!hit) capture_graph(e, d_a, d_b, d_c, batch);
        check(hipGraphLaunch(e->graph_exec, s), "graph launch");
        check(hipEventRecord(e->graph_done, s), "record graph launch");
        e->graph_launched = true;
    });
}

int bcrt_run(bcrt_engine* e, const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t batch) {
    return run_whole(e, a, b, c, batch, nullptr, nullptr);
}

int bcrt_run_sink(bcrt_engine* e, const uint32_t* a, const uint32_t* b, uint32_t batch, bcrt_sink sink, void* ctx) {
    if (!sink) return BCRT_EINVAL;
    return run_whole(e, a, b, nullptr, batch, sink, ctx);
}

int bcrt_synchronize(bcrt_engine* e) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] { check(hipStreamSynchronize(e->stream), "synchronize"); });
}

const char* bcrt_last_error(const bcrt_engine* e) { return e ? e->last_error.c_str() : "null engine"; }

int bcrt_geometry(const bcrt_engine* e, uint32_t* out) {
    if (!e || !out) return BCRT_EINVAL;
    const auto& g = e->plan.g;
    const uint32_t v[10] = {g.N, g.W, g.L, g.P, g.N1, g.N2, g.Kenc, g.Kcrt, g.Mcrt, g.product_bits};
    std::memcpy(out, v, sizeof(v));
    return BCRT_OK;
}

int bcrt_plan_tables(const bcrt_engine* e, uint32_t* primes, uint32_t* psi, uint32_t* crt_inverse,
                     uint32_t* weights) {
    if (!e) return BCRT_EINVAL;
    const auto& pl = e->plan;
    if (primes) std::memcpy(primes, pl.primes.data(), pl.primes.size() * 4);
    if (psi) std::memcpy(psi, pl.psi.data(), pl.psi.size() * 4);
    if (crt_inverse) std::memcpy(crt_inverse, pl.crt_inverse.data(), pl.crt_inverse.size() * 4);
    if (weights) std::memcpy(weights, pl.weights.data(), pl.weights.size() * 4);
    return BCRT_OK;
}

int bcrt_run_stages(bcrt_engine* e, uint32_t stage_mask, const uint32_t* d_a, const uint32_t* d_b,
                    uint32_t* d_c, uint32_t* d_planes, uint32_t batch, void* stream) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        require_batch(e, batch);
        if ((stage_mask & 1) && (!d_a || !d_b)) throw std::invalid_argument("ByteCRT encode needs a and b");
        if ((stage_mask & 16) && !d_c) throw std::invalid_argument("ByteCRT CRT needs c");
        if (!d_planes && !e->planes) throw std::invalid_argument("ByteCRT engine has no residue-plane workspace");
        const bcrt::BatchArgs args{d_a, d_b, d_c, d_planes ? d_planes : e->planes, batch};
        hipStream_t s = pick(e, stream);
        if ((stage_mask & 3) == 3) {
            check(bcrt::launch_encode_forward(e->tables, args, s), "encode and forward columns");
        } else {
            if (stage_mask & 1) check(bcrt::launch_encode(e->tables, args, s), "encode");
            if (stage_mask & 2) check(bcrt::launch_forward_columns(e->tables, args, s), "forward columns");
        }
        if (stage_mask & 4) check(bcrt::launch_middle(e->tables, args, s), "middle");
        if (stage_mask & 8) check(bcrt::launch_inverse_columns(e->tables, args, s), "inverse columns");
        if (stage_mask & 16) check(bcrt::launch_crt(e->tables, args, s), "crt");
    });
}

size_t bcrt_planes_words(const bcrt_engine* e, uint32_t batch) {
    return e ? size_t(2) * e->plan.g.P * e->plan.g.N * batch : 0;
}

int bcrt_mfma_probe(const int8_t* d_a, const int8_t* d_b, int32_t* d_c) {
    if (!d_a || !d_b || !d_c) return BCRT_EINVAL;
    if (bcrt::launch_mfma_probe(d_a, d_b, d_c, nullptr) != hipSuccess) return BCRT_ERUNTIME;
    return hipDeviceSynchronize() == hipSuccess ? BCRT_OK : BCRT_ERUNTIME;
}

}  // extern "C"

extern "C" int bcrt_plan_host(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, uint32_t* geometry,
                              uint32_t* primes, uint32_t* psi, uint32_t* crt_inverse, uint32_t* weights,
                              char* err, size_t errlen) {
    try {
        const bcrt::HostPlan pl = bcrt::make_plan(N, W, L, q);
        const auto& g = pl.g;
        if (geometry) {
            const uint32_t v[10] = {g.N, g.W, g.L, g.P, g.N1, g.N2, g.Kenc, g.Kcrt, g.Mcrt, g.product_bits};
            std::memcpy(geometry, v, sizeof(v));
        }
        if (primes) std::memcpy(primes, pl.primes.data(), pl.primes.size() * 4);
        if (psi) std::memcpy(psi, pl.psi.data(), pl.psi.size() * 4);
        if (crt_inverse) std::memcpy(crt_inverse, pl.crt_inverse.data(), pl.crt_inverse.size() * 4);
        if (weights) std::memcpy(weights, pl.weights.data(), pl.weights.size() * 4);
        return BCRT_OK;
    } catch (const std::invalid_argument& x) {
        if (err && errlen) std::snprintf(err, errlen, "%s", x.what());
        return BCRT_EINVAL;
    } catch (const std::exception& x) {
        if (err && errlen) std::snprintf(err, errlen, "%s", x.what());
        return BCRT_ERUNTIME;
    }
}

extern "C" int bcrt_near_power_gap(uint32_t W, uint32_t L, const uint32_t* q, uint32_t* gap) {
    if (!q || !gap) return BCRT_EINVAL;
    *gap = bcrt::near_power_gap(W, L, q);
    return BCRT_OK;
}

extern "C" int bcrt_set_pipeline(bcrt_engine* e, uint32_t chunk, uint32_t streams) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        if (chunk == 0 || streams == 0 || streams > 16)
            throw std::invalid_argument("ByteCRT pipeline needs chunk >= 1 and 1 <= streams <= 16");
        drop_graph(e);
        e->pipe_chunk = chunk;
        e->pipe_streams = streams;
    });
}

extern "C" int bcrt_set_encode_tile(bcrt_engine* e, uint32_t ts_log) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        const uint32_t n1 = e->tables.n1;
        if (ts_log > 7 || ts_log > n1)
            throw std::invalid_argument("ByteCRT encode tile needs 2^ts_log <= min(128, N/256)");
        drop_graph(e);  // captured graphs hold the tables by value
        e->tables.enc_ts_log = ts_log;
    });
}

extern "C" int bcrt_set_kernel_flags(bcrt_engine* e, uint32_t flags) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        if (flags > bcrt::kFlagsAll) throw std::invalid_argument("ByteCRT kernel flags must be within 0..127");
        drop_graph(e);  // captured graphs hold the tables by value
        e->tables.kflags = flags | (e->tables.kflags & bcrt::kPlanFlags);
    });
}

extern "C" int bcrt_set_fusion(bcrt_engine* e, uint32_t min_tiles) {
    if (!e) return BCRT_EINVAL;
    return guarded(e, [&] {
        drop_graph(e);  // captured graphs hold the tables by value
        e->tables.fuse_min_tiles = min_tiles;
    });
}

// ---------------------------------------------------------------------------
// Mixed batches
// ---------------------------------------------------------------------------
// Groups are classed by geometry (N, W); the products of each class are dealt
// into chunks (each group split evenly).  A chunk is one mixed launch per
// stage over its products, and chunks go round robin onto worker streams, so
// stages of different chunks and classes may overlap.  The schedule (chunks
// per class, streams, or every class in the same chunks) is set per mix;
// results do not depend on it.  The default, one chunk per class and one
// stream, was the fastest measured at 96 x (N=32,768) + 32 x (N=65,536)
// products: the kernels fill the device by themselves, and concurrent
// chunks only share it.
constexpr uint32_t kMixMaxChunks = 16;  // chunks per class, streams a schedule may ask for

struct bcrt_mix {
    std::vector<bcrt_group> groups;  // as created
    uint32_t sched_chunks = 1;       // chunks per class
    uint32_t sched_streams = 1;      // worker streams (at most one per chunk)
    uint32_t sched_flags = 0;        // BCRT_MIX_MERGE_CLASSES, BCRT_MIX_PERSISTENT
    std::vector<std::vector<bcrt::GroupDesc>> chunks;
    uint32_t* queues = nullptr;      // BCRT_MIX_PERSISTENT: kQueueWords counters per chunk
    std::vector<bcrt_engine*> engines;  // referenced (mix_refs) for the mix's lifetime
    hipStream_t stream = nullptr;
    hipEvent_t graph_done = nullptr;  // recorded behind every graph launch
    bool graph_launched = false;
    std::vector<hipStream_t> workers;
    EventSet eager_events, graph_events;  // fork/join events: eager runs, the capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::string last_error;
};

namespace {

// Waits for the mix's last launches, then frees its graph, fork/join events
// and worker streams (the layout of one schedule).
void drop_layout(bcrt_mix* m) {
    if (m->graph_launched) (void)hipEventSynchronize(m->graph_done);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (auto st : m->workers) (void)hipStreamSynchronize(st);
    m->graph_launched = false;
    if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->graph_exec = nullptr;
    m->graph = nullptr;
    destroy_events(m->eager_events);
    destroy_events(m->graph_events);
    for (auto st : m->workers) (void)hipStreamDestroy(st);
    m->workers.clear();
    m->chunks.clear();
    if (m->queues) (void)hipFree(m->queues);
    m->queues = nullptr;
}

// drop_layout, then the mix's own stream and event.
void release_mix(bcrt_mix* m) {
    drop_layout(m);
    if (m->graph_done) (void)hipEventDestroy(m->graph_done);
    m->graph_done = nullptr;
    if (m->stream) (void)hipStreamDestroy(m->stream);
    m->stream = nullptr;
}

// Deals the mix's groups into chunks under its schedule and creates the
// worker streams and events they need.
void layout_mix(bcrt_mix* m) {
    const auto& groups = m->groups;
    const uint32_t ngroups = uint32_t(groups.size());
    // Groups of one geometry (same kernel resource needs) share chunks unless
    // the schedule merges every class into the same chunks.
    std::vector<std::vector<uint32_t>> classes;
    for (uint32_t k = 0; k < ngroups; ++k) {
        const auto& t = groups[k].engine->tables;
        bool placed = false;
        for (auto& cl : classes) {
            const auto& u = groups[cl[0]].engine->tables;
            if ((m->sched_flags & BCRT_MIX_MERGE_CLASSES) || (u.N == t.N && u.W == t.W)) {
                cl.push_back(k);
                placed = true;
                break;
            }
        }
        if (!placed) classes.push_back({k});
    }
    for (const auto& cl : classes) {
        uint32_t in_class = 0;
        for (uint32_t k : cl) in_class += groups[k].batch;
        const uint32_t nchunks = std::min(m->sched_chunks, in_class);
        for (uint32_t c = 0; c < nchunks; ++c) {
            std::vector<bcrt::GroupDesc> chunk;
            uint32_t begin[bcrt::kGroupStages] = {};
            for (uint32_t k : cl) {
                const bcrt_group& g = groups[k];
                const uint32_t n = g.batch / nchunks + (c < g.batch % nchunks ? 1 : 0);
                if (n == 0) continue;
                const uint32_t first = c * (g.batch / nchunks) + std::min(c, g.batch % nchunks);
                const auto& t = g.engine->tables;
                const size_t io = size_t(t.N) * t.L, pl = size_t(2) * t.P * t.N;
                bcrt::GroupDesc d{};
                d.t = t;
                d.args = bcrt::BatchArgs{g.d_a + first * io, g.d_b + first * io, g.d_c + first * io,
                                         g.engine->planes + first * pl, n};
                for (int st = 0; st < bcrt::kGroupStages; ++st) {
                    d.begin[st] = begin[st];
                    begin[st] += bcrt::group_workgroups(d.t, n, st);
                }
                chunk.push_back(d);
            }
            if (!chunk.empty()) m->chunks.push_back(std::move(chunk));
        }
    }
    const uint32_t nchunks = uint32_t(m->chunks.size());
    const uint32_t nstreams = std::min(m->sched_streams, nchunks);
    if (m->sched_flags & BCRT_MIX_PERSISTENT)
        check(hipMalloc(reinterpret_cast<void**>(&m->queues), sizeof(uint32_t) * bcrt::kQueueWords * nchunks),
              "allocate mix work queues");
    if (nstreams > 1) {
        for (uint32_t c = 0; c < nstreams; ++c) {
            hipStream_t st = nullptr;
            check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "create mix worker stream");
            m->workers.push_back(st);
        }
        ensure_events(m->eager_events, nstreams);
        ensure_events(m->graph_events, nstreams);
    }
}

// Enqueue every chunk on `origin`, or forked over the worker streams with the
// events of `ev` when the schedule has several (also valid inside a stream
// capture).
void enqueue_mix(bcrt_mix* m, hipStream_t origin, EventSet& ev) {
    auto queue = [m](size_t c) { return m->queues ? m->queues + c * bcrt::kQueueWords : nullptr; };
    if (m->workers.empty()) {
        for (size_t c = 0; c < m->chunks.size(); ++c)
            check(bcrt::launch_product_groups(m->chunks[c].data(), uint32_t(m->chunks[c].size()), origin, queue(c)),
                  "launch mixed batch");
        return;
    }
    const size_t ns = m->workers.size();
    ensure_events(ev, ns);
    check(hipEventRecord(ev.fork, origin), "fork");
    for (size_t w = 0; w < ns; ++w) check(hipStreamWaitEvent(m->workers[w], ev.fork, 0), "fork wait");
    for (size_t c = 0; c < m->chunks.size(); ++c)
        check(bcrt::launch_product_groups(m->chunks[c].data(), uint32_t(m->chunks[c].size()), m->workers[c % ns],
                                          queue(c)),
              "launch mixed chunk");
    for (size_t w = 0; w < ns; ++w) {
        check(hipEventRecord(ev.joins[w], m->workers[w]), "join record");
        check(hipStreamWaitEvent(origin, ev.joins[w], 0), "join wait");
    }
}

}  // namespace

extern "C" int bcrt_mix_create(bcrt_mix** out, const bcrt_group* groups, uint32_t ngroups, char* err,
                               size_t errlen) {
    if (!out) return BCRT_EINVAL;
    *out = nullptr;
    auto* m = new bcrt_mix();
    int rc = BCRT_OK;
    try {
        if (!groups || ngroups == 0 || ngroups > bcrt::kMaxGroups)
            throw std::invalid_argument("ByteCRT mix needs 1.." + std::to_string(bcrt::kMaxGroups) + " groups");
        uint32_t total = 0;
        for (uint32_t k = 0; k < ngroups; ++k) {
            const bcrt_group& g = groups[k];
            if (!g.engine || !g.d_a || !g.d_b || !g.d_c)
                throw std::invalid_argument("ByteCRT mix group " + std::to_string(k) + " has a null pointer");
            for (uint32_t i = 0; i < k; ++i)
                if (groups[i].engine == g.engine)
                    throw std::invalid_argument("ByteCRT mix groups must use distinct engines");
            require_batch(g.engine, g.batch);
            total += g.batch;
        }
        (void)total;
        m->groups.assign(groups, groups + ngroups);
        check(hipStreamCreate(&m->stream), "create mix stream");
        check(hipEventCreateWithFlags(&m->graph_done, hipEventDisableTiming), "create graph event");
        layout_mix(m);
    } catch (const std::invalid_argument& x) {
        rc = BCRT_EINVAL;
        m->last_error = x.what();
    } catch (const std::exception& x) {
        rc = BCRT_ERUNTIME;
        m->last_error = x.what();
    }
    if (rc != BCRT_OK) {
        if (err && errlen) std::snprintf(err, errlen, "%s", m->last_error.c_str());
        release_mix(m);
        delete m;
        return rc;
    }
    {
        // The mix's launches read its engines' tables and planes: keep them
        // alive until the mix is destroyed.
        std::lock_guard<std::mutex> g(g_refs);
        for (uint32_t k = 0; k < ngroups; ++k) {
            m->engines.push_back(groups[k].engine);
            ++groups[k].engine->mix_refs;
        }
    }
    *out = m;
    return BCRT_OK;
}

extern "C" int bcrt_mix_run(bcrt_mix* m, void* stream, int graph) {
    if (!m) return BCRT_EINVAL;
    try {
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m->stream;
        if (!graph) {
            enqueue_mix(m, s, m->eager_events);
        } else {
            if (!m->graph_exec) {
                // Nothing the mix enqueued earlier is pending when the capture starts.
                (void)hipStreamSynchronize(m->stream);
                for (auto st : m->workers) (void)hipStreamSynchronize(st);
                check(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal), "begin capture");
                std::string fail;
                try {
                    enqueue_mix(m, m->stream, m->graph_events);
                } catch (const std::exception& x) {
                    fail = x.what();
                }
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(m->stream, &g);
                if (!fail.empty()) {
                    if (g) (void)hipGraphDestroy(g);
                    throw std::runtime_error(fail + " (capture)");
                }
                check(ce, "end capture");
                m->graph = g;
                check(hipGraphInstantiate(&m->graph_exec, g, nullptr, nullptr, 0), "instantiate graph");
            }
            check(hipGraphLaunch(m->graph_exec, s), "graph launch");
            check(hipEventRecord(m->graph_done, s), "record graph launch");
            m->graph_launched = true;
        }
        m->last_error.clear();
        return BCRT_OK;
    } catch (const std::exception& x) {
        m->last_error = x.what();
        return BCRT_ERUNTIME;
    }
}

extern "C" const char* bcrt_mix_last_error(const bcrt_mix* m) { return m ? m->last_error.c_str() : "null mix"; }

extern "C" int bcrt_mix_set_schedule(bcrt_mix* m, uint32_t chunks, uint32_t streams, uint32_t flags) {
    if (!m) return BCRT_EINVAL;
    if (chunks < 1 || chunks > kMixMaxChunks || streams < 1 || streams > kMixMaxChunks ||
        (flags & ~uint32_t(BCRT_MIX_MERGE_CLASSES | BCRT_MIX_PERSISTENT))) {
        m->last_error = "ByteCRT mix schedule needs 1 <= chunks, streams <= " + std::to_string(kMixMaxChunks) +
                        " and flags within BCRT_MIX_MERGE_CLASSES | BCRT_MIX_PERSISTENT";
        return BCRT_EINVAL;
    }
    try {
        drop_layout(m);
        m->sched_chunks = chunks;
        m->sched_streams = streams;
        m->sched_flags = flags;
        layout_mix(m);
        m->last_error.clear();
        return BCRT_OK;
    } catch (const std::exception& x) {
        m->last_error = x.what();
        return BCRT_ERUNTIME;
    }
}

extern "C" int bcrt_mix_chunks(const bcrt_mix* m) { return m ? int(m->chunks.size()) : BCRT_EINVAL; }

extern "C" void bcrt_mix_destroy(bcrt_mix* m) {
    if (!m) return;
    release_mix(m);
    std::vector<bcrt_engine*> orphans;  // engines destroyed while this mix held them
    {
        std::lock_guard<std::mutex> g(g_refs);
        for (bcrt_engine* e : m->engines)
            if (--e->mix_refs == 0 && e->destroy_pending) orphans.push_back(e);
    }
    for (bcrt_engine* e : orphans) {
        release(e);
        delete e;
    }
    delete m;
}

namespace bcrt {
const DevTables& engine_tables(const bcrt_engine* e) { return e->tables; }
}  // namespace bcrt

namespace bcrt {

int engine_device(const bcrt_engine* e) { return e->device; }

int engine_fail(bcrt_engine* e, int code, const std::string& msg) { return fail(e, code, msg); }

// The product for a batch with caller-given residue planes (2 * P * N * batch
// words), on one stream: no pipelining, no workspace, nothing allocated.
int run_with_planes(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t* planes,
                    uint32_t batch, void* stream) {
    return guarded(e, [&] {
        const bcrt::BatchArgs args{d_a, d_b, d_c, planes, batch};
        check(bcrt::launch_product(e->tables, args, pick(e, stream)), "launch product");
    });
}

}  // namespace bcrt
