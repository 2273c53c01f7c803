/*
 * bytecrt.h -- C ABI of the MI355X (gfx950) ByteCRT exact negacyclic product
 * engine: c = a * b mod (X^N + 1, q) for N wide coefficients of W bits stored
 * as L = ceil(W/32) little-endian u32 limbs.  Plain pointers and sizes only.
 *
 * The reference exposes this path through three C++ entry points over
 * std::vector tensors (native-batched/fherma.h:47-49):
 *     void* fherma_init(const fherma::Point&);
 *     fherma::Outputs fherma_run(void*, const fherma::Inputs&);
 *     void fherma_free(void*);
 * Those are provided unchanged by include/fherma.h (implemented on top of this
 * ABI in fherma-bytecrt_amd/csrc/fherma_solve.cpp).  Every entry point below
 * names the reference interface it replaces.
 *
 * Status codes: BCRT_OK, BCRT_EINVAL (what the reference raises as
 * std::invalid_argument, and its size/batch std::runtime_error checks),
 * BCRT_ERUNTIME (HIP / plan failures; std::runtime_error in the reference).
 * The message of the last failure is bcrt_last_error().
 */
#ifndef BYTECRT_H
#define BYTECRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BCRT_OK 0
#define BCRT_EINVAL 1
#define BCRT_ERUNTIME 2

typedef struct bcrt_engine bcrt_engine;

/* Replaces fherma_init (native-batched/fherma.h:47; native-batched/solve.cu:160-198,
 * core/solve.cu:295-397).  Builds the RNS plan for (N, W, L, q), uploads the
 * twiddle / MFMA weight tables and reserves device workspace for up to
 * max_batch products.  q: L limbs, bit_length(q) == W (any such q; the
 * reference requires q = 2^W - gap).  N = 2^k, 4096 <= N <= 65536.
 * On failure *out is NULL and err (if given) holds the message. */
int bcrt_create(bcrt_engine** out, uint32_t max_batch, uint32_t N, uint32_t W, uint32_t L,
                const uint32_t* q, char* err, size_t errlen);

/* Replaces fherma_free (native-batched/fherma.h:49; native-batched/solve.cu:246). */
void bcrt_destroy(bcrt_engine* e);

/* Replaces fherma_run (native-batched/fherma.h:48; native-batched/solve.cu:200-244,
 * core/solve.cu:506-525): whole call on host buffers a, b, c of
 * batch * N * L u32 each ([batch][N][L] row-major).  Includes the host-to-
 * device copies, all GPU work and the device-to-host copy; returns when c is
 * complete.  batch <= max_batch. */
int bcrt_run(bcrt_engine* e, const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t batch);

/* Replaces the owned-output drain of the batched adapter
 * (native-batched/solve.cu:138-155, which appends each finished chunk to the
 * output std::vector): the same whole call, but the [batch][N][L] result is
 * handed over in order as consecutive pieces, sink(ctx, words, nwords),
 * called from one engine thread while the call runs.  The call returns after
 * the last sink call. */
typedef void (*bcrt_sink)(void* ctx, const uint32_t* words, size_t nwords);
int bcrt_run_sink(bcrt_engine* e, const uint32_t* a, const uint32_t* b, uint32_t batch, bcrt_sink sink, void* ctx);

/* Device-resident form of fherma_run's GPU work (core/solve.cu:401-430 without
 * the PCIe legs): a, b, c are device pointers ([batch][N][L] u32); the five
 * kernels are enqueued on `stream` (a hipStream_t; NULL selects the engine's
 * own stream) and the call returns without waiting. */
int bcrt_run_device(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                    uint32_t batch, void* stream);

/* Same work as bcrt_run_device, replayed from a hipGraph captured on the first
 * call for this (pointers, batch) key (the batched adapter's stream/slot
 * orchestration, native-batched/solve.cu:56-156, becomes one graph launch). */
int bcrt_run_device_graph(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                          uint32_t batch, void* stream);

/* Allocation-free variant of bcrt_run_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * residue planes (2 * P * N * batch words) live in a 64 MiB device array that
 * the library holds on each GPU from its first call there until the process
 * ends, instead of the engine's workspace; the call allocates nothing, runs
 * on `stream` without pipelining and ignores max_batch.  A batch whose planes
 * exceed 64 MiB is refused (BCRT_EINVAL): one product fits at every geometry,
 * four at N = 32,768, W = 868.  Calls that overlap on one GPU share the array
 * and race. */
int bcrt_run_device_noalloc(bcrt_engine* e, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                            uint32_t batch, void* stream);

/* Batch pipelining (no reference counterpart; the reference overlapped four
 * product slots on three streams, native-batched/solve.cu:17-18,56-156):
 * batches larger than `chunk` products are split into chunks that run on
 * `streams` worker streams (1..16, default priority), forked from and joined
 * back into the caller's stream.  Defaults: chunk 32, streams 2.  Results do
 * not depend on the setting. */
int bcrt_set_pipeline(bcrt_engine* e, uint32_t chunk, uint32_t streams);

/* Encode tile shape (the autotuner's kernel-level knob; no reference
 * counterpart: encode_mma32 fixes its tile at compile time,
 * rns-encode-mma.cuh:148-252).  Each encode workgroup takes 128 coefficients
 * as 2^ts_log consecutive s (index div 256) by 128 >> ts_log consecutive r
 * (index mod 256); ts_log in [0, min(7, log2(N/256))], default min(5, ...).
 * Results do not depend on it.  Applies to later runs and later mixes. */
int bcrt_set_encode_tile(bcrt_engine* e, uint32_t ts_log);

/* Kernel structure of the first two stages (the autotuner's other
 * kernel-level knob; no reference counterpart, the reference always runs
 * encode_mma32 and forward_first_coalesced as two kernels,
 * ntt-pipeline.cuh:73-94).  For N >= 16,384 encode and the forward column
 * levels run as one fused kernel when it has at least `min_tiles` workgroups
 * ((N/256) * 2 * batch; default 256), otherwise as two kernels.  0 fuses
 * whenever the geometry allows, 0xFFFFFFFF never.  Results do not depend on
 * it.  Applies to later runs and later mixes. */
int bcrt_set_fusion(bcrt_engine* e, uint32_t min_tiles);

/* Kernel-level choices (the autotuner's third knob; no reference counterpart:
 * the reference fixes its kernels' tiling and scheduling at compile time,
 * ntt-pipeline.cuh:14-15, rns-encode-mma.cuh:148-252).  Bits: 1 launches the
 * workgroups in plain order instead of renumbering the tiles so that each
 * XCD (and its L2) runs one contiguous eighth; 2 drops the raised issue
 * priority the fused encode gives its MFMA blocks; 4 makes the fused encode's
 * waves load their weight fragments from L2 instead of staging each pair of
 * weight tiles in LDS; 8 runs the CRT built for five workgroups per CU with
 * its weight fragments loaded one k-step ahead (instead of four per CU, four
 * k-steps ahead); 16 runs the separate encode (N < 16,384 or batches too
 * small for the fused kernel) with one MFMA chain per wave at four
 * workgroups per CU (instead of two chains at three); 32 makes the CRT reduce
 * a near-power modulus q = 2^W - gap (0 < gap < 2^32) by the generic-q steps
 * instead of its near-power steps (g1-crt.cuh:15-68; the tuner leaves this bit
 * to tests, which compare the two); 64 makes the NTT kernels multiply by
 * twiddles with the six-operation TwoProduct form instead of the
 * five-operation form exact for primes p = 1 mod 2N (tests compare the two).
 * Default 0.  Results do not depend on them.  Applies to later runs and later
 * mixes (bits 8, 16 and 64 act on single-plan launches only). */
int bcrt_set_kernel_flags(bcrt_engine* e, uint32_t flags);

/* ---- Mixed batches (products of different N, W, q, one launch per stage) ----
 * No reference counterpart: the reference runs one fixed geometry per process
 * (native-batched/solve.cu:163-164).  Each group names an engine (its plan
 * and workspace; engines must be distinct; at most 16 groups) and a device
 * batch for it.  Under the default schedule (bcrt_mix_set_schedule) the
 * five stages of each geometry class run as ONE launch each, whose
 * workgroups cover every product of every modulus of that class. */
typedef struct bcrt_group {
    bcrt_engine* engine;
    const uint32_t* d_a;  /* [batch][N][L] device */
    const uint32_t* d_b;
    uint32_t* d_c;
    uint32_t batch;       /* <= the engine's max_batch */
} bcrt_group;

typedef struct bcrt_mix bcrt_mix;

/* The mix holds its engines' tables and workspace: bcrt_destroy of an engine
 * that a live mix names is deferred until the last such mix is destroyed. */
int bcrt_mix_create(bcrt_mix** out, const bcrt_group* groups, uint32_t ngroups, char* err, size_t errlen);
/* Enqueue the mixed batch on `stream` (hipStream_t, NULL = default); graph
 * != 0 replays a hipGraph captured on first use. */
int bcrt_mix_run(bcrt_mix* m, void* stream, int graph);
const char* bcrt_mix_last_error(const bcrt_mix* m);
void bcrt_mix_destroy(bcrt_mix* m);

/* Schedule of a mix (no reference counterpart; the reference's batched driver
 * fixes one: 4 slots over 3 streams, native-batched/solve.cu:17-18).  Groups
 * of one geometry (N, W) form a class; each class is dealt into `chunks`
 * chunks (each group split evenly), a chunk being one launch per stage over
 * its products, and chunks go round robin onto `streams` worker streams
 * forked from the caller's stream (1: every chunk on the caller's stream, in
 * order).  BCRT_MIX_MERGE_CLASSES puts every class into the same chunks (one
 * launch per stage and chunk over all geometries; the kernels then switch
 * shape per workgroup).  BCRT_MIX_PERSISTENT runs every stage launch as a
 * resident grid (as many workgroups as the device holds at once) whose
 * workgroups pull (group, product, prime) tiles from per-XCD work counters
 * (the mix allocates 192 bytes of counters per chunk; a small kernel clears
 * them at the start of each chunk, inside the captured graph too).
 * 1 <= chunks, streams <= 16; a new mix has chunks = streams = 1, no flags.
 * Waits for the mix's pending work and drops its graph; results do not depend
 * on the schedule. */
#define BCRT_MIX_MERGE_CLASSES 1u
#define BCRT_MIX_PERSISTENT 2u
int bcrt_mix_set_schedule(bcrt_mix* m, uint32_t chunks, uint32_t streams, uint32_t flags);
/* Launch sequences (chunks) of the mix's current schedule; BCRT_EINVAL for a
 * null mix. */
int bcrt_mix_chunks(const bcrt_mix* m);

/* ---- Serving: continuous batching of single-product requests ----
 * No reference counterpart (the reference serves one fixed-size batch per
 * fherma_run call; its per-init state is built once, native-batched/solve.cu:160-198).
 * hbm_budget_bytes bounds the server's device memory: three quarters are the
 * request arena (each request's operands, result and residue planes), one
 * quarter holds resident plan tables (twiddles and MFMA weights per (N, W, q),
 * no workspace).  Requests of any geometry are admitted, in arrival order,
 * into waves while the arena and pinned staging have room; each wave runs as
 * one mixed launch per stage on one of two compute streams, so waves admitted
 * while others are on the GPU overlap with them.  A request whose plan is not
 * resident builds it, first evicting idle plans (no request pending or in
 * flight) least recently used first until its tables fit; submit blocks
 * while every resident plan that could make room is still in use.  A request
 * larger than the arena, or a plan larger than the table share, is refused
 * with BCRT_EINVAL before anything is built.
 * a, b, c are host [N][L] buffers that must stay valid until the request is
 * waited on. */
typedef struct bcrt_server bcrt_server;
int bcrt_server_create(bcrt_server** out, size_t hbm_budget_bytes, uint32_t max_wave, char* err, size_t errlen);
int bcrt_server_submit(bcrt_server* s, uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, const uint32_t* a,
                       const uint32_t* b, uint32_t* c, uint64_t* ticket);
/* Blocks until the request is complete; latency_s = submit-to-completion. */
int bcrt_server_wait(bcrt_server* s, uint64_t ticket, double* latency_s);
/* out[0..7] = completed requests, waves launched, largest wave, resident
 * plans, plan evictions, table bytes resident, table share, arena bytes. */
int bcrt_server_stats(bcrt_server* s, uint64_t* out);
const char* bcrt_server_last_error(const bcrt_server* s);
void bcrt_server_destroy(bcrt_server* s);

/* Waits for all work the engine enqueued on its own stream. */
int bcrt_synchronize(bcrt_engine* e);

/* Message of the last failing call on this engine ("" if none). */
const char* bcrt_last_error(const bcrt_engine* e);

/* Geometry: out[0..9] = N, W, L, P, N1, N2, Kenc, Kcrt, Mcrt, product_bits. */
int bcrt_geometry(const bcrt_engine* e, uint32_t* out);

/* Plan tables (host copies): primes[P], psi[P], crt_inverse[P] and the CRT
 * weights[(P+1) * L] (P_j mod q, last row (-P) mod q) -- the values of
 * RnsPlan in native-smoke/rns-plan.hpp:13-23 and reference.py:50-53. */
int bcrt_plan_tables(const bcrt_engine* e, uint32_t* primes, uint32_t* psi, uint32_t* crt_inverse,
                     uint32_t* weights);

/* Host-only plan (no device work): geometry out[10] as bcrt_geometry, and
 * optionally primes[P], psi[P], crt_inverse[P], weights[(P+1)*L] as
 * bcrt_plan_tables.  Pass NULL for tables not wanted; call once with only
 * geometry to learn P.  Restates make_rns_plan (native-smoke/rns-plan.hpp:174-324)
 * for any q of W bits. */
int bcrt_plan_host(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q, uint32_t* geometry,
                   uint32_t* primes, uint32_t* psi, uint32_t* crt_inverse, uint32_t* weights,
                   char* err, size_t errlen);

/* *gap = 2^W - q when that is below 2^32 (q of L >= 3 limbs and bit length
 * W), else 0.  A plan for such a q reduces the CRT sum by the near-power
 * steps (q = 2^W - gap is the only modulus form make_rns_plan accepts,
 * rns-plan.hpp:186-197, reduced by reduce_near_power, g1-crt.cuh:45-68);
 * any other q takes the generic-q steps. */
int bcrt_near_power_gap(uint32_t W, uint32_t L, const uint32_t* q, uint32_t* gap);

/* Stage-level access for parity tests and profiling.  Residue planes
 * ([batch][2][P][N] u32, column-major positions) live in d_planes (NULL = the
 * engine workspace).  stage_mask bits: 1 encode, 2 forward columns, 4 middle
 * (forward blocks, pointwise, inverse blocks), 8 inverse columns + CRT scaling,
 * 16 CRT.  A mask with both 1 and 2 runs them as one fused kernel when
 * N >= 16,384 and (N/256) * 2 * batch >= 256 (a workgroup per 256
 * coefficients of whole columns), as the product path does; 1 or 2 alone runs
 * the separate kernels, which write the same planes.
 * The stages replace, in order, encode_mma32 (rns-encode-mma.cuh:148-252),
 * forward_first_coalesced (ntt-pipeline.cuh:73-94), fused_middle
 * (ntt-pipeline.cuh:96-120), inverse_weighted (ntt-weighted.cuh:10-44) and
 * g1_crt32 (g1-crt.cuh:201-335). */
int bcrt_run_stages(bcrt_engine* e, uint32_t stage_mask, const uint32_t* d_a, const uint32_t* d_b,
                    uint32_t* d_c, uint32_t* d_planes, uint32_t batch, void* stream);

/* Words of residue-plane workspace a batch needs (2 * P * N * batch). */
size_t bcrt_planes_words(const bcrt_engine* e, uint32_t batch);

/* One 32x32x32 int8 MFMA tile on the device: c[i][j] = sum_k a[i][k] * b[j][k]
 * (row-major int8 32x32 inputs, int32 output); the lane-layout check behind
 * the encode and CRT GEMMs (replaces native-smoke/mma-smoke.cu). */
int bcrt_mfma_probe(const int8_t* d_a, const int8_t* d_b, int32_t* d_c);

#ifdef __cplusplus
}
#endif

#endif /* BYTECRT_H */
