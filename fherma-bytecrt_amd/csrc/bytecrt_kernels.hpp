// Kernel launch interface of the MI355X ByteCRT engine (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bcrt {

// Device-resident constant tables of one plan, passed by value to every kernel.
struct DevTables {
    uint32_t N, W, L, P, N1, n1, N2, Kenc, PTenc, Kcrt, Mcrt, qshift;
    uint32_t enc_ts_log;           // encode tile: 2^enc_ts_log consecutive s by 128 >> enc_ts_log r
    uint32_t fuse_min_tiles;       // host: fused encode + forward columns from this many workgroups
    uint32_t kflags;               // kernel-level choices (kFlag*), 0 = the measured defaults
    uint32_t gap;                  // q = 2^W - gap, 0 < gap < 2^32, L >= 3 (near-power CRT reduction); else 0
    uint64_t qt1;                  // (q >> qshift) + 1
    double inv_qt1;                // 1 / qt1
    const double* pd;              // [P][2]: p, 1/p
    const uint32_t* zf;            // [P][N]: psi^brv(k) mod p
    const uint32_t* zi;            // [P][N]: psi^-brv(k) mod p
    const double* scale;           // [P]: (N/2)^-1 (P/p)^-1 mod p
    const double* scale_z1;        // [P]: scale psi^-brv(1) mod p
    const double* zf16;            // [P][16]: psi^brv(k) mod p, k < 16
    const void* enc_prime;         // [P] x 24 bytes: off (f64), 1/p rounded down (f64), p, off mod 2^32
    const int8_t* wenc;            // [PTenc][32][Kenc]
    const int8_t* wcrt;            // [Mcrt][Kcrt]
    const int64_t* corr;           // [Mcrt/4]
    const double* invp;            // [P]
    const uint32_t* q;             // [L]
};

// Kernel-level choices the autotuner can change at run time (DevTables::kflags).
// Results never depend on them.
constexpr uint32_t kFlagNoXcd = 1;      // plain workgroup order instead of the XCD tile renumbering
constexpr uint32_t kFlagNoPrio = 2;     // fused encode: no raised issue priority over its MFMA block
constexpr uint32_t kFlagWeightsL2 = 4;  // fused encode: weight fragments from L2, not staged in LDS
constexpr uint32_t kFlagCrtWide = 8;    // CRT built for five workgroups per CU, weights one k-step ahead
constexpr uint32_t kFlagEncOneChain = 16;  // separate encode: one MFMA chain per wave, four waves per SIMD
constexpr uint32_t kFlagGenericQ = 32;     // CRT: the generic-q reduction also for q = 2^W - gap
constexpr uint32_t kFlagMulModTw6 = 64;    // NTT kernels: six-operation twiddle products (mul_mod) throughout
constexpr uint32_t kFlagsAll = 127;
// Set by the engine from its plan, not a choice: every prime allows the
// five-operation twiddle products in the middle kernel's last four inverse
// levels (|q| k < 2^53 for |q| < 192 p + 3 p / 32 + 1, where p - 1 = k 2^s).
constexpr uint32_t kPlanTw5Late = 1u << 31;
constexpr uint32_t kPlanFlags = kPlanTw5Late;

// One batch of products.  a, b, out: [batch][N][L] u32 coefficient-major.
// planes: [batch][2][P][N] u32 workspace (operand A planes end up holding the
// CRT terms).  product_stride_in/out are the element strides between products
// (N*L for dense batches).
struct BatchArgs {
    const uint32_t* a;
    const uint32_t* b;
    uint32_t* out;
    uint32_t* planes;
    uint32_t batch;
};

// Stages of a mixed launch: 0 encode, 1 forward columns, 2 middle, 3 inverse
// columns, 4 CRT, and 5 encode fused with the forward columns (used instead
// of 0 and 1 when every group has N1 = 256).
constexpr int kGroupStages = 6;

// One group of a mixed launch: a plan and a batch of products under it.
// begin[s] is the group's first workgroup in stage s (prefix sums over the
// groups, in order; group_workgroups gives each group's count).  The table
// travels by value in the kernel arguments (16 groups fit the 4 KiB limit),
// so its pointers are known to be global and nothing is uploaded per launch.
constexpr uint32_t kMaxGroups = 16;
struct GroupDesc {
    DevTables t;
    BatchArgs args;
    uint32_t begin[kGroupStages];
};
struct GroupTable {
    GroupDesc g[kMaxGroups];
    uint32_t n;
    // Work queue (null: one workgroup per tile).  Otherwise each stage runs
    // as a resident grid whose workgroups pull tiles from queue[8 s + x],
    // x = the workgroup's XCD, until total[s] is reached; the counters start
    // at zero (launch_product_groups clears them with a kernel first).
    uint32_t* queue;
    uint32_t total[kGroupStages];
};
static_assert(sizeof(GroupTable) <= 4096, "kernel argument limit");

// Workgroups of stage `stage` for `batch` products of plan t in a mixed
// launch (padded to a multiple of 8; the padding workgroups exit at once).
uint32_t group_workgroups(const DevTables& t, uint32_t batch, int stage);

// Enqueue the five stages for groups[0..ngroups): one launch per stage
// covering every product of every group.  With `queue` (kQueueWords device
// words, owned by the caller, not shared by launches that may overlap) every
// stage runs as a persistent grid pulling (group, product, prime) tiles.
constexpr uint32_t kQueueWords = 8 * kGroupStages;
hipError_t launch_product_groups(const GroupDesc* groups, uint32_t ngroups, hipStream_t stream,
                                 uint32_t* queue = nullptr);

// Enqueue the whole product pipeline for a batch on `stream` (no host sync).
hipError_t launch_product(const DevTables& t, const BatchArgs& args, hipStream_t stream);

// Individual stages (tests and profiling).
hipError_t launch_encode(const DevTables& t, const BatchArgs& args, hipStream_t stream);
hipError_t launch_forward_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream);
// Encode and forward column levels in one kernel (N1 >= 64: one workgroup per
// 256 coefficients of whole columns of one operand) when that is at least
// t.fuse_min_tiles workgroups; otherwise the two kernels.
constexpr uint32_t kFusedMinTiles = 256;
hipError_t launch_encode_forward(const DevTables& t, const BatchArgs& args, hipStream_t stream);
hipError_t launch_middle(const DevTables& t, const BatchArgs& args, hipStream_t stream);
hipError_t launch_inverse_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream);
hipError_t launch_crt(const DevTables& t, const BatchArgs& args, hipStream_t stream);

// 32x32x32 signed int8 MFMA tile: c[32][32] = a[32][32] * b[32][32]^T (b given
// as [col][k]); one wave.  Lane-layout check for tests.
hipError_t launch_mfma_probe(const int8_t* a, const int8_t* b, int32_t* c, hipStream_t stream);

}  // namespace bcrt
