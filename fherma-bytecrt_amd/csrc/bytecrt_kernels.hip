// gfx950 kernels of the ByteCRT exact negacyclic product.
//
// Pipeline per product (all integer, bit-exact):
//   k_encode      wide coefficients -> residues mod each prime (int8 MFMA byte-power GEMM)
//   k_fwd_cols    forward NTT levels 0..n1-1 (columns of N1 elements, stride N2)
//   k_middle      forward levels n1..n-1 of A and B, pointwise product, inverse levels n-1..n1
//   k_inv_cols    inverse levels n1-1..0, scale by (N/2)^-1 * (P/p)^-1 (CRT terms)
//   k_crt         centered alpha + int8 MFMA CRT GEMM + carry + generic mod-q reduction
// For N >= 16,384 (N1 >= 64) the product path runs k_encode_fwd, one kernel
// for the first two (a workgroup per 256 coefficients of whole columns); the
// separate ones serve smaller N and batches, and the stage-level entry points.
//
// Residue planes use the "column-major" position f = r * N1 + s for natural
// coefficient index r + N2 * s (r < N2 = 256, s < N1 = N / 256): the column
// transforms read/write N1 contiguous words and the block transforms read a
// tile of 16 consecutive s for each r.
//
// The NTT is the merged negacyclic Cooley-Tukey / Gentleman-Sande pair with
// psi^{+-bitrev(k)} twiddles (no separate twist pass, bit-reversed spectrum
// consumed in place).  Inside a kernel residues are FP64 integers in a lazy
// balanced range and every modular product is exact (see mul_mod).  Between
// kernels the encode output and the CRT terms are canonical u32 in [0, p);
// the forward-column and middle outputs, read only by the next NTT kernel,
// are balanced int32 (|r| <= p/2 + 1; half of the middle kernel's words are
// its last modular products, |r| <= p/2 + 2^-12 p), which is cheaper to form.
#include "bytecrt_kernels.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace bcrt {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint32_t kThreads = 256;
constexpr uint32_t kEncCoefs = 128;  // coefficients per encode workgroup
constexpr uint32_t kCrtCoefs = 64;   // coefficients per CRT workgroup
constexpr uint32_t kMidBlocks = 16;  // 256-element blocks per middle workgroup
constexpr uint32_t kMaxPrimesCrt = 128;

// ---------------------------------------------------------------------------
// Exact modular arithmetic in FP64, p in (2^30, 2^31).
//
// mul_mod(x, w): x an integer-valued double with |x| <= 2^40, w in [0, p).
//   h + l == x * w exactly (TwoProduct by FMA); q = rint(h / p) from one FMA
//   against 1.5 * 2^52 (|h / p| < 2^51 keeps the sum in the binade where the
//   spacing is 1); r = (h - q p) + l.  h - q p is an integer of magnitude
//   below 2^31 + 2^18, so the FMA computes it exactly, and the final sum is
//   exact too.  Hence r == x w (mod p) exactly, with |r| <= p/2 + 2^-12 p.
// canon_product(r): a mul_mod result -> the representative in [0, p).
// balanced(x): |x| <= 2^40 -> a representative with |r| <= p/2 + 1 (int32).
// Forward butterflies grow |x| by at most 0.51 p per level, inverse (GS) sums
// at most double it: eight levels from |x| < p stay below 2^39.
// Larger inputs: for |x| < 2^48 the same steps stay exact (|h / p| < 2^49,
// the quotient off by at most 2^-4, |l| <= 2^26), with |r| <= 0.63 p, and
// balanced() still returns |r| <= p/2 + 1.  The fused encode relies on this:
// it feeds its unreduced dot sums (u < 2^47) straight into the forward
// column levels, which then stay below 2^48.
// ---------------------------------------------------------------------------
constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52

__device__ __forceinline__ double mul_mod(double x, double w, double p, double pinv) {
    const double h = x * w;
    const double l = fma(x, w, -h);
    const double q = fma(h, pinv, kRound) - kRound;
    return fma(-q, p, h) + l;
}

// |r| <= p/2 + 2^-12 p < p for a mul_mod result: one conditional add of p.
__device__ __forceinline__ uint32_t canon_product(double r, double p) { return uint32_t(r < 0.0 ? r + p : r); }

__device__ __forceinline__ int32_t balanced(double x, double p, double pinv) {
    const double q = fma(x, pinv, kRound) - kRound;
    return int32_t(fma(-q, p, x));  // exact, |r| <= p/2 + 1
}

// A residue word of a plane or exchange buffer as an FP64 integer: canonical
// u32 and balanced int32 words agree on [0, 2^31).
__device__ __forceinline__ double word_value(uint32_t w) { return double(int32_t(w)); }

// ---------------------------------------------------------------------------
// Twiddle products in five FP64 operations (mul_mod takes six).  Every prime
// of a plan has p = 1 (mod 2N), so p - 1 = k 2^s with s >= log2(2N) and
// k < 2^(31 - s).  For w in [0, p), wq = fl(w * pinv) (one operation per
// twiddle, shared by all its butterflies):
//   q = rint(x wq): the FMA against 1.5 * 2^52 rounds the exact x wq once,
//       and |wq - w/p| <= 2^-52, so |q - x w / p| <= 1/2 + |x| 2^-52;
//   c = q p - q = q (p - 1) = (q k) 2^s by one FMA: exact while |q| k < 2^53;
//   t = x w - c = (x w - q p) + q: one FMA, exact (an integer below 2^40);
//   r = t - q = x w - q p, exact, with |r| <= p/2 + |x| 2^-52 p, which is
//       below p/2 + 2^-13 p for |x| < 2^39.
// No TwoProduct is needed: the product's low part never has to be rounded.
// Each caller states the largest |x| it passes and why |q| k < 2^53 holds
// there (|q| < |x| + 1); mul_mod stays where inputs reach 2^48 (the fused
// encode) or both factors vary (the pointwise product).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mul_mod_tw(double x, double w, double wq, double p) {
    const double q = fma(x, wq, kRound) - kRound;
    return fma(x, w, -fma(q, p, -q)) - q;
}

// A twiddle's modular product: the five-operation form (FAST, wq given) or mul_mod.
template <bool FAST>
__device__ __forceinline__ double tw_mul(double x, double w, double wq, double p, double pinv) {
    if constexpr (FAST)
        return mul_mod_tw(x, w, wq, p);
    else
        return mul_mod(x, w, p, pinv);
}

// (x, y) -> (x + w y, x - w y)
template <bool FAST = false>
__device__ __forceinline__ void ct_bfly(double& x, double& y, double w, double wq, double p, double pinv) {
    const double v = tw_mul<FAST>(y, w, wq, p, pinv);
    y = x - v;
    x = x + v;
}

__device__ __forceinline__ void ct_bfly(double& x, double& y, double w, double p, double pinv) {
    ct_bfly<false>(x, y, w, 0.0, p, pinv);
}

// (x, y) -> (x + y, (x - y) w)
template <bool FAST = false>
__device__ __forceinline__ void gs_bfly(double& x, double& y, double w, double wq, double p, double pinv) {
    const double d = x - y;
    x = x + y;
    y = tw_mul<FAST>(d, w, wq, p, pinv);
}

// The twiddle's quotient factor for the five-operation form (unused otherwise).
template <bool FAST>
__device__ __forceinline__ double tw_quot(double w, double pinv) {
    return FAST ? w * pinv : 0.0;
}

// K transform levels on the thread's 16 registers, which hold (16 >> K)
// sub-blocks of 2^K consecutive registers.  Local level l is level e0 + l of
// a length-M transform; the sub-block's outer index is o0 + sb; the twiddle
// of group g is z[(mult << (e0 + l)) + ((o0 + sb) << l) + g].
// (FAST: five-operation twiddle products.)
template <int K, bool FAST = false>
__device__ __forceinline__ void ct_round(double (&v)[16], const uint32_t* __restrict__ z, uint32_t mult,
                                         uint32_t e0, uint32_t o0, double p, double pinv) {
#pragma unroll
    for (int l = 0; l < K; ++l) {
        const int half = 1 << (K - 1 - l);
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb) {
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) {
                const double w = double(z[(mult << (e0 + l)) + ((o0 + sb) << l) + g]);
                const double wq = tw_quot<FAST>(w, pinv);
#pragma unroll
                for (int x = 0; x < half; ++x) {
                    const int i0 = (sb << K) + g * 2 * half + x;
                    ct_bfly<FAST>(v[i0], v[i0 + half], w, wq, p, pinv);
                }
            }
        }
    }
}

// (LMIN > 0 stops above the round's lowest LMIN levels.)
template <int K, int LMIN = 0, bool FAST = false>
__device__ __forceinline__ void gs_round(double (&v)[16], const uint32_t* __restrict__ z, uint32_t mult,
                                         uint32_t e0, uint32_t o0, double p, double pinv) {
#pragma unroll
    for (int l = K - 1; l >= LMIN; --l) {
        const int half = 1 << (K - 1 - l);
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb) {
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) {
                const double w = double(z[(mult << (e0 + l)) + ((o0 + sb) << l) + g]);
                const double wq = tw_quot<FAST>(w, pinv);
#pragma unroll
                for (int x = 0; x < half; ++x) {
                    const int i0 = (sb << K) + g * 2 * half + x;
                    gs_bfly<FAST>(v[i0], v[i0 + half], w, wq, p, pinv);
                }
            }
        }
    }
}

struct EncPrime {
    double off;       // offset making the biased dot sum non-negative (< 2^47, exact)
    double inv;       // a double <= 1/p
    uint32_t p;
    uint32_t off_lo;  // off mod 2^32
};

__device__ __forceinline__ uint32_t* plane_ptr(const DevTables& t, uint32_t* planes, uint32_t prod,
                                               uint32_t op, uint32_t j) {
    return planes + ((size_t(prod) * 2 + op) * t.P + j) * t.N;
}

// ---------------------------------------------------------------------------
// Encode: residues of 128 coefficients (consecutive column-major positions)
// for every prime.  GEMM C[w-row][coef] = sum_k Wenc[w-row][k] * byte_k(x_coef),
// w-row = 4 * prime + digit, with balanced signed weight digits and biased data
// bytes (x_byte ^ 0x80 = x_byte - 128); per prime
//   x = sum_d 256^d C[4j+d] + 128 sum_k r_k  (mod p),  |sum_d 256^d C| < 2^45.
// ---------------------------------------------------------------------------
size_t enc_smem_bytes(const DevTables& t) {
    return size_t(kEncCoefs) * (t.Kenc + 16) + sizeof(EncPrime) * kMaxPrimesCrt;
}

// One encode workgroup: tile bx of operand op of product prod.
// Encode tile bx: TR consecutive r (index mod 256) by TS consecutive s (index
// div 256), TS * TR = 128 with TS = 2^enc_ts_log <= N1 chosen per plan (32
// unless tuned): each s brings TR consecutive coefficients
// (TR * L contiguous input words), each (prime, r) gets TS consecutive plane
// positions.  Tile row m is (r0 + m / TS, s0 + m % TS).
struct EncGeom {
    uint32_t ts_log, TS, TR, r0, s0, row_words, stride_words;
    __device__ __forceinline__ EncGeom(const DevTables& t, uint32_t ts, uint32_t r, uint32_t s) {
        ts_log = ts;
        TS = 1u << ts_log;
        TR = kEncCoefs >> ts_log;
        r0 = r;
        s0 = s;
        row_words = t.Kenc / 4;
        stride_words = row_words + 4;
    }
    __device__ __forceinline__ EncGeom(const DevTables& t, uint32_t bx) {
        ts_log = t.enc_ts_log;
        TS = 1u << ts_log;
        TR = kEncCoefs >> ts_log;
        const uint32_t rtiles = t.N2 / TR;
        r0 = (bx % rtiles) * TR;
        s0 = (bx / rtiles) * TS;
        row_words = t.Kenc / 4;
        stride_words = row_words + 4;
    }
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Input staging of an encode tile.  In LDS each of the tile's 128
// coefficients has a row of Kenc bytes (8 KS words: its L limbs, then
// padding), moved as 2 KS pieces of 16 bytes, KS per thread: one buffer load
// and one 16-byte LDS store per piece, every load in flight before the first
// store.  Piece P = threadIdx.x + kThreads * it is piece q = P mod 2 KS of tile
// coefficient c = P / 2 KS, which is row ro = c & (TR - 1) of group gi =
// c >> (7 - ts_log) (one s; a group's TR coefficients are contiguous in the
// operand, groups N2 L words apart): it starts at source word
// (ro + N2 gi) L + 4 q from the tile's first coefficient and lands at word 4 q
// of LDS row (ro << ts_log) + gi.  The padding words of a row receive
// whatever follows the coefficient in the operand (the next coefficient's
// first limbs, or 0 past the operand's end, where buffer loads return 0):
// the plan fills only columns k < 4 L of the weight tiles, so padding bytes
// add nothing to any dot sum.  The loads are raw buffer loads with 32-bit
// offsets (N L < 2^22) over the operand from the tile's first coefficient on.
template <int KS>
struct EncPiece {
    uint32_t src, row;  // source word offset, LDS row
    __device__ __forceinline__ EncPiece(const DevTables& t, const EncGeom& e, uint32_t it) {
        constexpr uint32_t PPC = 2 * KS;
        // P / PPC for P < 2^12 as a 24-bit multiply by ceil(2^20 / PPC) (exact:
        // the rounding adds less than 2^-8 < 1 / PPC)
        constexpr uint32_t kInv = ((1u << 20) + PPC - 1) / PPC;
        const uint32_t P = threadIdx.x + kThreads * it;
        const uint32_t c = __umul24(P, kInv) >> 20, q = P - c * PPC;
        const uint32_t tr_log = 7 - e.ts_log, gi = c >> tr_log, ro = c & (e.TR - 1);
        src = __umul24(ro + __umul24(gi, t.N2), t.L) + 4 * q;
        row = __umul24((ro << e.ts_log) + gi, e.stride_words) + 4 * q;
    }
};

template <int KS>
__device__ __forceinline__ void enc_load(const DevTables& t, const uint32_t* __restrict__ in, const EncGeom& e,
                                         u32x4 (&regs)[KS]) {
    const uint32_t first = (e.r0 + t.N2 * e.s0) * t.L;  // word offset of the tile in the operand
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(in + first), 0, int((t.N * t.L - first) * 4), 0x00020000);
#pragma unroll
    for (uint32_t it = 0; it < KS; ++it)
        regs[it] = __builtin_amdgcn_raw_buffer_load_b128(rs, 4 * EncPiece<KS>(t, e, it).src, 0, 0);
}

// The loaded pieces into the LDS tile as biased bytes (x ^ 0x80 = x - 128, for
// the signed MFMA).
template <int KS>
__device__ __forceinline__ void enc_store(const DevTables& t, const EncGeom& e, const u32x4 (&regs)[KS],
                                          uint32_t* enc_tile) {
#pragma unroll
    for (uint32_t it = 0; it < KS; ++it)
        *reinterpret_cast<u32x4*>(enc_tile + EncPiece<KS>(t, e, it).row) = regs[it] ^ 0x80808080u;
}

// The per-prime reduction constants {off, 1/p rounded down, p, off mod 2^32}
// into LDS (after the data tile).
__device__ __forceinline__ EncPrime* enc_constants(const DevTables& t, uint8_t* smem) {
    EncPrime* cst = reinterpret_cast<EncPrime*>(smem + size_t(kEncCoefs) * (t.Kenc + 16));
    for (uint32_t j = threadIdx.x; j < t.P; j += kThreads)
        cst[j] = reinterpret_cast<const EncPrime*>(t.enc_prime)[j];
    return cst;
}

// GEMM and reduction of the tile in LDS, residues to the planes.  C = 2: wave w
// takes the coefficient sub-tiles 2 (w & 1) and 2 (w & 1) + 1 (two independent
// accumulator chains) and the weight tiles pt = w >> 1, +2, ...; C = 1: wave w
// takes sub-tile w and every weight tile (half the accumulators, so the
// kernel fits four workgroups per CU, but each weight tile is loaded by four
// waves instead of two).
template <int KS, int C>
__device__ __forceinline__ void enc_compute(const DevTables& t, uint32_t* __restrict__ planes, const EncGeom& e,
                                            uint32_t op, uint32_t prod, const uint8_t* smem, const EncPrime* cst) {
    const uint32_t* enc_tile = reinterpret_cast<const uint32_t*>(smem);
    const uint32_t ts_log = e.ts_log, TS = e.TS, r0 = e.r0, s0 = e.s0, stride_words = e.stride_words;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 31, kh = lane >> 5;
    const uint8_t* tile_bytes = reinterpret_cast<const uint8_t*>(enc_tile);
    const uint32_t stride_bytes = stride_words * 4;
    uint32_t* out = plane_ptr(t, planes, prod, op, 0);  // prime j at out + j * N (P * N < 2^32)
    static_assert(C == 1 || C == 2, "encode chains");
    const uint32_t c0 = C == 2 ? 2 * (wave & 1) : wave, pt0 = C == 2 ? wave >> 1 : 0, pstep = C;
    // weight fragments of the next tile are loaded while this one computes
    v4i A[KS];
    {
        const uint32_t pt = min(pt0, t.PTenc - 1);
        const int8_t* wrow = t.wenc + (size_t(pt) * 32 + row) * t.Kenc + 16 * kh;
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2) A[s2] = *reinterpret_cast<const v4i*>(wrow + 32 * s2);
    }
    for (uint32_t pt = pt0; pt < t.PTenc; pt += pstep) {
        v4i An[KS];
        {
            const uint32_t pn = min(pt + pstep, t.PTenc - 1);  // clamped: the last tile reloads itself
            const int8_t* wrow = t.wenc + (size_t(pn) * 32 + row) * t.Kenc + 16 * kh;
#pragma unroll
            for (int s2 = 0; s2 < KS; ++s2) An[s2] = *reinterpret_cast<const v4i*>(wrow + 32 * s2);
        }
        // constants of this lane's four primes j = 8 pt + 2 g + kh (clamped;
        // the stores of primes past P are masked)
        double off[4], inv[4], pd[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const EncPrime c = cst[min(pt * 8 + 2 * g + kh, t.P - 1)];
            off[g] = c.off;
            inv[g] = c.inv;
            pd[g] = double(c.p);
        }
        v16i acc[C];
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] = v16i{};
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2) {
            v4i B[C];
#pragma unroll
            for (int c = 0; c < C; ++c)
                B[c] = *reinterpret_cast<const v4i*>(tile_bytes + ((c0 + c) * 32 + row) * stride_bytes + 16 * kh +
                                                     32 * s2);
#pragma unroll
            for (int c = 0; c < C; ++c)
                acc[c] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[s2], B[c], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const uint32_t m = (c0 + c) * 32 + row;
            const uint32_t f = (r0 + (m >> ts_log)) * t.N1 + s0 + (m & (TS - 1));
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                // digit sums |C| < 2^22: the two byte pairs fit int32
                const int32_t lo = acc[c][4 * g] + acc[c][4 * g + 1] * 256;
                const int32_t hi = acc[c][4 * g + 2] + acc[c][4 * g + 3] * 256;
                // u = lo + 2^16 hi + off, 0 <= u < 2^47, exact in FP64; inv <= 1/p makes
                // qd = floor(u / p) or one less, so u - qd p (exact: both below 2^47)
                // is in [0, 2p).
                const double ud = fma(double(hi), 65536.0, double(lo)) + off[g];
                const double qd = __builtin_trunc(ud * inv[g]);
                const uint32_t r = uint32_t(fma(-qd, pd[g], ud));
                const uint32_t j = pt * 8 + 2 * g + kh;
                if (j < t.P) out[j * t.N + f] = min(r, r - uint32_t(pd[g]));
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < KS; ++s2) A[s2] = An[s2];
    }
}

// One encode tile (bx, op, prod) with C MFMA chains per wave.
template <int KS, int C = 2>
__device__ __forceinline__ void encode_wg(const DevTables& t, const uint32_t* __restrict__ a,
                                          const uint32_t* __restrict__ b, uint32_t* __restrict__ planes,
                                          uint32_t bx, uint32_t op, uint32_t prod, uint8_t* smem) {
    const EncPrime* cst = enc_constants(t, smem);
    const EncGeom e(t, bx);
    u32x4 regs[KS];
    enc_load<KS>(t, (op ? b : a) + size_t(prod) * t.N * t.L, e, regs);
    enc_store<KS>(t, e, regs, reinterpret_cast<uint32_t*>(smem));
    __syncthreads();
    enc_compute<KS, C>(t, planes, e, op, prod, smem, cst);
}

// Encode with the k-step count (Kenc / 32, 1..8) as a compile-time constant.
template <int C = 2>
__device__ __forceinline__ void encode_dispatch(const DevTables& t, const uint32_t* __restrict__ a,
                                                const uint32_t* __restrict__ b, uint32_t* __restrict__ planes,
                                                uint32_t bx, uint32_t op, uint32_t prod, uint8_t* smem) {
    switch (t.Kenc / 32) {
        case 1: encode_wg<1, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 2: encode_wg<2, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 3: encode_wg<3, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 4: encode_wg<4, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 5: encode_wg<5, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 6: encode_wg<6, C>(t, a, b, planes, bx, op, prod, smem); break;
        case 7: encode_wg<7, C>(t, a, b, planes, bx, op, prod, smem); break;
        default: encode_wg<8, C>(t, a, b, planes, bx, op, prod, smem); break;
    }
}

// ---------------------------------------------------------------------------
// LDS exchange layout shared by the NTT kernels: element u of a length-M
// sequence c sits at c * S + u + (u >> 4).  With S >= M + M/16 and
// S == T (mod 32), where T = M/16 threads own one sequence, both exchange
// patterns (thread tau holding u = tau + T x, or u = 16 tau + x) are
// bank-conflict free for 32 banks per half-wave and 64 banks per wave.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t skew(uint32_t u) { return u + (u >> 4); }

// Orders LDS accesses among the lanes of one wavefront (the LDS unit runs a
// wave's instructions in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// makes every wave wait until its global stores are acknowledged
// (s_waitcnt vmcnt(0) before s_barrier); this one waits for LDS operations
// (lgkmcnt) alone, so stores issued just before it drain in the background.
// Direct global -> LDS loads are vector-memory operations: a wave that issued
// some waits for them itself (wait_vmem) before such a barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s_waitcnt vmcnt(0): every vector-memory operation of this wave (loads,
// direct global -> LDS loads, stores) has completed.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <uint32_t M>
constexpr uint32_t exchange_stride() {
    constexpr uint32_t T = M / 16, base = M + M / 16;
    return base + (T + 32 - base % 32) % 32;
}

// ---------------------------------------------------------------------------
// Column transforms (first / last n1 levels).  T = N1/16 threads per column,
// 16 registers each; 4096/N1 columns per workgroup, all of one prime.  The
// exchange between the two register rounds goes through LDS as FP64.
// ---------------------------------------------------------------------------
constexpr size_t kExchangeBytes = size_t(kThreads / 16) * exchange_stride<256>() * sizeof(double);
// The column and middle kernels hold kExchangeBytes (35 KB) of LDS per
// workgroup, so a CU runs four of them: four waves per SIMD.  Telling the
// compiler so (instead of letting it aim for eight, which these kernels can
// never reach) keeps it from trading latency for registers, such as issuing
// each twiddle load just before its use.
#define kNttOccupancy __attribute__((amdgpu_waves_per_eu(4, 4)))
static_assert(4 * kExchangeBytes <= 160 * 1024 && 5 * kExchangeBytes > 160 * 1024, "four per CU");

// Forward columns take canonical encode words; level l adds at most p/2 +
// 2^-13 p, so every twiddle product's input is below p + 7 (p/2 + 2^-13 p) <
// 4.6 p < 2^33.3 and |q| k < 2^33.3 2^18 < 2^53 for s >= 13 (FAST holds for
// every plan).
template <int NLOG1, bool FAST = true>
__device__ __forceinline__ void fwd_cols_wg(const DevTables& t, uint32_t* __restrict__ planes, uint32_t bx,
                                            uint32_t op, uint32_t prod, double* s) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T, S = exchange_stride<N1>();
    constexpr int K2 = NLOG1 - 4;
    static_assert(K2 == 0 || CPW * S * sizeof(double) <= kExchangeBytes, "exchange buffer");
    const uint32_t wg_per_prime = t.N2 / CPW;
    const uint32_t j = bx / wg_per_prime;
    const uint32_t r0 = (bx % wg_per_prime) * CPW;
    const uint32_t cw = threadIdx.x / T, tau = threadIdx.x % T;
    uint32_t* col = plane_ptr(t, planes, prod, op, j) + size_t(r0 + cw) * N1;
    const uint32_t* z = t.zf + size_t(j) * t.N;
    const double p = t.pd[2 * j], pinv = t.pd[2 * j + 1];
    double v[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = double(col[tau + T * x]);
    ct_round<4, FAST>(v, z, 1, 0, 0, p, pinv);
    if constexpr (K2 > 0) {
        double* sc = s + cw * S;
#pragma unroll
        for (int x = 0; x < 16; ++x) sc[skew(tau + T * x)] = v[x];
        wave_lds_sync();
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = sc[skew(16 * tau + x)];
        ct_round<K2, FAST>(v, z, 1, 4, tau * (16 >> K2), p, pinv);
#pragma unroll
        for (int x = 0; x < 16; x += 4)
            *reinterpret_cast<uint4*>(col + 16 * tau + x) =
                make_uint4(uint32_t(balanced(v[x], p, pinv)), uint32_t(balanced(v[x + 1], p, pinv)),
                           uint32_t(balanced(v[x + 2], p, pinv)), uint32_t(balanced(v[x + 3], p, pinv)));
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) col[x] = uint32_t(balanced(v[x], p, pinv));
    }
}

// Inverse columns take the middle kernel's words: balanced (|x| <= p/2 + 1)
// or its last modular products (|x| <= m0 = p/2 + 2^-12 p); each level at
// most doubles the largest value, so the inputs of the last (folded) level
// are below 2^n1 m0, |q| < 2^(n1-1) p (1 + 2^-11) + 1, and with s >= n1 + 9,
// |q| k < p^2 (1 + 2^-11) / 2^10 + 2^22 < 2^53 (FAST holds for every plan).
template <int NLOG1, bool FAST = true>
__device__ __forceinline__ void inv_cols_wg(const DevTables& t, uint32_t* __restrict__ planes, uint32_t bx,
                                            uint32_t prod, double* s) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T, S = exchange_stride<N1>();
    constexpr int K2 = NLOG1 - 4;
    const uint32_t wg_per_prime = t.N2 / CPW;
    const uint32_t j = bx / wg_per_prime;
    const uint32_t r0 = (bx % wg_per_prime) * CPW;
    const uint32_t cw = threadIdx.x / T, tau = threadIdx.x % T;
    uint32_t* col = plane_ptr(t, planes, prod, 0, j) + size_t(r0 + cw) * N1;
    const uint32_t* z = t.zi + size_t(j) * t.N;
    const double p = t.pd[2 * j], pinv = t.pd[2 * j + 1];
    double v[16];
    if constexpr (K2 > 0) {
#pragma unroll
        for (int x = 0; x < 16; x += 4) {
            const uint4 q4 = *reinterpret_cast<const uint4*>(col + 16 * tau + x);
            v[x] = word_value(q4.x);
            v[x + 1] = word_value(q4.y);
            v[x + 2] = word_value(q4.z);
            v[x + 3] = word_value(q4.w);
        }
        gs_round<K2, 0, FAST>(v, z, 1, 4, tau * (16 >> K2), p, pinv);
        double* sc = s + cw * S;
#pragma unroll
        for (int x = 0; x < 16; ++x) sc[skew(16 * tau + x)] = v[x];
        wave_lds_sync();
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = sc[skew(tau + T * x)];
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = word_value(col[x]);
    }
    gs_round<4, 1, FAST>(v, z, 1, 0, 0, p, pinv);
    // Level 0 (twiddle z[1]) with the scale folded in: (x + y) s and
    // (x - y) (z[1] s), one modular product per output instead of two for
    // the difference.
    const double sc2 = t.scale[j], sz = t.scale_z1[j];
    const double sc2q = tw_quot<FAST>(sc2, pinv), szq = tw_quot<FAST>(sz, pinv);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const double d = v[x] - v[x + 8];
        v[x] = tw_mul<FAST>(v[x] + v[x + 8], sc2, sc2q, p, pinv);
        v[x + 8] = tw_mul<FAST>(d, sz, szq, p, pinv);
    }
#pragma unroll
    for (int x = 0; x < 16; ++x) col[tau + T * x] = canon_product(v[x], p);
}

// Twiddles of a radix-16 round (K = 4, one sub-block): tw[2^l - 1 + g] is the
// twiddle of group g at local level l, loaded ahead of the round that uses it
// and widened to FP64 there (u32 halves the registers they hold meanwhile).
// (LEVELS < 4 loads only the round's first LEVELS levels.)
template <int LEVELS = 4>
__device__ __forceinline__ void load_tw16(uint32_t (&tw)[15], const uint32_t* __restrict__ z, uint32_t mult,
                                          uint32_t e0, uint32_t o0) {
#pragma unroll
    for (int l = 0; l < LEVELS; ++l)
#pragma unroll
        for (int g = 0; g < (1 << l); ++g) tw[(1 << l) - 1 + g] = z[(mult << (e0 + l)) + (o0 << l) + g];
}

// Forward radix-16 round on two operands of the same prime and block at once:
// each twiddle serves two independent butterfly chains.  (LEVELS = 3 leaves
// the last level to mid_pair_step.)
template <bool FAST, int LEVELS = 4>
__device__ __forceinline__ void ct16x2_tw(double (&v)[16], double (&u)[16], const uint32_t (&tw)[15], double p,
                                          double pinv) {
#pragma unroll
    for (int l = 0; l < LEVELS; ++l) {
        const int half = 1 << (3 - l);
#pragma unroll
        for (int g = 0; g < (1 << l); ++g) {
            const double w = double(tw[(1 << l) - 1 + g]), wq = tw_quot<FAST>(w, pinv);
#pragma unroll
            for (int x = 0; x < half; ++x) {
                const int i0 = g * 2 * half + x;
                ct_bfly<FAST>(v[i0], v[i0 + half], w, wq, p, pinv);
                ct_bfly<FAST>(u[i0], u[i0 + half], w, wq, p, pinv);
            }
        }
    }
}

// Inverse radix-16 round from local level TOP down (TOP = 2: the round's
// first level was done by mid_pair_step).
template <bool FAST, int TOP = 3>
__device__ __forceinline__ void gs16_tw(double (&v)[16], const uint32_t (&tw)[15], double p, double pinv) {
#pragma unroll
    for (int l = TOP; l >= 0; --l) {
        const int half = 1 << (3 - l);
#pragma unroll
        for (int g = 0; g < (1 << l); ++g) {
            const double w = double(tw[(1 << l) - 1 + g]), wq = tw_quot<FAST>(w, pinv);
#pragma unroll
            for (int x = 0; x < half; ++x) {
                const int i0 = g * 2 * half + x;
                gs_bfly<FAST>(v[i0], v[i0 + half], w, wq, p, pinv);
            }
        }
    }
}

// The last forward level, the pointwise product and the first inverse level
// of each pair (2g, 2g + 1), in one step.  With the pair's forward twiddle w
// (its inverse twiddle is w^-1), those three levels map (a0, a1), (b0, b1) to
//   2 (a0 b0 + w^2 a1 b1)  and  2 (a0 b1 + a1 b0).
// This returns both without the factor 2 (the plan's inverse scale is
// (N/2)^-1 for it), the second as (a0 + a1)(b0 + b1) - a0 b0 - a1 b1: three
// products where the levels take four twiddle and two pointwise products.
// w^2 needs no table: z[k] = psi^brv(k) has z[2m + b]^2 = (-1)^b z[m], so
// the square of group g's twiddle is (-1)^g times twiddle tw[3 + g/2] of the
// level before.  Inputs |a|, |b| < 4.01 p, so the sums' products stay below
// 2^68, where mul_mod is exact; the outputs are below 2 m and 3 m, m the
// largest modular product.
template <bool FAST>
__device__ __forceinline__ void mid_pair_step(double (&v)[16], const double (&u)[16], const uint32_t (&tw)[15],
                                              double p, double pinv) {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const double a0 = v[2 * g], a1 = v[2 * g + 1], b0 = u[2 * g], b1 = u[2 * g + 1];
        const double c00 = mul_mod(a0, b0, p, pinv);
        const double c11 = mul_mod(a1, b1, p, pinv);
        const double cs = mul_mod(a0 + a1, b0 + b1, p, pinv);
        const double w2 = double(tw[3 + (g >> 1)]);
        const double t = tw_mul<FAST>(c11, w2, tw_quot<FAST>(w2, pinv), p, pinv);
        v[2 * g] = (g & 1) ? c00 - t : c00 + t;
        v[2 * g + 1] = cs - c00 - c11;
    }
}

// ---------------------------------------------------------------------------
// Middle: 16 blocks (256 contiguous natural coefficients each) of one prime,
// forward levels n1..n-1 for A and B, pointwise product, inverse levels back
// (the last forward level, the product and the first inverse level as one
// pair step).
// Grid (N1/16, batch, P): workgroups of one prime run together so its block
// twiddles stay in L2 across the whole batch; each radix-16 round's 15
// twiddles are loaded while the round before it computes.  One exchange
// buffer holds the tile of A and B (u32), then each wave's rows of A, of B
// and of the product in turn (FP64).
//
// Twiddle products (MODE): 0 mul_mod throughout; 1 the five-operation form
// except in the last four inverse levels; 2 the five-operation form
// throughout.  Bounds, with m0 = p/2 + 2^-12 p the largest modular product:
// the forward levels start from balanced words and add at most m0 per level
// (the pair step's inputs < p/2 + 1 + 7 m0 < 4.01 p); the pair step returns
// words below 2 m0 and 3 m0, and each later inverse level at most doubles
// them, so the differences that inverse level i (i = 2 .. 8, the pair step
// being level 1) multiplies are below 3 m0 2^(i-1).  So |q| k < 12.01 p 2^18
// < 2^53 for the forward levels and for i <= 4 and every plan (s >= 13),
// while the last four levels reach |x| < 384 m0 < 192 p + 3 p / 32 (< 2^39),
// which the plan checks prime by prime (kPlanTw5Late; always true for
// N = 65,536, where s >= 17).
// ---------------------------------------------------------------------------
constexpr uint32_t kMidS = exchange_stride<256>();

// Thread u owns row u of the tile: 16 consecutive blocks (64 contiguous bytes).
__device__ __forceinline__ void mid_fetch(uint4 (&q4)[4], const uint32_t* __restrict__ plane, uint32_t N1,
                                          uint32_t b0) {
    const uint4* src = reinterpret_cast<const uint4*>(plane + size_t(threadIdx.x) * N1 + b0);
#pragma unroll
    for (int k = 0; k < 4; ++k) q4[k] = src[k];
}

template <int MODE>
__device__ __forceinline__ void middle_wg(const DevTables& t, uint32_t* __restrict__ planes, uint32_t bx,
                                          uint32_t prod, uint32_t j, double* sx) {
    static_assert(kMidBlocks * kMidS * sizeof(double) <= kExchangeBytes, "exchange buffer");
    static_assert(MODE >= 0 && MODE <= 2, "twiddle-product mode");
    constexpr bool F1 = MODE >= 1, F2 = MODE == 2;
    const uint32_t b0 = bx * kMidBlocks;
    const double p = t.pd[2 * j], pinv = t.pd[2 * j + 1];
    uint32_t* pa = plane_ptr(t, planes, prod, 0, j);
    uint32_t* pb = plane_ptr(t, planes, prod, 1, j);
    const uint32_t bl = threadIdx.x >> 4, tau = threadIdx.x & 15;
    const uint32_t mult = t.N1 + b0 + bl;
    const uint32_t* zf = t.zf + size_t(j) * t.N;
    const uint32_t* zi = t.zi + size_t(j) * t.N;
    // The exchange buffer first holds both operands as 32-bit residue words,
    // laid out per wave: wave w's four blocks of A, then its four of B, in
    // bytes [32 w kMidS, 32 (w + 1) kMidS), which is exactly where its blocks'
    // FP64 rows (kMidS doubles per block) lie.  So every later exchange of a
    // wave (A, then B, then the product, each as FP64) touches only its own
    // bytes, and needs only a wave-level sync.
    uint32_t* uw = reinterpret_cast<uint32_t*>(sx);
    auto arow = [](uint32_t blk) { return (blk >> 2) * (8 * kMidS) + (blk & 3) * kMidS; };
    uint32_t* ra = uw + arow(bl);
    uint32_t* rb = ra + 4 * kMidS;
    double* rx = sx + bl * kMidS;
    double v[16], w[16];
    uint32_t twa[15], twb[15];
    load_tw16(twa, zf, mult, 0, 0);
    {
        uint4 qa[4], qb[4];
        mid_fetch(qa, pa, t.N1, b0);
        mid_fetch(qb, pb, t.N1, b0);
        const uint32_t o = skew(threadIdx.x);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t* da = uw + k * (8 * kMidS) + o;  // blocks 4k .. 4k + 3 (wave k's), A then B
            da[0 * kMidS] = qa[k].x;
            da[1 * kMidS] = qa[k].y;
            da[2 * kMidS] = qa[k].z;
            da[3 * kMidS] = qa[k].w;
            da[4 * kMidS] = qb[k].x;
            da[5 * kMidS] = qb[k].y;
            da[6 * kMidS] = qb[k].z;
            da[7 * kMidS] = qb[k].w;
        }
    }
    __syncthreads();
    // forward sub-levels 0..3 (positions tau + 16x) of A and B together
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        v[x] = word_value(ra[skew(tau + 16 * x)]);
        w[x] = word_value(rb[skew(tau + 16 * x)]);
    }
    load_tw16<3>(twb, zf, mult, 4, tau);
    ct16x2_tw<F1>(v, w, twa, p, pinv);
    // Exchange A, then B, as FP64 through the block's row: no reduction (the
    // values stay below p/2 + 1 + 4 m0 < 2.51 p) and no conversions.
    wave_lds_sync();  // this wave's u32 rows (under its FP64 rows) are read
#pragma unroll
    for (int x = 0; x < 16; ++x) rx[skew(tau + 16 * x)] = v[x];
    wave_lds_sync();
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = rx[skew(16 * tau + x)];
    wave_lds_sync();
#pragma unroll
    for (int x = 0; x < 16; ++x) rx[skew(tau + 16 * x)] = w[x];
    wave_lds_sync();
    // sub-levels 4..6 (positions 16 tau + x), then sub-level 7, the
    // pointwise product and inverse sub-level 7 as one pair step
#pragma unroll
    for (int x = 0; x < 16; ++x) w[x] = rx[skew(16 * tau + x)];
    load_tw16<3>(twa, zi, mult, 4, tau);
    ct16x2_tw<F1, 3>(v, w, twb, p, pinv);
    mid_pair_step<F1>(v, w, twb, p, pinv);
    // inverse sub-levels 6..4
    load_tw16(twb, zi, mult, 0, 0);
    gs16_tw<F1, 2>(v, twa, p, pinv);
    wave_lds_sync();  // every lane of the wave has read B from the row
#pragma unroll
    for (int x = 0; x < 16; ++x) rx[skew(16 * tau + x)] = v[x];
    wave_lds_sync();
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = rx[skew(tau + 16 * x)];
    gs16_tw<F2>(v, twb, p, pinv);
    __syncthreads();
    {
        // balanced residues back in place, then row u out as four 16-byte stores;
        // v[8..15] come straight from the last level's modular products
        // (|r| <= p/2 + 2^-12 p, below 2^31), so they only need converting
        uint32_t* sy = reinterpret_cast<uint32_t*>(sx);
#pragma unroll
        for (int x = 0; x < 16; ++x)
            sy[bl * kMidS + skew(tau + 16 * x)] = uint32_t(x < 8 ? balanced(v[x], p, pinv) : int32_t(v[x]));
        __syncthreads();
        const uint32_t u = threadIdx.x;
        uint4* dst = reinterpret_cast<uint4*>(pa + size_t(u) * t.N1 + b0);
        const uint32_t o = skew(u);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            dst[k] = make_uint4(sy[(4 * k) * kMidS + o], sy[(4 * k + 1) * kMidS + o],
                                sy[(4 * k + 2) * kMidS + o], sy[(4 * k + 3) * kMidS + o]);
    }
}

// ---------------------------------------------------------------------------
// Encode fused with the forward column levels (N1 = 64, 128 or 256): one
// workgroup takes 256 coefficients, CPW = 256 / N1 whole columns r0 .. r0 +
// CPW - 1 (coefficients r + 256 s, s < N1) of one operand.  Their input bytes
// are staged through LDS in two halves of 128 coefficients; each wave keeps
// the MFMA data fragments of its 64 coefficients in registers for the whole
// kernel.  Then, per pair of weight tiles (16 primes), the waves' residues go
// into an FP64 exchange of 16 primes x CPW columns x N1 positions and each
// group of N1 / 16 threads runs the column transform of one (prime, column)
// exactly as k_fwd_cols does, storing balanced residues.  Everything a pair
// reads from global memory besides the data arrives ahead of use: its weight
// tiles by direct global -> LDS loads during the previous pair's transform,
// its primes' (p, 1/p) and first-round twiddles the same way during its MFMA
// block, the second-round twiddles into registers during the first round.
// ---------------------------------------------------------------------------
constexpr uint32_t kEfS = exchange_stride<256>();
constexpr uint32_t kEfTw = 16;  // first-round twiddle slots per prime: z_j[0 .. 15] as FP64 (z_j[0] unused)

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// Dynamic LDS of the fused kernel: the staging area (later the exchange),
// the per-prime dot-sum offsets, then one pair of weight tiles.  The current
// pair's (p, 1/p) and first-round twiddles (FP64) sit in 2,304 bytes of static LDS.
size_t ef_smem_bytes(const DevTables& t) {
    const size_t stage = size_t(kEncCoefs) * (t.Kenc + 16), exch = size_t(16) * kEfS * sizeof(double);
    // + one pair of weight tiles (64 rows of Kenc + 16 bytes) when Kenc <= 224;
    // at Kenc = 256 that would leave room for only two workgroups per CU, so
    // that width reads its weight fragments from global memory instead
    const size_t wtiles = t.Kenc <= 224 && !(t.kflags & kFlagWeightsL2) ? size_t(64) * (t.Kenc + 16) : 0;
    return std::max(stage, exch) + sizeof(double) * kMaxPrimesCrt + wtiles;
}

// The twiddles a ct_round<K>(v, z, mult, e0, o0) uses, in its order of use
// (loaded ahead of the round; (16 >> K) (2^K - 1) <= 15 of them).
template <int K>
__device__ __forceinline__ void load_round_tw(uint32_t (&tw)[15], const uint32_t* __restrict__ z, uint32_t mult,
                                              uint32_t e0, uint32_t o0) {
    int n = 0;
#pragma unroll
    for (int l = 0; l < K; ++l)
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb)
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) tw[n++] = z[(mult << (e0 + l)) + ((o0 + sb) << l) + g];
}

template <int K>
__device__ __forceinline__ void ct_round_pre(double (&v)[16], const uint32_t (&tw)[15], double p, double pinv) {
    int n = 0;
#pragma unroll
    for (int l = 0; l < K; ++l) {
        const int half = 1 << (K - 1 - l);
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb) {
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) {
                const double w = double(tw[n++]);
#pragma unroll
                for (int x = 0; x < half; ++x) {
                    const int i0 = (sb << K) + g * 2 * half + x;
                    ct_bfly(v[i0], v[i0 + half], w, p, pinv);
                }
            }
        }
    }
}

// One forward radix-16 round (column levels 0..3) with its 15 twiddles
// tw[k] = z[k], k = 1 .. 15, already widened to FP64 (in LDS).
__device__ __forceinline__ void ct_round_tw15(double (&v)[16], const double* tw, double p, double pinv) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int half = 8 >> l;
#pragma unroll
        for (int g = 0; g < (1 << l); ++g) {
            const double w = tw[(1 << l) + g];
#pragma unroll
            for (int x = 0; x < half; ++x) ct_bfly(v[g * 2 * half + x], v[g * 2 * half + x + half], w, p, pinv);
        }
    }
}

template <int KS, int NLOG1, bool STAGE_W>
__device__ __forceinline__ void encfwd_body(const DevTables& t, const uint32_t* __restrict__ a,
                                            const uint32_t* __restrict__ b, uint32_t* __restrict__ planes,
                                            uint32_t bx, uint32_t op, uint32_t prod, uint8_t* smem,
                                            double* pbuf, double* twbuf) {
    static_assert(NLOG1 >= 6 && NLOG1 <= 8, "fused encode covers N1 = 64 .. 256");
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = 256 / N1, S = exchange_stride<N1>();
    constexpr int K2 = NLOG1 - 4;
    constexpr uint32_t ts_log = NLOG1 < 7 ? NLOG1 : 7, TRh = kEncCoefs >> ts_log;  // r per staging half
    static_assert(16 * CPW * S == 16 * kEfS, "exchange size");
    const uint32_t r0 = bx * CPW;
    const size_t stage = size_t(kEncCoefs) * (t.Kenc + 16), exch = size_t(16) * kEfS * sizeof(double);
    double* offs = reinterpret_cast<double*>(smem + (stage > exch ? stage : exch));
    uint8_t* wbuf = reinterpret_cast<uint8_t*>(offs + kMaxPrimesCrt);  // weight tiles of the current pair
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Each pair's two weight tiles are copied into LDS once per workgroup (the
    // four waves then read their fragments there, instead of each loading
    // them from L2), the next pair's during the current transform, by direct
    // global -> LDS loads: no registers, and no wave waits for them before
    // the barrier that ends the pair.  Slot k (16 bytes) of the tiles is
    // byte 16 (k mod (2 KS + 1)) of row k / (2 KS + 1) (the last slot of a
    // row is padding); wave instruction n fills slots 64 n .. 64 n + 63.
    // (kFlagWeightsL2 switches the staging off: every wave loads its fragments from L2)
    constexpr bool stage_w = STAGE_W;
    auto load_weights = [&](uint32_t grp) {
        if constexpr (!stage_w) return;
        constexpr uint32_t SPR = 2 * KS + 1;
        // k / SPR for k < 2^12 as a 24-bit multiply by ceil(2^20 / SPR): exact,
        // since k (2^20 / SPR rounded up - 2^20 / SPR) / 2^20 < 2^-8 < 1 / SPR
        constexpr uint32_t kInvSpr = ((1u << 20) + SPR - 1) / SPR;
        const uint32_t slots = min(64u, (t.PTenc - 2 * grp) * 32) * SPR;
        const int8_t* src = t.wenc + size_t(2 * grp) * 32 * (32 * KS);
        for (uint32_t n = wave; n * 64 < slots; n += 4) {
            const uint32_t k = n * 64 + lane;
            if (k < slots) {
                const uint32_t rw = __umul24(k, kInvSpr) >> 20, c16 = min(k - __umul24(rw, SPR), 2 * KS - 1);
                __builtin_amdgcn_global_load_lds(src + rw * (32 * KS) + 16 * c16, (lds_void_ptr)(wbuf + n * 1024),
                                                 16, 0, 0);
            }
        }
    };
    // (p, 1/p) and the FP64 row z_j[0 .. 15] of the pair's primes j = 16 grp +
    // jt (clamped to the last prime), the same way: lanes 0..15 of wave 2
    // take one 16-byte (p, 1/p) each, and lane k of waves 0 and 1 the 16-byte
    // piece k & 7 of prime k >> 3's row (wave w's 64 pieces land at twbuf +
    // 128 w).
    auto load_pair_consts = [&](uint32_t grp) {
        if (wave == 2 && lane < 16)
            __builtin_amdgcn_global_load_lds(t.pd + 2 * min(16 * grp + lane, t.P - 1), (lds_void_ptr)pbuf, 16, 0, 0);
        if (wave < 2) {
            const uint32_t k = wave * 64 + lane;
            __builtin_amdgcn_global_load_lds(t.zf16 + size_t(min(16 * grp + (k >> 3), t.P - 1)) * kEfTw + 2 * (k & 7),
                                             (lds_void_ptr)(twbuf + wave * 128), 16, 0, 0);
        }
    };
    for (uint32_t j = threadIdx.x; j < t.P; j += kThreads)
        offs[j] = reinterpret_cast<const EncPrime*>(t.enc_prime)[j].off;
    const uint32_t* in = (op ? b : a) + size_t(prod) * t.N * t.L;
    const uint32_t row = lane & 31, kh = lane >> 5;
    const uint32_t stride_bytes = t.Kenc + 16;
    // wave w holds coefficients s = (2 w + c) * 32 + row, c = 0, 1 (half w >> 1)
    v4i B[2][KS];
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        // half h: coefficients 128 h .. 128 h + 127 of the workgroup
        const EncGeom e = NLOG1 == 8 ? EncGeom(t, 7, r0, kEncCoefs * h) : EncGeom(t, ts_log, r0 + h * TRh, 0);
        u32x4 regs[KS];
        enc_load<KS>(t, in, e, regs);
        if (h) __syncthreads();  // the first half's fragments have been read
        enc_store<KS>(t, e, regs, reinterpret_cast<uint32_t*>(smem));
        __syncthreads();
        if ((wave >> 1) == h) {
            const uint32_t c0 = 2 * (wave & 1);
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int s2 = 0; s2 < KS; ++s2)
                    B[c][s2] = *reinterpret_cast<const v4i*>(smem + ((c0 + c) * 32 + row) * stride_bytes + 16 * kh +
                                                             32 * s2);
        }
    }
    load_weights(0);
    __syncthreads();  // the staging area becomes the exchange; the first pair's weights are in
    double* X = reinterpret_cast<double*>(smem);
    const uint32_t q = threadIdx.x / T, jt = q / CPW, cq = q % CPW, tau = threadIdx.x % T;
    const uint32_t npairs = (t.PTenc + 1) / 2;
    for (uint32_t grp = 0; grp < npairs; ++grp) {
        load_pair_consts(grp);  // lands during the MFMA block (the barrier after it waits for it)
        // MFMA and VALU instructions share a SIMD's issue port: a wave in its
        // MFMA block goes first, so the matrix pipe stays fed while the other
        // resident workgroups' transforms fill the gaps (measured 1.5% faster)
        if (!(t.kflags & kFlagNoPrio)) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (uint32_t u = 0; u < 2; ++u) {
            const uint32_t pt = 2 * grp + u;
            if (pt < t.PTenc) {
                v4i A[KS];
                if constexpr (stage_w) {
                    // an LDS-typed pointer: the two branches must not merge into generic (flat) loads
                    typedef const __attribute__((address_space(3))) v4i lds_v4i;
                    lds_v4i* wrow = (lds_v4i*)(wbuf + (u * 32 + row) * stride_bytes + 16 * kh);
#pragma unroll
                    for (int s2 = 0; s2 < KS; ++s2) A[s2] = wrow[2 * s2];
                } else {
                    const int8_t* wrow = t.wenc + (size_t(pt) * 32 + row) * t.Kenc + 16 * kh;
#pragma unroll
                    for (int s2 = 0; s2 < KS; ++s2) A[s2] = *reinterpret_cast<const v4i*>(wrow + 32 * s2);
                }
                double off[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) off[g] = offs[min(pt * 8 + 2 * g + kh, t.P - 1)];
                v16i acc[2] = {v16i{}, v16i{}};
#pragma unroll
                for (int s2 = 0; s2 < KS; ++s2) {
                    acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[s2], B[0][s2], acc[0], 0, 0, 0);
                    acc[1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[s2], B[1][s2], acc[1], 0, 0, 0);
                }
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const uint32_t m = (2 * wave + c) * 32 + row;  // coefficient r0 + (m >> NLOG1) + 256 (m & (N1 - 1))
                    const uint32_t xrow = (m >> NLOG1) * S + skew(m & (N1 - 1));
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int32_t lo = acc[c][4 * g] + acc[c][4 * g + 1] * 256;
                        const int32_t hi = acc[c][4 * g + 2] + acc[c][4 * g + 3] * 256;
                        // u = x (mod p) in [0, 2^47), exact; the transform's mul_mod and
                        // sums stay exact on it (|u w / p| < 2^48), so no reduction here
                        X[(8 * u + 2 * g + kh) * CPW * S + xrow] = fma(double(hi), 65536.0, double(lo)) + off[g];
                    }
                }
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();  // every wave is done with this pair's weights; its constants are in
        if (grp + 1 < npairs) load_weights(grp + 1);  // lands during the transform
        // forward column levels 0..n1-1 of prime j, column r0 + cq (k_fwd_cols)
        const uint32_t j = 16 * grp + jt;
        double* xc = X + q * S;
        const bool live = j < t.P;
        const uint32_t jc = live ? j : 0;
        const uint32_t* z = t.zf + size_t(jc) * t.N;
        const double p = pbuf[2 * jt], pinv = pbuf[2 * jt + 1];
        double v[16];
        uint32_t tw2[15];  // the second round's twiddles, in flight during the first (when registers allow)
        if (live) {
            if constexpr (stage_w) load_round_tw<K2>(tw2, z, 1, 4, tau * (16 >> K2));
#pragma unroll
            for (int x = 0; x < 16; ++x) v[x] = xc[skew(tau + T * x)];
            ct_round_tw15(v, twbuf + jt * kEfTw, p, pinv);
#pragma unroll
            for (int x = 0; x < 16; ++x) xc[skew(tau + T * x)] = v[x];
        }
        wave_lds_sync();
        // This wave's direct-to-LDS loads of the next pair's weight tiles
        // (issued before the first round) are complete from here on (the
        // second round waits for its twiddle loads anyway), so the barrier at
        // the end of the pair waits for LDS operations only, not for the
        // stores below.
        wait_vmem();
        if (live) {
#pragma unroll
            for (int x = 0; x < 16; ++x) v[x] = xc[skew(16 * tau + x)];
            if constexpr (stage_w)
                ct_round_pre<K2>(v, tw2, p, pinv);
            else
                ct_round<K2>(v, z, 1, 4, tau * (16 >> K2), p, pinv);
            uint32_t* col = plane_ptr(t, planes, prod, op, j) + size_t(r0 + cq) * N1;
#pragma unroll
            for (int x = 0; x < 16; x += 4)
                *reinterpret_cast<uint4*>(col + 16 * tau + x) =
                    make_uint4(uint32_t(balanced(v[x], p, pinv)), uint32_t(balanced(v[x + 1], p, pinv)),
                               uint32_t(balanced(v[x + 2], p, pinv)), uint32_t(balanced(v[x + 3], p, pinv)));
        }
        lds_barrier();  // the exchange, constants and weights are rewritten for the next pair
    }
}

// The weight staging (off at Kenc = 256 and under kFlagWeightsL2) is chosen
// per launch; each choice is its own code path (a branch merged inside the
// pair loop would make every wave wait for the staging loads of the other).
template <int KS, int NLOG1>
__device__ __forceinline__ void encfwd_wg(const DevTables& t, const uint32_t* __restrict__ a,
                                          const uint32_t* __restrict__ b, uint32_t* __restrict__ planes,
                                          uint32_t bx, uint32_t op, uint32_t prod, uint8_t* smem) {
    // statically allocated, apart from the dynamic area, so the compiler sees
    // that their loads do not alias the exchange and the weight tiles
    __shared__ __align__(16) double pbuf[32];              // [16][2]: p, 1/p of the pair's primes
    __shared__ __align__(16) double twbuf[16 * kEfTw];    // [16][16]: z_j[0 .. 15]
    if (KS <= 7 && !(t.kflags & kFlagWeightsL2))
        encfwd_body<KS, NLOG1, KS <= 7>(t, a, b, planes, bx, op, prod, smem, pbuf, twbuf);
    else
        encfwd_body<KS, NLOG1, false>(t, a, b, planes, bx, op, prod, smem, pbuf, twbuf);
}

// ---------------------------------------------------------------------------
// CRT: 64 coefficients (consecutive column-major positions), four waves.
// Terms t_j are in operand-A planes.  alpha = round(sum t_j / p_j): the plan's
// guard P > 4 N (q-1)^2 puts the exact sum in [alpha, alpha + 1/4), and the
// computed one (FP64 per-wave shares rounded to FP32, summed in FP64) is
// within 2^-16 of it.  Then
//   z = sum_{j<=P, a<4} byte_a(t_j) * W[j,a],  W[j,a] = 256^a w_j mod q,
// as an int8 MFMA GEMM over balanced signed weight digits and biased bytes
// (wave w computes digit-row tiles w and w+4; the terms sit in LDS already
// biased, byte ^ 0x80).  Every dot product is below 2^23 in magnitude (at most
// 4 (P+1) <= 516 products of two values in [-128, 127]), so the int64 sum of a
// limb's four digit rows, 2^24 a3 + 2^16 a2 + 2^8 a1 + a0 plus the bias
// correction, stays below 2^49 and is formed exactly in FP64.  The quotient
// Q = floor(z / q) is estimated before any carry from the top four limb sums
// (truncation error below 2^-48 q), and one pass of exact FP64 limb steps
//   t = s_i - Q q_i + carry,  carry = floor(t / 2^32),  digit = t - carry 2^32
// (|t| < 2^52) yields z - Q q; a wrong estimate (z / q within 2^-40 of an
// integer) leaves it in [-q, 2q), which one rare add or subtract of q fixes.
// For q = 2^W - gap (0 < gap < 2^32, L >= 3) the pass uses z - Q q = z +
// Q gap - Q 2^W instead: Q < 2^18, so Q gap and Q 2^W are a few exact limbs
// added to limbs 0, 1, L-1 and the top, and the steps need no limbs of q.
//
// One LDS region serves three phases: the term rows (GEMM operand), then the
// int64 limb sums [L+1][64] (written once every wave has left the GEMM; the
// accumulators wait in registers across the barrier), then the output digits
// [L][65] u32 written in place over the limb sums by the carry chain.  Digit
// (i, c) occupies bytes [260 i + 4 c, +4), inside [260 i, 260 i + 256), and the
// limb sums not yet read at step i start at byte 512 (i + 1), so no digit write
// reaches an unread sum.  The stride 65 keeps the output stage's reads (65
// digits apart across a wave) on distinct banks.
// ---------------------------------------------------------------------------
constexpr uint32_t kCrtThreads = 256;
constexpr int kCrtAhead = 4;  // k-steps the weight loads run ahead of the GEMM
// Workgroups per CU the CRT kernels are built for: the prefetched weight
// fragments need up to 128 VGPRs (five per CU left 102 and serialised the
// loads instead; four with the prefetch were 10% faster at Kcrt = 480).
constexpr int kCrtWgPerCu = 4;
// The alternative the autotuner may pick (kFlagCrtWide): five workgroups per
// CU, with the weight fragments loaded one k-step ahead.
constexpr int kCrtAheadWide = 1;
constexpr int kCrtWgPerCuWide = 5;
constexpr uint32_t kCrtDigitStride = kCrtCoefs + 1;

size_t crt_smem_bytes(const DevTables& t) {
    // term rows (their last 16 bytes hold the waves' partial alpha sums), later
    // the limb sums
    return std::max(size_t(kCrtCoefs) * (t.Kcrt + 16), size_t(t.L + 1) * kCrtCoefs * 8);
}

// Row tiles per wave the instantiations cover: the digit-row tiles Mcrt / 32
// are at most 4 when Kcrt / 32 <= 8 and at most 8 otherwise.
__host__ __device__ constexpr uint32_t crt_tiles_per_wave(uint32_t ks, uint32_t mt) {
    return ks <= 8 ? 1 : (ks == 9 && mt <= 4 ? 1 : 2);
}

// Sum of the four digit rows 4 g .. 4 g + 3 of an MFMA accumulator, weighted by
// 256^row, exactly in FP64 (|a| < 2^23 keeps every partial sum below 2^48).
__device__ __forceinline__ double limb_sum(const v16i& a, int g) {
    return fma(double(a[4 * g + 3]), 16777216.0,
               fma(double(a[4 * g + 2]), 65536.0, fma(double(a[4 * g + 1]), 256.0, double(a[4 * g]))));
}

// CRT for 64 coefficients with the k-loop of the GEMM (Kcrt / 32 steps), the
// row tiles per wave (KT, four waves) and the weight look-ahead (D k-steps)
// fixed.
template <int KS, int KT, int D>
__device__ __forceinline__ void crt_wg(const DevTables& t, const uint32_t* __restrict__ planes,
                                       uint32_t* __restrict__ out, uint32_t bx, uint32_t prod,
                                       uint8_t* crt_smem) {
    const uint32_t L = t.L, P = t.P;
    const uint32_t tstride = t.Kcrt + 16;  // bytes per term row (odd multiple of 16: conflict-free b128)
    const uint32_t tw = tstride / 4;
    const uint32_t groups = t.Kcrt / 16;   // 4-term groups per row
    uint8_t* tmat = crt_smem;
    uint32_t* tmw = reinterpret_cast<uint32_t*>(crt_smem);
    double* lo = reinterpret_cast<double*>(crt_smem);    // [L+1][64] after the GEMM
    uint32_t* dig = tmw;                                 // [L][65] over the limb sums
    const uint32_t f0 = bx * kCrtCoefs;
    const uint32_t* pa = planes + size_t(prod) * 2 * P * t.N + f0;
    {
        // lane c gathers four consecutive terms of its coefficient (coalesced per
        // plane); every load of a thread is issued before its LDS stores.  The
        // wave's share of alpha comes straight from these registers (the term
        // indices, hence the 1/p_j loads, are uniform across the wave).
        constexpr uint32_t kIters = (KS * 2 + 3) / 4;
        const uint32_t c = threadIdx.x & 63, g0 = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        uint32_t v[kIters][4];
#pragma unroll
        for (uint32_t it = 0; it < kIters; ++it) {
            const uint32_t g = g0 + 4 * it;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t j = min(4 * g + k, P - 1);  // clamped: every load unconditional
                v[it][k] = pa[size_t(j) * t.N + c];
            }
        }
#pragma unroll
        for (uint32_t it = 0; it < kIters; ++it) {
            const uint32_t g = g0 + 4 * it;
            if (g < groups) {
                uint32_t w4[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) w4[k] = (4 * g + k < P ? v[it][k] : 0u) ^ 0x80808080u;
                *reinterpret_cast<uint4*>(tmw + c * tw + 4 * g) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
            }
        }
        double part = 0.0;
#pragma unroll
        for (uint32_t it = 0; it < kIters; ++it) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t j = 4 * (g0 + 4 * it) + k;
                if (j < P) part = fma(double(v[it][k]), t.invp[j], part);  // one rounding per term
            }
        }
        // A share (at most 36 terms, each t_j / p_j < 1) is below 2^6, so FP32
        // keeps it within 2^-18; the sum's margin to the next integer is 3/4.
        reinterpret_cast<float*>(tmw + c * tw + t.Kcrt / 4)[g0] = float(part);
    }
    __syncthreads();
    if (threadIdx.x < kCrtCoefs) {
        // alpha = round(sum_j t_j / p_j); the sum lies within 1/4 above an
        // integer, far beyond the rounding of the shares, so their order is free.
        const uint32_t c = threadIdx.x;
        const float* pr = reinterpret_cast<const float*>(tmw + c * tw + t.Kcrt / 4);
        const double a = (double(pr[0]) + double(pr[1])) + (double(pr[2]) + double(pr[3]));
        tmw[c * tw + P] = uint32_t(a + 0.5) ^ 0x80808080u;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 31, kh = lane >> 5;
    const uint32_t MT = t.Mcrt / 32;
    v16i acc0[KT], acc1[KT];
    const int8_t* wrow[KT];
    bool live[KT];
#pragma unroll
    for (int i = 0; i < KT; ++i) {
        acc0[i] = v16i{};
        acc1[i] = v16i{};
        const uint32_t mt = wave + 4 * i;
        live[i] = mt < MT;
        wrow[i] = t.wcrt + size_t(min(mt, MT - 1) * 32 + row) * t.Kcrt + 16 * kh;  // clamped: in bounds
    }
    // Weight fragments (from L2) are loaded D k-steps ahead of the MFMAs that
    // use them, for both tiles of the wave at once; left to itself the
    // compiler issued each load just before its MFMA and waited an L2 round
    // trip per k-step (0.8 ms of a 2.55 ms batch).
    v4i A[KT][KS];
#pragma unroll
    for (int k = 0; k < KS && k < D; ++k)
#pragma unroll
        for (int i = 0; i < KT; ++i) A[i][k] = *reinterpret_cast<const v4i*>(wrow[i] + 32 * k);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        if (k + D < KS) {
#pragma unroll
            for (int i = 0; i < KT; ++i) A[i][k + D] = *reinterpret_cast<const v4i*>(wrow[i] + 32 * (k + D));
        }
        __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of the MFMAs below
        const v4i B0 = *reinterpret_cast<const v4i*>(tmat + row * tstride + 32 * k + 16 * kh);
        const v4i B1 = *reinterpret_cast<const v4i*>(tmat + (32 + row) * tstride + 32 * k + 16 * kh);
#pragma unroll
        for (int i = 0; i < KT; ++i) {
            if (live[i]) {
                acc0[i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][k], B0, acc0[i], 0, 0, 0);
                acc1[i] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i][k], B1, acc1[i], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // the term rows are dead from here on
#pragma unroll
    for (int i = 0; i < KT; ++i) {
        const uint32_t mt = wave + 4 * i;
        if (mt < MT) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const uint32_t limb = mt * 8 + 2 * g + kh;
                if (limb <= L) {
                    const double cr = double(t.corr[limb]);
                    lo[limb * kCrtCoefs + row] = limb_sum(acc0[i], g) + cr;
                    lo[limb * kCrtCoefs + 32 + row] = limb_sum(acc1[i], g) + cr;
                }
            }
        }
    }
    __syncthreads();

    if (threadIdx.x < kCrtCoefs) {
        // One wave: lanes step through the limbs together, so digit i is
        // written only after every lane has read limb sums 0..i.
        const uint32_t c = threadIdx.x;
        uint32_t* d = dig + c;  // digit i at d[i * kCrtDigitStride]
        double* s = lo + c;     // limb sum i at s[i * kCrtCoefs]
        // Q from the top four limb sums and limbs of q (units of 2^(32 (L-3)))
        const uint32_t b = L - 3;
        const double zt = ((s[L * kCrtCoefs] * 4294967296.0 + s[(L - 1) * kCrtCoefs]) * 4294967296.0 +
                           s[(L - 2) * kCrtCoefs]) * 4294967296.0 + s[b * kCrtCoefs];
        const double qe = (double(t.q[L - 1]) * 4294967296.0 + double(t.q[L - 2])) * 4294967296.0 + double(t.q[b]);
        const double Q = fmax(floor(zt / qe), 0.0);
        double carry = 0.0, top = s[L * kCrtCoefs];
        if (t.gap && !(t.kflags & kFlagGenericQ)) {
            // near-power q: Q gap and Q 2^r (r = W - 32 (L-1), the top limb's
            // bits) are below 2^50, exact, and split exactly into 32-bit halves
            const double qg = Q * double(t.gap), g1 = floor(qg * 0x1p-32), g0 = fma(-g1, 4294967296.0, qg);
            const double h = ldexp(Q, int(t.W - 32 * (L - 1))), h1 = floor(h * 0x1p-32);
            const double h0 = fma(-h1, 4294967296.0, h);
            s[0] += g0;
            s[kCrtCoefs] += g1;
            s[(L - 1) * kCrtCoefs] -= h0;
            top -= h1;
            for (uint32_t i0 = 0; i0 < L; i0 += 8) {
                double sv[8];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k) sv[k] = s[min(i0 + k, L) * kCrtCoefs];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k) {
                    const uint32_t i = i0 + k;
                    if (i < L) {
                        const double tt = sv[k] + carry;
                        carry = floor(tt * 0x1p-32);
                        d[i * kCrtDigitStride] = uint32_t(fma(-carry, 4294967296.0, tt));
                    }
                }
            }
        } else {
            const uint32_t qv = t.q[min(c, L - 1)];  // lane i holds q_i (L <= 64)
            for (uint32_t i0 = 0; i0 < L; i0 += 8) {
                // eight limb sums in flight (index clamped to L: every read in
                // bounds and not yet overwritten by a digit)
                double sv[8];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k) sv[k] = s[min(i0 + k, L) * kCrtCoefs];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k) {
                    const uint32_t i = i0 + k;
                    if (i < L) {
                        const double qi = double(uint32_t(__builtin_amdgcn_readlane(int(qv), int(i))));
                        const double tt = fma(-Q, qi, sv[k]) + carry;
                        carry = floor(tt * 0x1p-32);
                        d[i * kCrtDigitStride] = uint32_t(fma(-carry, 4294967296.0, tt));
                    }
                }
            }
        }
        // r = z - Q q = sum d_i 2^(32 i) + top 2^(32 L), -q <= r < 2q
        top += carry;
        bool sub = top > 0.0, add = top < 0.0;
        if (top == 0.0) {
            sub = true;  // r == q unless a limb differs
            for (int i = int(L) - 1; i >= 0; --i) {
                const uint32_t di = d[i * kCrtDigitStride];
                if (di != t.q[i]) {
                    sub = di > t.q[i];
                    break;
                }
            }
        }
        if (sub || add) {
            int64_t cy = 0;
            for (uint32_t i = 0; i < L; ++i) {
                const int64_t dd = int64_t(d[i * kCrtDigitStride]) + (add ? int64_t(t.q[i]) : -int64_t(t.q[i])) + cy;
                cy = dd >> 32;
                d[i * kCrtDigitStride] = uint32_t(dd);
            }
        }
    }
    __syncthreads();
    uint32_t* dst = out + size_t(prod) * t.N * L;
    const uint32_t w = threadIdx.x & 63;
    if (w < L) {
        for (uint32_t cc = threadIdx.x >> 6; cc < kCrtCoefs; cc += kCrtThreads / 64) {
            const uint32_t f = f0 + cc;
            const uint32_t n = (f >> t.n1) + t.N2 * (f & (t.N1 - 1));
            dst[size_t(n) * L + w] = dig[w * kCrtDigitStride + cc];
        }
    }
}

// CRT with the k-step count (Kcrt / 32, 1..17), the row tiles per wave and
// the weight look-ahead as compile-time constants.
template <int KS, int D = kCrtAhead>
__device__ __forceinline__ void crt_tiles(const DevTables& t, const uint32_t* __restrict__ planes,
                                          uint32_t* __restrict__ out, uint32_t bx, uint32_t prod, uint8_t* smem) {
    if constexpr (KS <= 8) {
        crt_wg<KS, 1, D>(t, planes, out, bx, prod, smem);
    } else if constexpr (KS == 9) {
        if (t.Mcrt / 32 <= 4)
            crt_wg<KS, 1, D>(t, planes, out, bx, prod, smem);
        else
            crt_wg<KS, 2, D>(t, planes, out, bx, prod, smem);
    } else {
        crt_wg<KS, 2, D>(t, planes, out, bx, prod, smem);
    }
}

__device__ __forceinline__ void crt_dispatch(const DevTables& t, const uint32_t* __restrict__ planes,
                                             uint32_t* __restrict__ out, uint32_t bx, uint32_t prod,
                                             uint8_t* smem) {
    switch (t.Kcrt / 32) {
        case 1: crt_tiles<1>(t, planes, out, bx, prod, smem); break;
        case 2: crt_tiles<2>(t, planes, out, bx, prod, smem); break;
        case 3: crt_tiles<3>(t, planes, out, bx, prod, smem); break;
        case 4: crt_tiles<4>(t, planes, out, bx, prod, smem); break;
        case 5: crt_tiles<5>(t, planes, out, bx, prod, smem); break;
        case 6: crt_tiles<6>(t, planes, out, bx, prod, smem); break;
        case 7: crt_tiles<7>(t, planes, out, bx, prod, smem); break;
        case 8: crt_tiles<8>(t, planes, out, bx, prod, smem); break;
        case 9: crt_tiles<9>(t, planes, out, bx, prod, smem); break;
        case 10: crt_tiles<10>(t, planes, out, bx, prod, smem); break;
        case 11: crt_tiles<11>(t, planes, out, bx, prod, smem); break;
        case 12: crt_tiles<12>(t, planes, out, bx, prod, smem); break;
        case 13: crt_tiles<13>(t, planes, out, bx, prod, smem); break;
        case 14: crt_tiles<14>(t, planes, out, bx, prod, smem); break;
        case 15: crt_tiles<15>(t, planes, out, bx, prod, smem); break;
        case 16: crt_tiles<16>(t, planes, out, bx, prod, smem); break;
        default: crt_tiles<17>(t, planes, out, bx, prod, smem); break;
    }
}

__global__ void k_mfma_probe(const int8_t* a, const int8_t* b, int32_t* c) {
    const uint32_t lane = threadIdx.x & 63, row = lane & 31, kh = lane >> 5;
    const v4i A = *reinterpret_cast<const v4i*>(a + row * 32 + 16 * kh);
    const v4i B = *reinterpret_cast<const v4i*>(b + row * 32 + 16 * kh);
    v16i acc = v16i{};
    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const uint32_t r = 8 * (i / 4) + 4 * kh + (i % 4);
        c[r * 32 + row] = acc[i];
    }
}

// Workgroups of each stage for one group: (x, y, z) of the single-plan grid.
__host__ __device__ inline uint32_t cols_per_wg(uint32_t n1) { return kThreads / ((1u << n1) / 16); }

__host__ __device__ inline void stage_grid(const DevTables& t, uint32_t batch, int stage, uint32_t* x,
                                           uint32_t* y, uint32_t* z) {
    switch (stage) {
        case 0: *x = t.N / kEncCoefs; *y = 2; *z = batch; break;
        case 1: *x = t.P * (t.N2 / cols_per_wg(t.n1)); *y = 2; *z = batch; break;
        case 2: *x = t.N1 / kMidBlocks; *y = batch; *z = t.P; break;
        case 3: *x = t.P * (t.N2 / cols_per_wg(t.n1)); *y = 1; *z = batch; break;
        case 4: *x = t.N / kCrtCoefs; *y = 1; *z = batch; break;
        default: *x = t.N1; *y = 2; *z = batch; break;  // encode + forward columns: 256 / N1 columns per tile
    }
}

// The dispatcher deals workgroups to the 8 XCDs round-robin (id % 8), and each
// XCD has its own L2.  Renumbering so that XCD k runs the k-th contiguous
// eighth of the tiles keeps neighbouring tiles -- which share cache lines of
// their operands and the prime's twiddles -- in one L2 at about the same time.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t id, uint32_t total, uint32_t kflags) {
    if (kflags & kFlagNoXcd) return id;
    const uint32_t per = total / 8, full = per * 8;
    return id < full ? (id % 8) * per + id / 8 : id;
}

// (x, y, z) of tile `tile` in a stage grid of width gx and height gy.
__device__ __forceinline__ void tile_xyz(uint32_t tile, uint32_t gx, uint32_t gy, uint32_t* x, uint32_t* y,
                                         uint32_t* z) {
    *x = tile % gx;
    *y = (tile / gx) % gy;
    *z = tile / (gx * gy);
}

// This workgroup's (x, y, z) in the single-plan grid of `stage` (1-D launch).
__device__ __forceinline__ void single_xyz(const DevTables& t, int stage, uint32_t* x, uint32_t* y, uint32_t* z) {
    uint32_t gx, gy, gz;
    stage_grid(t, 1, stage, &gx, &gy, &gz);
    stage_grid(t, gridDim.x / (gx * gy * gz), stage, &gx, &gy, &gz);  // the launch covers whole products
    tile_xyz(xcd_tile(blockIdx.x, gridDim.x, t.kflags), gx, gy, x, y, z);
    *x = __builtin_amdgcn_readfirstlane(*x);  // uniform: keep them in scalar registers
    *y = __builtin_amdgcn_readfirstlane(*y);
    *z = __builtin_amdgcn_readfirstlane(*z);
}

// ---------------------------------------------------------------------------
// Entry kernels, all on 1-D grids of XCD-renumbered tiles.  Single-plan
// kernels take the plan by value; group kernels take a table of (plan, batch)
// groups whose tiles are numbered group after group, so one launch per stage
// covers products of different geometries and moduli.
// ---------------------------------------------------------------------------
// C MFMA chains per wave (kFlagEncOneChain: C = 1, four workgroups per CU).
template <int C>
__global__ __launch_bounds__(kThreads, C == 1 ? 4 : 3) void k_encode(DevTables t, const uint32_t* __restrict__ a,
                                                                  const uint32_t* __restrict__ b,
                                                                  uint32_t* __restrict__ planes) {
    extern __shared__ __align__(16) uint8_t smem_enc[];
    uint32_t x, y, z;
    single_xyz(t, 0, &x, &y, &z);
    encode_dispatch<C>(t, a, b, planes, x, y, z, smem_enc);
}

template <int KS, int NLOG1>
__global__ __launch_bounds__(kThreads, 3) void k_encode_fwd(DevTables t, const uint32_t* __restrict__ a,
                                                         const uint32_t* __restrict__ b,
                                                         uint32_t* __restrict__ planes) {
    extern __shared__ __align__(16) uint8_t smem_ef[];
    uint32_t r, op, prod;
    tile_xyz(xcd_tile(blockIdx.x, gridDim.x, t.kflags), t.N1, 2, &r, &op, &prod);
    // uniform over the workgroup: scalar registers, not VGPRs held all kernel
    r = __builtin_amdgcn_readfirstlane(r);
    op = __builtin_amdgcn_readfirstlane(op);
    prod = __builtin_amdgcn_readfirstlane(prod);
    encfwd_wg<KS, NLOG1>(t, a, b, planes, r, op, prod, smem_ef);
}

template <int NLOG1, bool FAST>
__global__ __launch_bounds__(kThreads) kNttOccupancy void k_fwd_cols(DevTables t, uint32_t* __restrict__ planes) {
    extern __shared__ __align__(16) uint8_t smem_fc[];
    uint32_t x, y, z;
    single_xyz(t, 1, &x, &y, &z);
    fwd_cols_wg<NLOG1, FAST>(t, planes, x, y, z, reinterpret_cast<double*>(smem_fc));
}

template <int NLOG1, bool FAST>
__global__ __launch_bounds__(kThreads) kNttOccupancy void k_inv_cols(DevTables t, uint32_t* __restrict__ planes) {
    extern __shared__ __align__(16) uint8_t smem_ic[];
    uint32_t x, y, z;
    single_xyz(t, 3, &x, &y, &z);
    inv_cols_wg<NLOG1, FAST>(t, planes, x, z, reinterpret_cast<double*>(smem_ic));
}

template <int MODE>
__global__ __launch_bounds__(kThreads) kNttOccupancy void k_middle(DevTables t, uint32_t* __restrict__ planes) {
    extern __shared__ __align__(16) uint8_t smem_mid[];
    uint32_t x, y, z;
    single_xyz(t, 2, &x, &y, &z);
    middle_wg<MODE>(t, planes, x, y, z, reinterpret_cast<double*>(smem_mid));
}

// One kernel per k-step count, chosen on the host: each gets its own register
// allocation (one kernel switching over all seventeen spilled 88 VGPRs; at
// Kcrt = 480 this one spills none).  WIDE: the five-per-CU build.
template <int KS, bool WIDE>
__global__ __launch_bounds__(kCrtThreads, WIDE ? kCrtWgPerCuWide : kCrtWgPerCu) void k_crt(
    DevTables t, const uint32_t* __restrict__ planes, uint32_t* __restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem_crt[];
    uint32_t x, y, z;
    single_xyz(t, 4, &x, &y, &z);
    crt_tiles<KS, WIDE ? kCrtAheadWide : kCrtAhead>(t, planes, out, x, z, smem_crt);
}

// Finds the group of workgroup slot `id` of a mixed launch and the slot's
// (x, y, z) in that group's grid, or returns false for the padding slots.
// Each group's slots start at a multiple of 8 (group_workgroups pads the
// count), so the XCD renumbering is applied inside every group: each XCD gets
// an eighth of every group's tiles, whatever their cost.
// (The table is only ever indexed, never addressed, so it stays in the kernel
// argument segment.)
__device__ __forceinline__ bool locate(const GroupTable& g, int stage, uint32_t id, uint32_t* k_out, uint32_t* x,
                                       uint32_t* y, uint32_t* z) {
    uint32_t k = 0;
    while (k + 1 < g.n && id >= g.g[k + 1].begin[stage]) ++k;
    uint32_t gx, gy, gz;
    stage_grid(g.g[k].t, g.g[k].args.batch, stage, &gx, &gy, &gz);
    const uint32_t count = gx * gy * gz;
    const uint32_t tile = xcd_tile(id - g.g[k].begin[stage], (count + 7) / 8 * 8, g.g[k].t.kflags);
    *k_out = __builtin_amdgcn_readfirstlane(k);  // uniform: scalar registers
    tile_xyz(tile, gx, gy, x, y, z);
    *x = __builtin_amdgcn_readfirstlane(*x);
    *y = __builtin_amdgcn_readfirstlane(*y);
    *z = __builtin_amdgcn_readfirstlane(*z);
    return tile < count;
}

// Runs body(d, x, y, z) for this workgroup's tile(s) of stage `stage`, d
// being a copy of the tile's group descriptor (a copy: its fields are not
// loaded again after every barrier): the slot blockIdx.x, or, with a work
// queue, slots pulled one at a time from the counter of this workgroup's XCD
// (workgroup b runs on XCD b mod 8; slot ids 8 n + x go to XCD x, which keeps
// the renumbered tiles of each XCD on it) until the stage's slots run out.
// Every counter only grows and every workgroup stops at the first slot past
// the end, so the grid always drains.  (The body captures nothing: a lambda
// holding a reference to the by-value table would copy it to scratch.)  The
// two modes are separate kernels (PQ): one function holding both copies of a
// body gets the registers of the larger.
template <bool PQ, class Body>
__device__ __forceinline__ void for_each_tile(const GroupTable& g, int stage, Body body) {
    uint32_t k, x, y, z;
    if constexpr (!PQ) {
        if (locate(g, stage, blockIdx.x, &k, &x, &y, &z)) {
            const GroupDesc d = g.g[k];
            body(d, x, y, z);
        }
    } else {
        __shared__ uint32_t next_slot;
        const uint32_t xcd = blockIdx.x & 7;
        uint32_t* ctr = g.queue + 8 * stage + xcd;
        for (;;) {
            if (threadIdx.x == 0) next_slot = atomicAdd(ctr, 1u);
            __syncthreads();
            const uint32_t id = __builtin_amdgcn_readfirstlane(next_slot) * 8 + xcd;
            if (id >= g.total[stage]) return;  // uniform: every wave read the same slot
            if (locate(g, stage, id, &k, &x, &y, &z)) {
                const GroupDesc d = g.g[k];
                body(d, x, y, z);
            }
            __syncthreads();  // the slot word and the body's LDS are free for the next tile
        }
    }
}

__global__ void k_queue_reset(uint32_t* __restrict__ queue) {
    if (threadIdx.x < kQueueWords) queue[threadIdx.x] = 0;
}

// KS / NLOG1 = 0: any mix of shapes (switch per workgroup); otherwise every
// group of the launch has that shape and the code is specialised for it.
template <int KS, bool PQ>
__global__ __launch_bounds__(kThreads, 3) void k_encode_groups(const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_enc_g[];
    for_each_tile<PQ>(g, 0, [](const GroupDesc& d, uint32_t x, uint32_t y, uint32_t z) {
        if constexpr (KS == 0)
            encode_dispatch(d.t, d.args.a, d.args.b, d.args.planes, x, y, z, smem_enc_g);
        else
            encode_wg<KS>(d.t, d.args.a, d.args.b, d.args.planes, x, y, z, smem_enc_g);
    });
}

template <int KS, int NLOG1, bool PQ>
__global__ __launch_bounds__(kThreads, 3) void k_encode_fwd_groups(const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_ef_g[];
    for_each_tile<PQ>(g, 5, [](const GroupDesc& d, uint32_t r, uint32_t op, uint32_t prod) {
        encfwd_wg<KS, NLOG1>(d.t, d.args.a, d.args.b, d.args.planes, r, op, prod, smem_ef_g);
    });
}

template <int NLOG1, bool PQ>
__global__ __launch_bounds__(kThreads) kNttOccupancy void k_fwd_cols_groups(const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_fc_g[];
    for_each_tile<PQ>(g, 1, [](const GroupDesc& d, uint32_t x, uint32_t y, uint32_t z) {
        double* s = reinterpret_cast<double*>(smem_fc_g);
        if constexpr (NLOG1 != 0) {
            fwd_cols_wg<NLOG1>(d.t, d.args.planes, x, y, z, s);
        } else {
            switch (d.t.n1) {
                case 4: fwd_cols_wg<4>(d.t, d.args.planes, x, y, z, s); break;
                case 5: fwd_cols_wg<5>(d.t, d.args.planes, x, y, z, s); break;
                case 6: fwd_cols_wg<6>(d.t, d.args.planes, x, y, z, s); break;
                case 7: fwd_cols_wg<7>(d.t, d.args.planes, x, y, z, s); break;
                default: fwd_cols_wg<8>(d.t, d.args.planes, x, y, z, s); break;
            }
        }
    });
}

// (mixes use the five-operation twiddle products; MODE 2 when every group's
// plan allows them in the last four inverse levels too)
template <bool PQ, int MODE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(PQ ? 3 : 4, 4))) void k_middle_groups(
    const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_mid_g[];
    for_each_tile<PQ>(g, 2, [](const GroupDesc& d, uint32_t x, uint32_t y, uint32_t z) {
        middle_wg<MODE>(d.t, d.args.planes, x, y, z, reinterpret_cast<double*>(smem_mid_g));
    });
}

template <int NLOG1, bool PQ>
__global__ __launch_bounds__(kThreads) kNttOccupancy void k_inv_cols_groups(const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_ic_g[];
    for_each_tile<PQ>(g, 3, [](const GroupDesc& d, uint32_t x, uint32_t, uint32_t z) {
        double* s = reinterpret_cast<double*>(smem_ic_g);
        if constexpr (NLOG1 != 0) {
            inv_cols_wg<NLOG1>(d.t, d.args.planes, x, z, s);
        } else {
            switch (d.t.n1) {
                case 4: inv_cols_wg<4>(d.t, d.args.planes, x, z, s); break;
                case 5: inv_cols_wg<5>(d.t, d.args.planes, x, z, s); break;
                case 6: inv_cols_wg<6>(d.t, d.args.planes, x, z, s); break;
                case 7: inv_cols_wg<7>(d.t, d.args.planes, x, z, s); break;
                default: inv_cols_wg<8>(d.t, d.args.planes, x, z, s); break;
            }
        }
    });
}

template <int KS, bool PQ>
__global__ __launch_bounds__(kCrtThreads, kCrtWgPerCu) void k_crt_groups(const GroupTable g) {
    extern __shared__ __align__(16) uint8_t smem_crt_g[];
    for_each_tile<PQ>(g, 4, [](const GroupDesc& d, uint32_t x, uint32_t, uint32_t z) {
        if constexpr (KS == 0)
            crt_dispatch(d.t, d.args.planes, d.args.out, x, z, smem_crt_g);
        else
            crt_tiles<KS>(d.t, d.args.planes, d.args.out, x, z, smem_crt_g);
    });
}

bool shapes_ok(const DevTables& t, const BatchArgs& args) {
    return args.batch >= 1 && args.batch <= 65535 && t.N2 == 256 && t.N1 >= 16 && t.N1 <= 256 &&
           t.N == t.N1 * t.N2 && t.N % kEncCoefs == 0 && t.Kenc % 32 == 0 && t.Kenc <= 256 &&
           t.Kcrt % 32 == 0 && t.Mcrt % 32 == 0 && t.P >= 1 && t.P <= kMaxPrimesCrt && t.L >= 3 &&
           t.L <= 64 && t.Kcrt >= 4 * (t.P + 1) && t.Mcrt >= 4 * t.L + 1 && t.Kenc >= 4 * t.L &&
           t.Kcrt <= 32 * 17 && t.Mcrt <= 32 * 4 * crt_tiles_per_wave(t.Kcrt / 32, t.Mcrt / 32) &&
           t.qshift + 32 == t.W && t.enc_ts_log <= t.n1 && t.enc_ts_log <= 7 && (t.kflags & ~kPlanFlags) <= kFlagsAll;
}

// Twiddle-product form of the single-plan NTT kernels (see middle_wg): the
// five-operation form unless kFlagMulModTw6 asks for mul_mod throughout; in
// the middle kernel's last four inverse levels only where the plan allows it.
int tw_mode(const DevTables& t) {
    return (t.kflags & kFlagMulModTw6) ? 0 : (t.kflags & kPlanTw5Late) ? 2 : 1;
}

template <int NLOG1>
hipError_t launch_cols(const DevTables& t, const BatchArgs& args, hipStream_t stream, bool forward) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T;
    const dim3 grid(t.P * (t.N2 / CPW) * (forward ? 2 : 1) * args.batch);
    const bool fast = tw_mode(t) > 0;
    if (forward && fast)
        hipLaunchKernelGGL((k_fwd_cols<NLOG1, true>), grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes);
    else if (forward)
        hipLaunchKernelGGL((k_fwd_cols<NLOG1, false>), grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes);
    else if (fast)
        hipLaunchKernelGGL((k_inv_cols<NLOG1, true>), grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes);
    else
        hipLaunchKernelGGL((k_inv_cols<NLOG1, false>), grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes);
    return hipGetLastError();
}

hipError_t dispatch_cols(const DevTables& t, const BatchArgs& args, hipStream_t stream, bool forward) {
    switch (t.n1) {
        case 4: return launch_cols<4>(t, args, stream, forward);
        case 5: return launch_cols<5>(t, args, stream, forward);
        case 6: return launch_cols<6>(t, args, stream, forward);
        case 7: return launch_cols<7>(t, args, stream, forward);
        case 8: return launch_cols<8>(t, args, stream, forward);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_encode(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    const dim3 grid(t.N / kEncCoefs * 2 * args.batch);
    if (t.kflags & kFlagEncOneChain)
        hipLaunchKernelGGL(k_encode<1>, grid, dim3(kThreads), enc_smem_bytes(t), stream, t, args.a, args.b,
                           args.planes);
    else
        hipLaunchKernelGGL(k_encode<2>, grid, dim3(kThreads), enc_smem_bytes(t), stream, t, args.a, args.b,
                           args.planes);
    return hipGetLastError();
}

template <int NLOG1>
void launch_encfwd_n(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    const dim3 g(t.N1 * 2 * args.batch), blk(kThreads);
    const size_t sm = ef_smem_bytes(t);
    switch (t.Kenc / 32) {
        case 1: hipLaunchKernelGGL((k_encode_fwd<1, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 2: hipLaunchKernelGGL((k_encode_fwd<2, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 3: hipLaunchKernelGGL((k_encode_fwd<3, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 4: hipLaunchKernelGGL((k_encode_fwd<4, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 5: hipLaunchKernelGGL((k_encode_fwd<5, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 6: hipLaunchKernelGGL((k_encode_fwd<6, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        case 7: hipLaunchKernelGGL((k_encode_fwd<7, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
        default: hipLaunchKernelGGL((k_encode_fwd<8, NLOG1>), g, blk, sm, stream, t, args.a, args.b, args.planes); break;
    }
}

hipError_t launch_encode_forward(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    // The fused kernel has N1 * 2 * batch workgroups; below about one per CU
    // (N = 16,384 at B = 1) the separate kernels, with more and smaller
    // workgroups, are faster (default threshold kFusedMinTiles; tunable).
    switch (t.N1 * 2 * args.batch >= t.fuse_min_tiles ? t.n1 : 0) {
        case 6: launch_encfwd_n<6>(t, args, stream); break;
        case 7: launch_encfwd_n<7>(t, args, stream); break;
        case 8: launch_encfwd_n<8>(t, args, stream); break;
        default: {
            const hipError_t e = launch_encode(t, args, stream);
            return e != hipSuccess ? e : launch_forward_columns(t, args, stream);
        }
    }
    return hipGetLastError();
}

hipError_t launch_forward_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    return dispatch_cols(t, args, stream, true);
}

hipError_t launch_middle(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    const dim3 grid(t.N1 / kMidBlocks * args.batch * t.P);
    switch (tw_mode(t)) {
        case 0: hipLaunchKernelGGL(k_middle<0>, grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes); break;
        case 1: hipLaunchKernelGGL(k_middle<1>, grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes); break;
        default: hipLaunchKernelGGL(k_middle<2>, grid, dim3(kThreads), kExchangeBytes, stream, t, args.planes); break;
    }
    return hipGetLastError();
}

hipError_t launch_inverse_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    return dispatch_cols(t, args, stream, false);
}

template <bool WIDE>
void launch_crt_ks(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    const dim3 g(t.N / kCrtCoefs * args.batch), blk(kCrtThreads);
    const size_t sm = crt_smem_bytes(t);
    switch (t.Kcrt / 32) {
        case 1: hipLaunchKernelGGL((k_crt<1, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 2: hipLaunchKernelGGL((k_crt<2, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 3: hipLaunchKernelGGL((k_crt<3, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 4: hipLaunchKernelGGL((k_crt<4, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 5: hipLaunchKernelGGL((k_crt<5, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 6: hipLaunchKernelGGL((k_crt<6, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 7: hipLaunchKernelGGL((k_crt<7, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 8: hipLaunchKernelGGL((k_crt<8, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 9: hipLaunchKernelGGL((k_crt<9, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 10: hipLaunchKernelGGL((k_crt<10, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 11: hipLaunchKernelGGL((k_crt<11, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 12: hipLaunchKernelGGL((k_crt<12, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 13: hipLaunchKernelGGL((k_crt<13, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 14: hipLaunchKernelGGL((k_crt<14, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 15: hipLaunchKernelGGL((k_crt<15, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        case 16: hipLaunchKernelGGL((k_crt<16, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
        default: hipLaunchKernelGGL((k_crt<17, WIDE>), g, blk, sm, stream, t, args.planes, args.out); break;
    }
}

hipError_t launch_crt(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    if (t.kflags & kFlagCrtWide)
        launch_crt_ks<true>(t, args, stream);
    else
        launch_crt_ks<false>(t, args, stream);
    return hipGetLastError();
}

hipError_t launch_product(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    hipError_t e;
    if ((e = launch_encode_forward(t, args, stream)) != hipSuccess) return e;
    if ((e = launch_middle(t, args, stream)) != hipSuccess) return e;
    if ((e = launch_inverse_columns(t, args, stream)) != hipSuccess) return e;
    return launch_crt(t, args, stream);
}

hipError_t launch_mfma_probe(const int8_t* a, const int8_t* b, int32_t* c, hipStream_t stream) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, stream, a, b, c);
    return hipGetLastError();
}

uint32_t group_workgroups(const DevTables& t, uint32_t batch, int stage) {
    uint32_t x, y, z;
    stage_grid(t, batch, stage, &x, &y, &z);
    return (x * y * z + 7) / 8 * 8;  // padded to whole rounds of the 8 XCDs
}

// Workgroups of `kern` (`threads` each, `sm` bytes of dynamic LDS) the
// current device holds at once, rounded down to whole rounds of the 8 XCDs.
uint32_t resident_grid(const void* kern, uint32_t threads, size_t sm) {
    static std::mutex m;
    static std::map<std::tuple<int, const void*, uint32_t, size_t>, uint32_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 8;
    const auto key = std::make_tuple(dev, kern, threads, sm);
    std::lock_guard<std::mutex> g(m);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, int(threads), sm) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        per_cu = cus = 1;
    const uint32_t n = std::max<uint32_t>(8, uint32_t(std::max(per_cu, 1) * std::max(cus, 1)) / 8 * 8);
    cache.emplace(key, n);
    return n;
}

// Launches group kernel `kern` over `tiles` workgroup slots: one workgroup per
// slot, or, when the table has a work queue, a resident grid whose
// workgroups pull the slots.
void launch_g(const void* kern, uint32_t tiles, uint32_t threads, size_t sm, GroupTable& table,
              hipStream_t stream) {
    const uint32_t grid = table.queue ? std::min(tiles, resident_grid(kern, threads, sm)) : tiles;
    void* args[] = {&table};
    (void)hipLaunchKernel(kern, dim3(grid), dim3(threads), args, sm, stream);
}

template <int NLOG1, bool PQ>
void launch_encfwd_groups(uint32_t ks, uint32_t tiles, size_t sm, GroupTable& table, hipStream_t stream) {
    const void* k = nullptr;
    switch (ks) {
        case 1: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<1, NLOG1, PQ>); break;
        case 2: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<2, NLOG1, PQ>); break;
        case 3: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<3, NLOG1, PQ>); break;
        case 4: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<4, NLOG1, PQ>); break;
        case 5: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<5, NLOG1, PQ>); break;
        case 6: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<6, NLOG1, PQ>); break;
        case 7: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<7, NLOG1, PQ>); break;
        default: k = reinterpret_cast<const void*>(&k_encode_fwd_groups<8, NLOG1, PQ>); break;
    }
    launch_g(k, tiles, kThreads, sm, table, stream);
}

// The stage launches of launch_product_groups, with the one-tile-per-workgroup
// (PQ = false) or the work-queue (PQ = true) kernels.
template <bool PQ>
void enqueue_groups(const GroupDesc* groups, uint32_t ngroups, GroupTable& table,
                    const uint32_t (&total)[kGroupStages], size_t enc, size_t crt, size_t ef, hipStream_t stream) {
    auto go = [&](const void* kern, uint32_t tiles, uint32_t threads, size_t sm) {
        launch_g(kern, tiles, threads, sm, table, stream);
    };
    // Specialised kernels when every group shares the shape they depend on.
    bool same_k = true, same_n1 = true, same_kc = true;
    for (uint32_t k = 1; k < ngroups; ++k) {
        same_k = same_k && groups[k].t.Kenc == groups[0].t.Kenc;
        same_n1 = same_n1 && groups[k].t.n1 == groups[0].t.n1;
        same_kc = same_kc && groups[k].t.Kcrt == groups[0].t.Kcrt;
    }
    const uint32_t ks = same_k ? groups[0].t.Kenc / 32 : 0, n1 = same_n1 ? groups[0].t.n1 : 0;
    const size_t xb = kExchangeBytes;
    if (ks != 0 && n1 >= 6 && total[5] >= groups[0].t.fuse_min_tiles) {
        // every group has the same Kenc and N1 >= 64: encode and forward columns in one kernel
        switch (n1) {
            case 6: launch_encfwd_groups<6, PQ>(ks, total[5], ef, table, stream); break;
            case 7: launch_encfwd_groups<7, PQ>(ks, total[5], ef, table, stream); break;
            default: launch_encfwd_groups<8, PQ>(ks, total[5], ef, table, stream); break;
        }
    } else {
        switch (ks) {
            case 1: go(reinterpret_cast<const void*>(&k_encode_groups<1, PQ>), total[0], kThreads, enc); break;
            case 2: go(reinterpret_cast<const void*>(&k_encode_groups<2, PQ>), total[0], kThreads, enc); break;
            case 3: go(reinterpret_cast<const void*>(&k_encode_groups<3, PQ>), total[0], kThreads, enc); break;
            case 4: go(reinterpret_cast<const void*>(&k_encode_groups<4, PQ>), total[0], kThreads, enc); break;
            case 5: go(reinterpret_cast<const void*>(&k_encode_groups<5, PQ>), total[0], kThreads, enc); break;
            case 6: go(reinterpret_cast<const void*>(&k_encode_groups<6, PQ>), total[0], kThreads, enc); break;
            case 7: go(reinterpret_cast<const void*>(&k_encode_groups<7, PQ>), total[0], kThreads, enc); break;
            case 8: go(reinterpret_cast<const void*>(&k_encode_groups<8, PQ>), total[0], kThreads, enc); break;
            default: go(reinterpret_cast<const void*>(&k_encode_groups<0, PQ>), total[0], kThreads, enc); break;
        }
        switch (n1) {
            case 4: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<4, PQ>), total[1], kThreads, xb); break;
            case 5: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<5, PQ>), total[1], kThreads, xb); break;
            case 6: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<6, PQ>), total[1], kThreads, xb); break;
            case 7: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<7, PQ>), total[1], kThreads, xb); break;
            case 8: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<8, PQ>), total[1], kThreads, xb); break;
            default: go(reinterpret_cast<const void*>(&k_fwd_cols_groups<0, PQ>), total[1], kThreads, xb); break;
        }
    }
    bool late = true;
    for (uint32_t k = 0; k < ngroups; ++k) late = late && (groups[k].t.kflags & kPlanTw5Late);
    if (late)
        go(reinterpret_cast<const void*>(&k_middle_groups<PQ, 2>), total[2], kThreads, xb);
    else
        go(reinterpret_cast<const void*>(&k_middle_groups<PQ, 1>), total[2], kThreads, xb);
    switch (n1) {
        case 4: go(reinterpret_cast<const void*>(&k_inv_cols_groups<4, PQ>), total[3], kThreads, xb); break;
        case 5: go(reinterpret_cast<const void*>(&k_inv_cols_groups<5, PQ>), total[3], kThreads, xb); break;
        case 6: go(reinterpret_cast<const void*>(&k_inv_cols_groups<6, PQ>), total[3], kThreads, xb); break;
        case 7: go(reinterpret_cast<const void*>(&k_inv_cols_groups<7, PQ>), total[3], kThreads, xb); break;
        case 8: go(reinterpret_cast<const void*>(&k_inv_cols_groups<8, PQ>), total[3], kThreads, xb); break;
        default: go(reinterpret_cast<const void*>(&k_inv_cols_groups<0, PQ>), total[3], kThreads, xb); break;
    }
    switch (same_kc ? groups[0].t.Kcrt / 32 : 0) {
        case 1: go(reinterpret_cast<const void*>(&k_crt_groups<1, PQ>), total[4], kCrtThreads, crt); break;
        case 2: go(reinterpret_cast<const void*>(&k_crt_groups<2, PQ>), total[4], kCrtThreads, crt); break;
        case 3: go(reinterpret_cast<const void*>(&k_crt_groups<3, PQ>), total[4], kCrtThreads, crt); break;
        case 4: go(reinterpret_cast<const void*>(&k_crt_groups<4, PQ>), total[4], kCrtThreads, crt); break;
        case 5: go(reinterpret_cast<const void*>(&k_crt_groups<5, PQ>), total[4], kCrtThreads, crt); break;
        case 6: go(reinterpret_cast<const void*>(&k_crt_groups<6, PQ>), total[4], kCrtThreads, crt); break;
        case 7: go(reinterpret_cast<const void*>(&k_crt_groups<7, PQ>), total[4], kCrtThreads, crt); break;
        case 8: go(reinterpret_cast<const void*>(&k_crt_groups<8, PQ>), total[4], kCrtThreads, crt); break;
        case 9: go(reinterpret_cast<const void*>(&k_crt_groups<9, PQ>), total[4], kCrtThreads, crt); break;
        case 10: go(reinterpret_cast<const void*>(&k_crt_groups<10, PQ>), total[4], kCrtThreads, crt); break;
        case 11: go(reinterpret_cast<const void*>(&k_crt_groups<11, PQ>), total[4], kCrtThreads, crt); break;
        case 12: go(reinterpret_cast<const void*>(&k_crt_groups<12, PQ>), total[4], kCrtThreads, crt); break;
        case 13: go(reinterpret_cast<const void*>(&k_crt_groups<13, PQ>), total[4], kCrtThreads, crt); break;
        case 14: go(reinterpret_cast<const void*>(&k_crt_groups<14, PQ>), total[4], kCrtThreads, crt); break;
        case 15: go(reinterpret_cast<const void*>(&k_crt_groups<15, PQ>), total[4], kCrtThreads, crt); break;
        case 16: go(reinterpret_cast<const void*>(&k_crt_groups<16, PQ>), total[4], kCrtThreads, crt); break;
        case 17: go(reinterpret_cast<const void*>(&k_crt_groups<17, PQ>), total[4], kCrtThreads, crt); break;
        default: go(reinterpret_cast<const void*>(&k_crt_groups<0, PQ>), total[4], kCrtThreads, crt); break;
    }
}

hipError_t launch_product_groups(const GroupDesc* groups, uint32_t ngroups, hipStream_t stream, uint32_t* queue) {
    if (ngroups == 0 || ngroups > kMaxGroups) return hipErrorInvalidValue;
    GroupTable table{};
    table.n = ngroups;
    table.queue = queue;
    size_t enc = 0, crt = 0, ef = 0;
    uint32_t total[kGroupStages] = {};
    for (uint32_t k = 0; k < ngroups; ++k) {
        const GroupDesc& g = groups[k];
        if (!shapes_ok(g.t, g.args)) return hipErrorInvalidValue;
        enc = std::max(enc, enc_smem_bytes(g.t));
        crt = std::max(crt, crt_smem_bytes(g.t));
        ef = std::max(ef, ef_smem_bytes(g.t));
        for (int st = 0; st < kGroupStages; ++st) {
            if (g.begin[st] != total[st]) return hipErrorInvalidValue;
            total[st] += group_workgroups(g.t, g.args.batch, st);
        }
        table.g[k] = g;
    }
    for (int st = 0; st < kGroupStages; ++st) table.total[st] = total[st];
    if (queue) {
        hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, stream, queue);  // a kernel, not a memset node
        enqueue_groups<true>(groups, ngroups, table, total, enc, crt, ef, stream);
    } else {
        enqueue_groups<false>(groups, ngroups, table, total, enc, crt, ef, stream);
    }
    return hipGetLastError();
}

hipError_t configure_kernels() {
    // Every kernel's dynamic LDS stays within the 64 KiB default (the largest,
    // the fused encode at Kenc = 224 with its staged weight tiles, takes 50 KB
    // besides 2.3 KB of static LDS; the CRT at most 35.8 KB), so no attribute
    // needs raising.
    static_assert(size_t(kEncCoefs) * (256 + 16) + sizeof(EncPrime) * kMaxPrimesCrt <= 65536, "encode LDS");
    static_assert(size_t(16) * kEfS * sizeof(double) + sizeof(double) * kMaxPrimesCrt + 64 * (224 + 16) <= 65536,
                  "fused encode LDS");
    static_assert(size_t(kCrtCoefs) * (32 * 17 + 16) <= 65536 && size_t(65) * kCrtCoefs * 8 <= 65536, "CRT LDS");
    return hipSuccess;
}

}  // namespace bcrt
