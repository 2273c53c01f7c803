// Host-side RNS plan and device table layouts for the MI355X ByteCRT engine.
//
// The plan follows the reference's init-time setup (native-smoke/rns-plan.hpp:174-324,
// reference.py:39-54): greedy descending NTT primes p == 1 mod 2N in (2^30, 2^31)
// until bit_length(prod p) >= 2W + bit_length(N) + 2, the smallest-base primitive
// 2N-th root psi, CRT inverses (P/p_j)^-1 mod p_j, weights P_j mod q and the alpha
// weight (-P) mod q.  It is generalised to any q with bit_length(q) == W (the
// reference only accepts q = 2^W - gap) and to any N = 2^k, 4096 <= N <= 65536.
//
// On top of that it lays out what the gfx950 kernels read:
//  * psi-power twiddles in bit-reversed order with Shoup companions (forward and
//    inverse merged negacyclic NTT, no twist pass);
//  * signed int8 weight tiles for the encode and CRT MFMA GEMMs (balanced base-256
//    digits, so the signed MFMA is exact; the unsigned data bytes are biased by
//    -128 and a per-column constant restores the sum);
//  * the generic-q quotient estimate constants for the CRT epilogue.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bcrt {

constexpr uint32_t kMaxPrimes = 128;
constexpr uint32_t kMaxLimbs = 64;
constexpr uint32_t kBlockLen = 256;  // N2: contiguous block length of the second NTT pass

struct Geometry {
    uint32_t N = 0, W = 0, L = 0, P = 0;
    uint32_t log_n = 0;
    uint32_t N1 = 0, n1 = 0;  // column length / log2 (first NTT pass)
    uint32_t N2 = 0;          // block length (second NTT pass)
    uint32_t Kenc = 0;        // encode inner axis: 4L padded to 32
    uint32_t PTenc = 0;       // encode prime tiles (8 primes each)
    uint32_t Kcrt = 0;        // CRT inner axis: 4(P+1) padded to 32
    uint32_t Mcrt = 0;        // CRT output digit rows: 4L+1 padded to 32
    uint32_t product_bits = 0;
};

struct HostPlan {
    Geometry g;
    std::vector<uint32_t> q;            // L limbs
    std::vector<uint32_t> primes;       // P
    std::vector<uint32_t> psi;          // P: primitive 2N-th roots
    std::vector<uint32_t> crt_inverse;  // P: (P/p_j)^-1 mod p_j
    std::vector<uint32_t> weights;      // (P+1) x L: P_j mod q, last row (-P) mod q
    std::vector<double> pd;             // P x 2: (p, 1/p) for the FP64 modular products
    std::vector<double> scale;          // P: (N/2)^-1 * crt_inverse mod p
    std::vector<double> scale_z1;       // P: scale * psi^-brv(1) mod p (last inverse level, scale folded in)
    std::vector<double> zf16;           // P x 16: psi^brv(k) mod p for k < 16 (first forward column round)
    std::vector<uint32_t> zf, zi;       // P x N: psi^{+-brv(k)} mod p
    std::vector<int8_t> wenc;           // PTenc x 32 x Kenc
    std::vector<uint64_t> enc_off;      // P
    std::vector<double> enc_inv;        // P: 1/p rounded down (FP64 quotient estimate)
    std::vector<uint8_t> enc_prime;     // P x 24 bytes: {double off, double inv, u32 p, u32 off mod 2^32}
    std::vector<int8_t> wcrt;           // Mcrt x Kcrt
    std::vector<int64_t> corr;          // Mcrt/4 limb corrections
    std::vector<double> invp;           // P
    uint32_t qshift = 0;                // W - 32
    uint64_t qt1 = 0;                   // (q >> qshift) + 1
    double inv_qt1 = 0;                 // 1 / qt1
    uint32_t gap = 0;                   // 2^W - q when below 2^32 and L >= 3 (near-power CRT reduction), else 0
    bool tw5_late = false;              // every prime: |q| k < 2^53 for |q| < 192 p + 3 p / 32 + 1 (p - 1 = k 2^s)
};

// Throws std::invalid_argument (bad geometry or q) or std::runtime_error (no plan).
HostPlan make_plan(uint32_t N, uint32_t W, uint32_t L, const uint32_t* q);

// Number of primes the plan would use for (N, W) (no q needed).
uint32_t plan_prime_count(uint32_t N, uint32_t W);

// gap when q (L limbs, bit length W, L >= 3) is 2^W - gap with gap < 2^32 --
// the form the CRT's near-power reduction takes, and the only one the
// reference accepts (rns-plan.hpp:186-197); otherwise 0.
uint32_t near_power_gap(uint32_t W, uint32_t L, const uint32_t* q);

}  // namespace bcrt
