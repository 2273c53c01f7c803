// gfx950 kernels of the ByteCRT exact negacyclic product.
//
// Pipeline per product (all integer, bit-exact):
//   k_encode      wide coefficients -> residues mod each prime (int8 MFMA byte-power GEMM)
//   k_fwd_cols    forward NTT levels 0..n1-1 (columns of N1 elements, stride N2)
//   k_middle      forward levels n1..n-1 of A and B, pointwise product, inverse levels n-1..n1
//   k_inv_cols    inverse levels n1-1..0, scale by N^-1 * (P/p)^-1 (CRT terms)
//   k_crt         centered alpha + int8 MFMA CRT GEMM + carry + generic mod-q reduction
//
// Residue planes use the "column-major" position f = r * N1 + s for natural
// coefficient index r + N2 * s (r < N2 = 256, s < N1 = N / 256): the column
// transforms read/write N1 contiguous words and the block transforms read a
// tile of 16 consecutive s for each r.
//
// The NTT is the merged negacyclic Cooley-Tukey / Gentleman-Sande pair with
// psi^{+-bitrev(k)} twiddles (no separate twist pass, bit-reversed spectrum
// consumed in place), Shoup multiplication (v_mul_hi_u32) for twiddles and a
// Montgomery product for the pointwise step; all values stay canonical in [0, p).
#include "bytecrt_kernels.hpp"

namespace bcrt {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint32_t kThreads = 256;
constexpr uint32_t kEncCoefs = 128;  // coefficients per encode workgroup
constexpr uint32_t kCrtCoefs = 64;   // coefficients per CRT workgroup
constexpr uint32_t kMidBlocks = 16;  // 256-element blocks per middle workgroup
constexpr uint32_t kMidStride = 257;

// ---------------------------------------------------------------------------
// Modular arithmetic, p in (2^30, 2^31)
// ---------------------------------------------------------------------------
// x * w mod p for any x < 2^32, w < p, ws = floor(w 2^32 / p); the raw
// remainder is in [0, 2p) and one min() folds it (Shoup 2008 / Harvey 2014).
__device__ __forceinline__ uint32_t mul_shoup(uint32_t x, uint2 w, uint32_t p) {
    const uint32_t qh = __umulhi(x, w.y);
    const uint32_t r = x * w.x - qh * p;
    return min(r, r - p);
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t s = a + b;
    return min(s, s - p);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b + p;
    return min(d, d - p);
}

// a * b * 2^-32 mod p (Montgomery 1985), a, b < p; ninv = -p^-1 mod 2^32.
__device__ __forceinline__ uint32_t mont_mul(uint32_t a, uint32_t b, uint32_t p, uint32_t ninv) {
    const uint64_t t = uint64_t(a) * b;
    const uint32_t m = uint32_t(t) * ninv;
    const uint32_t r = uint32_t((t + uint64_t(m) * p) >> 32);
    return min(r, r - p);
}

// (x, y) -> (x + w y, x - w y)
__device__ __forceinline__ void ct_bfly(uint32_t& x, uint32_t& y, uint2 w, uint32_t p) {
    const uint32_t v = mul_shoup(y, w, p);
    const uint32_t s = add_mod(x, v, p);
    y = sub_mod(x, v, p);
    x = s;
}

// (x, y) -> (x + y, (x - y) w)
__device__ __forceinline__ void gs_bfly(uint32_t& x, uint32_t& y, uint2 w, uint32_t p) {
    const uint32_t s = add_mod(x, y, p);
    y = mul_shoup(x - y + p, w, p);
    x = s;
}

// K transform levels on the thread's 16 registers, which hold (16 >> K)
// sub-blocks of 2^K consecutive registers.  Local level l is level e0 + l of
// a length-M transform; the sub-block's outer index is o0 + sb; the twiddle
// of group g is z[(mult << (e0 + l)) + ((o0 + sb) << l) + g].
template <int K>
__device__ __forceinline__ void ct_round(uint32_t (&v)[16], const uint2* __restrict__ z,
                                         uint32_t mult, uint32_t e0, uint32_t o0, uint32_t p) {
#pragma unroll
    for (int l = 0; l < K; ++l) {
        constexpr int dummy = 0;
        (void)dummy;
        const int half = 1 << (K - 1 - l);
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb) {
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) {
                const uint2 w = z[(mult << (e0 + l)) + ((o0 + sb) << l) + g];
#pragma unroll
                for (int x = 0; x < half; ++x) {
                    const int i0 = (sb << K) + g * 2 * half + x;
                    ct_bfly(v[i0], v[i0 + half], w, p);
                }
            }
        }
    }
}

template <int K>
__device__ __forceinline__ void gs_round(uint32_t (&v)[16], const uint2* __restrict__ z,
                                         uint32_t mult, uint32_t e0, uint32_t o0, uint32_t p) {
#pragma unroll
    for (int l = K - 1; l >= 0; --l) {
        const int half = 1 << (K - 1 - l);
#pragma unroll
        for (int sb = 0; sb < (16 >> K); ++sb) {
#pragma unroll
            for (int g = 0; g < (1 << l); ++g) {
                const uint2 w = z[(mult << (e0 + l)) + ((o0 + sb) << l) + g];
#pragma unroll
                for (int x = 0; x < half; ++x) {
                    const int i0 = (sb << K) + g * 2 * half + x;
                    gs_bfly(v[i0], v[i0 + half], w, p);
                }
            }
        }
    }
}

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
__global__ __launch_bounds__(kThreads) void k_encode(DevTables t, const uint32_t* __restrict__ a,
                                                     const uint32_t* __restrict__ b,
                                                     uint32_t* __restrict__ planes) {
    extern __shared__ uint32_t enc_tile[];
    const uint32_t op = blockIdx.y, prod = blockIdx.z;
    const uint32_t* in = (op ? b : a) + size_t(prod) * t.N * t.L;
    const uint32_t f0 = blockIdx.x * kEncCoefs;
    const uint32_t row_words = t.Kenc / 4;
    const uint32_t stride_words = row_words + 4;
    for (uint32_t i = threadIdx.x; i < kEncCoefs * row_words; i += kThreads) {
        const uint32_t c = i / row_words, w = i - c * row_words;
        const uint32_t f = f0 + c;
        const uint32_t n = (f >> t.n1) + t.N2 * (f & (t.N1 - 1));
        enc_tile[c * stride_words + w] = w < t.L ? in[size_t(n) * t.L + w] : 0u;
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 31, kh = lane >> 5;
    const uint8_t* tile_bytes = reinterpret_cast<const uint8_t*>(enc_tile);
    const uint32_t stride_bytes = stride_words * 4;
    const v4i bias = {int(0x80808080u), int(0x80808080u), int(0x80808080u), int(0x80808080u)};
    for (uint32_t pt = wave; pt < t.PTenc; pt += 4) {
        v16i acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = v16i{};
        const int8_t* wrow = t.wenc + (size_t(pt) * 32 + row) * t.Kenc + 16 * kh;
        for (uint32_t kk = 0; kk < t.Kenc; kk += 32) {
            const v4i A = *reinterpret_cast<const v4i*>(wrow + kk);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const v4i B = *reinterpret_cast<const v4i*>(tile_bytes + (ct * 32 + row) * stride_bytes +
                                                            kk + 16 * kh) ^ bias;
                acc[ct] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B, acc[ct], 0, 0, 0);
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint32_t j = pt * 8 + 2 * g + kh;
            if (j < t.P) {
                const uint32_t p = t.primes[j];
                const uint64_t off = t.enc_off[j], mu = t.enc_mu[j];
                uint32_t* dst = plane_ptr(t, planes, prod, op, j) + f0 + row;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const int64_t v = int64_t(acc[ct][4 * g]) + (int64_t(acc[ct][4 * g + 1]) << 8) +
                                      (int64_t(acc[ct][4 * g + 2]) << 16) +
                                      (int64_t(acc[ct][4 * g + 3]) << 24);
                    const uint64_t u = uint64_t(v + int64_t(off));
                    // u < 2^47: qh = floor(u mu / 2^64) is floor(u/p) or one less.
                    const uint64_t qh = __umul64hi(u, mu);
                    const uint32_t r = uint32_t(u) - uint32_t(qh) * p;
                    dst[ct * 32] = min(r, r - p);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Column transforms (first / last n1 levels).  T = N1/16 threads per column,
// 16 registers each; 4096/N1 columns per workgroup, all of one prime.
// ---------------------------------------------------------------------------
template <int NLOG1>
__global__ __launch_bounds__(kThreads) void k_fwd_cols(DevTables t, uint32_t* __restrict__ planes) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T, S = N1 + 1;
    constexpr int K2 = NLOG1 - 4;
    __shared__ uint32_t s[K2 > 0 ? CPW * S : 1];
    const uint32_t wg_per_prime = t.N2 / CPW;
    const uint32_t j = blockIdx.x / wg_per_prime;
    const uint32_t r0 = (blockIdx.x % wg_per_prime) * CPW;
    const uint32_t cw = threadIdx.x / T, tau = threadIdx.x % T;
    uint32_t* col = plane_ptr(t, planes, blockIdx.z, blockIdx.y, j) + size_t(r0 + cw) * N1;
    const uint2* z = t.zf + size_t(j) * t.N;
    const uint32_t p = t.primes[j];
    uint32_t v[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = col[tau + T * x];
    ct_round<4>(v, z, 1, 0, 0, p);
    if constexpr (K2 > 0) {
#pragma unroll
        for (int x = 0; x < 16; ++x) s[cw * S + tau + T * x] = v[x];
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = s[cw * S + 16 * tau + x];
        ct_round<K2>(v, z, 1, 4, tau * (16 >> K2), p);
#pragma unroll
        for (int x = 0; x < 16; ++x) col[16 * tau + x] = v[x];
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) col[x] = v[x];
    }
}

template <int NLOG1>
__global__ __launch_bounds__(kThreads) void k_inv_cols(DevTables t, uint32_t* __restrict__ planes) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T, S = N1 + 1;
    constexpr int K2 = NLOG1 - 4;
    __shared__ uint32_t s[K2 > 0 ? CPW * S : 1];
    const uint32_t wg_per_prime = t.N2 / CPW;
    const uint32_t j = blockIdx.x / wg_per_prime;
    const uint32_t r0 = (blockIdx.x % wg_per_prime) * CPW;
    const uint32_t cw = threadIdx.x / T, tau = threadIdx.x % T;
    uint32_t* col = plane_ptr(t, planes, blockIdx.z, 0, j) + size_t(r0 + cw) * N1;
    const uint2* z = t.zi + size_t(j) * t.N;
    const uint32_t p = t.primes[j];
    uint32_t v[16];
    if constexpr (K2 > 0) {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = col[16 * tau + x];
        gs_round<K2>(v, z, 1, 4, tau * (16 >> K2), p);
#pragma unroll
        for (int x = 0; x < 16; ++x) s[cw * S + 16 * tau + x] = v[x];
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = s[cw * S + tau + T * x];
    } else {
#pragma unroll
        for (int x = 0; x < 16; ++x) v[x] = col[x];
    }
    gs_round<4>(v, z, 1, 0, 0, p);
    const uint2 sc = t.scale[j];
#pragma unroll
    for (int x = 0; x < 16; ++x) col[tau + T * x] = mul_shoup(v[x], sc, p);
}

// ---------------------------------------------------------------------------
// Middle: 16 blocks (256 contiguous natural coefficients each) of one prime,
// forward levels n1..n-1 for A and B, pointwise product, inverse levels back.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_middle(DevTables t, uint32_t* __restrict__ planes) {
    __shared__ uint32_t sa[kMidBlocks * kMidStride];
    __shared__ uint32_t sb[kMidBlocks * kMidStride];
    const uint32_t groups = t.N1 / kMidBlocks;
    const uint32_t j = blockIdx.x / groups;
    const uint32_t b0 = (blockIdx.x % groups) * kMidBlocks;
    const uint32_t prod = blockIdx.z;
    const uint32_t p = t.primes[j];
    uint32_t* pa = plane_ptr(t, planes, prod, 0, j);
    uint32_t* pb = plane_ptr(t, planes, prod, 1, j);
    for (uint32_t i = threadIdx.x; i < kMidBlocks * 256; i += kThreads) {
        const uint32_t u = i >> 4, bl = i & 15;
        sa[bl * kMidStride + u] = pa[size_t(u) * t.N1 + b0 + bl];
        sb[bl * kMidStride + u] = pb[size_t(u) * t.N1 + b0 + bl];
    }
    __syncthreads();
    const uint32_t bl = threadIdx.x >> 4, tau = threadIdx.x & 15;
    const uint32_t mult = t.N1 + b0 + bl;
    const uint2* zf = t.zf + size_t(j) * t.N;
    const uint2* zi = t.zi + size_t(j) * t.N;
    uint32_t* ra = sa + bl * kMidStride;
    uint32_t* rb = sb + bl * kMidStride;
    uint32_t v[16], w[16];
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        v[x] = ra[tau + 16 * x];
        w[x] = rb[tau + 16 * x];
    }
    ct_round<4>(v, zf, mult, 0, 0, p);
    ct_round<4>(w, zf, mult, 0, 0, p);
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        ra[tau + 16 * x] = v[x];
        rb[tau + 16 * x] = w[x];
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 16; ++x) {
        v[x] = ra[16 * tau + x];
        w[x] = rb[16 * tau + x];
    }
    ct_round<4>(v, zf, mult, 4, tau, p);
    ct_round<4>(w, zf, mult, 4, tau, p);
    const uint32_t ninv = t.mont_neg_inv[j];
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = mont_mul(v[x], w[x], p, ninv);
    gs_round<4>(v, zi, mult, 4, tau, p);
#pragma unroll
    for (int x = 0; x < 16; ++x) ra[16 * tau + x] = v[x];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 16; ++x) v[x] = ra[tau + 16 * x];
    gs_round<4>(v, zi, mult, 0, 0, p);
#pragma unroll
    for (int x = 0; x < 16; ++x) ra[tau + 16 * x] = v[x];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kMidBlocks * 256; i += kThreads) {
        const uint32_t u = i >> 4, b = i & 15;
        pa[size_t(u) * t.N1 + b0 + b] = sa[b * kMidStride + u];
    }
}

// ---------------------------------------------------------------------------
// CRT: 64 coefficients (consecutive column-major positions).  Terms t_j are
// in operand-A planes.  alpha = round(sum t_j / p_j) in FP64 (the exact value
// is within 1/4 of an integer; the FP64 error is below 2^-37).  Then
//   z = sum_{j<=P, a<4} byte_a(t_j) * W[j,a],  W[j,a] = 256^a w_j mod q,
// as an int8 MFMA GEMM over balanced signed weight digits and biased bytes,
// a signed carry chain over 4-digit limbs, and r = z - Q q with Q from the
// top 64 bits (three conditional subtractions at most).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_crt(DevTables t, const uint32_t* __restrict__ planes,
                                                  uint32_t* __restrict__ out) {
    extern __shared__ __align__(16) uint8_t crt_smem[];
    const uint32_t tstride = t.Kcrt + 16;  // bytes per term row
    const uint32_t tw = tstride / 4;
    const uint32_t kw = t.Kcrt / 4;
    uint8_t* tmat = crt_smem;
    uint32_t* tmw = reinterpret_cast<uint32_t*>(crt_smem);
    int64_t* lo = reinterpret_cast<int64_t*>(crt_smem + kCrtCoefs * tstride);
    const uint32_t L = t.L, P = t.P, L1 = L + 1;
    const uint32_t f0 = blockIdx.x * kCrtCoefs, prod = blockIdx.z;
    const uint32_t* pa = planes + size_t(prod) * 2 * P * t.N;
    for (uint32_t i = threadIdx.x; i < P * kCrtCoefs; i += kThreads) {
        const uint32_t j = i >> 6, c = i & 63;
        tmw[c * tw + j] = pa[size_t(j) * t.N + f0 + c];
    }
    const uint32_t padw = kw - P;
    for (uint32_t i = threadIdx.x; i < kCrtCoefs * padw; i += kThreads) {
        const uint32_t c = i / padw, w = P + (i - c * padw);
        tmw[c * tw + w] = 0u;
    }
    __syncthreads();
    {
        const uint32_t c = threadIdx.x >> 2, part = threadIdx.x & 3;
        double s = 0.0;
        for (uint32_t j = part; j < P; j += 4) s += double(tmw[c * tw + j]) * t.invp[j];
        s += __shfl_xor(s, 1, 4);
        s += __shfl_xor(s, 2, 4);
        if (part == 0) tmw[c * tw + P] = uint32_t(s + 0.5);
    }
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row = lane & 31, kh = lane >> 5;
    const v4i bias = {int(0x80808080u), int(0x80808080u), int(0x80808080u), int(0x80808080u)};
    for (uint32_t mt = wave; mt < t.Mcrt / 32; mt += 4) {
        v16i acc0 = v16i{}, acc1 = v16i{};
        const int8_t* wrow = t.wcrt + size_t(mt * 32 + row) * t.Kcrt + 16 * kh;
        for (uint32_t kk = 0; kk < t.Kcrt; kk += 32) {
            const v4i A = *reinterpret_cast<const v4i*>(wrow + kk);
            const v4i B0 = *reinterpret_cast<const v4i*>(tmat + row * tstride + kk + 16 * kh) ^ bias;
            const v4i B1 = *reinterpret_cast<const v4i*>(tmat + (32 + row) * tstride + kk + 16 * kh) ^ bias;
            acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B1, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint32_t limb = mt * 8 + 2 * g + kh;
            if (limb <= L) {
                const int64_t cr = t.corr[limb];
                lo[row * L1 + limb] = int64_t(acc0[4 * g]) + (int64_t(acc0[4 * g + 1]) << 8) +
                                      (int64_t(acc0[4 * g + 2]) << 16) +
                                      (int64_t(acc0[4 * g + 3]) << 24) + cr;
                lo[(32 + row) * L1 + limb] = int64_t(acc1[4 * g]) + (int64_t(acc1[4 * g + 1]) << 8) +
                                             (int64_t(acc1[4 * g + 2]) << 16) +
                                             (int64_t(acc1[4 * g + 3]) << 24) + cr;
            }
        }
    }
    __syncthreads();

    if (threadIdx.x < kCrtCoefs) {
        int64_t* z = lo + threadIdx.x * L1;
        int64_t carry = 0;
        for (uint32_t i = 0; i < L; ++i) {
            const int64_t tot = carry + z[i];
            z[i] = int64_t(uint32_t(tot));
            carry = tot >> 32;
        }
        int64_t top = carry + z[L];  // z = sum digits 2^(32 i) + top 2^(32 L), 0 <= z < 2^(W+17)
        auto word = [&](uint32_t i) -> uint64_t {
            return i < L ? uint64_t(uint32_t(z[i])) : (i == L ? uint64_t(top) : 0);
        };
        const uint32_t a = t.qshift >> 5, bsh = t.qshift & 31;
        uint64_t zt = (word(a) | (word(a + 1) << 32)) >> bsh;
        if (bsh) zt |= word(a + 2) << (64 - bsh);
        const uint64_t qf = uint64_t(double(zt) * t.inv_qt1);
        const uint64_t Q = qf ? qf - 1 : 0;
        uint64_t mc = 0;
        int64_t br = 0;
        for (uint32_t i = 0; i < L; ++i) {
            const uint64_t pq = Q * t.q[i] + mc;
            mc = pq >> 32;
            const int64_t d = int64_t(uint32_t(z[i])) - int64_t(uint32_t(pq)) - br;
            br = d < 0;
            z[i] = int64_t(uint32_t(d));
        }
        top -= int64_t(mc) + br;
        for (int it = 0; it < 3; ++it) {
            bool ge = top > 0;
            if (top == 0) {
                ge = true;
                for (int i = int(L) - 1; i >= 0; --i) {
                    const uint32_t zi = uint32_t(z[i]), qi = t.q[i];
                    if (zi != qi) {
                        ge = zi > qi;
                        break;
                    }
                }
            }
            if (ge) {
                int64_t b2 = 0;
                for (uint32_t i = 0; i < L; ++i) {
                    const int64_t d = int64_t(uint32_t(z[i])) - int64_t(t.q[i]) - b2;
                    b2 = d < 0;
                    z[i] = int64_t(uint32_t(d));
                }
                top -= b2;
            }
        }
    }
    __syncthreads();
    uint32_t* dst = out + size_t(prod) * t.N * L;
    for (uint32_t i = threadIdx.x; i < kCrtCoefs * L; i += kThreads) {
        const uint32_t c = i / L, limb = i - c * L;
        const uint32_t f = f0 + c;
        const uint32_t n = (f >> t.n1) + t.N2 * (f & (t.N1 - 1));
        dst[size_t(n) * L + limb] = uint32_t(lo[c * L1 + limb]);
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

template <int NLOG1>
hipError_t launch_cols(const DevTables& t, const BatchArgs& args, hipStream_t stream, bool forward) {
    constexpr uint32_t N1 = 1u << NLOG1, T = N1 / 16, CPW = kThreads / T;
    const dim3 grid(t.P * (t.N2 / CPW), forward ? 2 : 1, args.batch);
    if (forward)
        hipLaunchKernelGGL(k_fwd_cols<NLOG1>, grid, dim3(kThreads), 0, stream, t, args.planes);
    else
        hipLaunchKernelGGL(k_inv_cols<NLOG1>, grid, dim3(kThreads), 0, stream, t, args.planes);
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

size_t crt_smem_bytes(const DevTables& t) {
    return size_t(kCrtCoefs) * (t.Kcrt + 16) + size_t(kCrtCoefs) * (t.L + 1) * 8;
}

size_t enc_smem_bytes(const DevTables& t) { return size_t(kEncCoefs) * (t.Kenc + 16); }

bool shapes_ok(const DevTables& t, const BatchArgs& args) {
    return args.batch >= 1 && args.batch <= 65535 && t.N2 == 256 && t.N1 >= 16 && t.N1 <= 256 &&
           t.N == t.N1 * t.N2 && t.N % kEncCoefs == 0 && t.Kenc % 32 == 0 && t.Kcrt % 32 == 0 &&
           t.Mcrt % 32 == 0 && t.P >= 1 && t.P <= 128 && t.L >= 3 && t.L <= 64 &&
           t.Kcrt >= 4 * (t.P + 1) && t.Mcrt >= 4 * t.L + 1 && t.Kenc >= 4 * t.L && t.qshift + 32 == t.W;
}

}  // namespace

hipError_t launch_encode(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    const dim3 grid(t.N / kEncCoefs, 2, args.batch);
    hipLaunchKernelGGL(k_encode, grid, dim3(kThreads), enc_smem_bytes(t), stream, t, args.a, args.b,
                       args.planes);
    return hipGetLastError();
}

hipError_t launch_forward_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    return dispatch_cols(t, args, stream, true);
}

hipError_t launch_middle(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    const dim3 grid(t.P * (t.N1 / kMidBlocks), 1, args.batch);
    hipLaunchKernelGGL(k_middle, grid, dim3(kThreads), 0, stream, t, args.planes);
    return hipGetLastError();
}

hipError_t launch_inverse_columns(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    return dispatch_cols(t, args, stream, false);
}

hipError_t launch_crt(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    if (!shapes_ok(t, args)) return hipErrorInvalidValue;
    const dim3 grid(t.N / kCrtCoefs, 1, args.batch);
    hipLaunchKernelGGL(k_crt, grid, dim3(kThreads), crt_smem_bytes(t), stream, t, args.planes, args.out);
    return hipGetLastError();
}

hipError_t launch_product(const DevTables& t, const BatchArgs& args, hipStream_t stream) {
    hipError_t e;
    if ((e = launch_encode(t, args, stream)) != hipSuccess) return e;
    if ((e = launch_forward_columns(t, args, stream)) != hipSuccess) return e;
    if ((e = launch_middle(t, args, stream)) != hipSuccess) return e;
    if ((e = launch_inverse_columns(t, args, stream)) != hipSuccess) return e;
    return launch_crt(t, args, stream);
}

hipError_t launch_mfma_probe(const int8_t* a, const int8_t* b, int32_t* c, hipStream_t stream) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, stream, a, b, c);
    return hipGetLastError();
}

hipError_t configure_kernels() {
    // CRT may need more than the 64 KiB default dynamic LDS.
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k_crt),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

}  // namespace bcrt
