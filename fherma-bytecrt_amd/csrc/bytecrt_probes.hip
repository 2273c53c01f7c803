// Measurement probes for gfx950 (include/bytecrt_probes.h): HBM stream
// kernels, butterfly kernels for three exact modular multiplications, and the
// overlap of MFMA and VALU work from different waves of one SIMD.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bytecrt.h"
#include "../../include/bytecrt_probes.h"

namespace {

constexpr uint32_t kThreads = 256;
constexpr int kInFlight = 4;       // 16-byte accesses per thread in flight
constexpr uint32_t kGridPerCu = 32;  // workgroups per CU (a sweep of 4..32 on MI355X: 32 is best)

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Grid-stride over 16-byte words; each step issues kInFlight loads before its
// stores (copy), or only loads (read), or only stores (write).  Copy and read
// use non-temporal accesses (the data is touched once).
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_stream(const v4u* __restrict__ src, v4u* __restrict__ dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * kThreads;
    const size_t tid = size_t(blockIdx.x) * kThreads + threadIdx.x;
    v4u acc = v4u{0, 0, 0, 0};
    size_t i = tid;
    for (; i + (kInFlight - 1) * stride < n; i += kInFlight * stride) {
        if constexpr (MODE == 2) {
#pragma unroll
            for (int k = 0; k < kInFlight; ++k) dst[i + k * stride] = v4u{uint32_t(i), 1u, 2u, uint32_t(k)};
        } else {
            v4u v[kInFlight];
#pragma unroll
            for (int k = 0; k < kInFlight; ++k) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
            for (int k = 0; k < kInFlight; ++k) {
                if constexpr (MODE == 0)
                    __builtin_nontemporal_store(v[k], dst + i + k * stride);
                else
                    acc ^= v[k];
            }
        }
    }
    for (; i < n; i += stride) {
        if constexpr (MODE == 0) dst[i] = src[i];
        if constexpr (MODE == 1) acc ^= src[i];
        if constexpr (MODE == 2) dst[i] = v4u{uint32_t(i), 1u, 2u, 3u};
    }
    if constexpr (MODE == 1) dst[tid] = acc;
}

// ---- butterfly forms -------------------------------------------------------
struct Twiddles {
    uint32_t w[15];      // w in [0, p)
    uint32_t shoup[15];  // floor(w 2^32 / p)
    uint32_t mont[15];   // w 2^32 mod p
    uint32_t p;
    uint32_t pneg_inv;   // -p^-1 mod 2^32
    double pd, pinv;
};

constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52

// The product kernels' FP64 form: exact for |x| <= 2^40, |r| <= p/2 + 2^-12 p.
__device__ __forceinline__ double mul_mod_f64(double x, double w, double p, double pinv) {
    const double h = x * w;
    const double l = fma(x, w, -h);
    const double q = fma(h, pinv, kRound) - kRound;
    return fma(-q, p, h) + l;
}

// The same with the product x w formed by an FMA with a -0.0 addend (the
// same double, bit for bit): v_fma_f64 instead of v_mul_f64.  nz must be
// -0.0 held where the compiler cannot see it (it would fold the FMA back).
__device__ __forceinline__ double mul_mod_f64_fma(double x, double w, double p, double pinv, double nz) {
    const double h = fma(x, w, nz);
    const double l = fma(x, w, -h);
    const double q = fma(h, pinv, kRound) - kRound;
    return fma(-q, p, h) + l;
}

// The NTT kernels' five-operation FP64 form for p - 1 = k 2^s: wq = fl(w/p),
// q = rint(x wq), x w - q p = fma(x, w, -(q p - q)) - q, exact while |q| k < 2^53
// (the probe's inputs stay below 2.1 p; the entry point checks 3 p k < 2^53).
__device__ __forceinline__ double mul_mod_tw(double x, double w, double wq, double p) {
    const double q = fma(x, wq, kRound) - kRound;
    return fma(x, w, -fma(q, p, -q)) - q;
}

// x in [0, p), w in [0, p), ws = floor(w 2^32 / p): x w mod p in [0, p).
__device__ __forceinline__ uint32_t mul_mod_shoup(uint32_t x, uint32_t w, uint32_t ws, uint32_t p) {
    const uint32_t q = __umulhi(x, ws);
    const uint32_t r = x * w - q * p;  // in [0, 2p), exact mod 2^32
    return min(r, r - p);
}

// x in [0, p), wm = w 2^32 mod p: REDC(x wm) = x w mod p in [0, p).
__device__ __forceinline__ uint32_t mul_mod_mont(uint32_t x, uint32_t wm, uint32_t p, uint32_t pneg_inv) {
    const uint64_t t = uint64_t(x) * wm;
    const uint32_t m = uint32_t(t) * pneg_inv;
    const uint64_t u = (t + uint64_t(m) * p) >> 32;  // < 2p
    const uint32_t r = uint32_t(u);
    return min(r, r - p);
}

__device__ __forceinline__ uint32_t add_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t s = a + b;  // < 2^32 for p < 2^31
    return min(s, s - p);
}

__device__ __forceinline__ uint32_t sub_mod(uint32_t a, uint32_t b, uint32_t p) {
    const uint32_t d = a - b;
    return min(d, d + p);
}

// 16 starting residues of a thread, canonical.
__device__ __forceinline__ uint32_t start_value(size_t tid, int k, uint32_t p) {
    uint64_t z = uint64_t(tid) * 16 + uint64_t(k) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return uint32_t((z ^ (z >> 31)) % p);
}

template <int FORM>
__global__ __launch_bounds__(kThreads) void k_bfly(uint32_t* __restrict__ out, const Twiddles tw, uint32_t rounds) {
    const size_t tid = size_t(blockIdx.x) * kThreads + threadIdx.x;
    const uint32_t p = tw.p;
    if constexpr (FORM == 3 || FORM == 4) {
        // five-operation form; FORM 4 writes the two butterflies of each pair
        // (consecutive in the level's order) step by step, interleaved
        double v[16], w[15], wq[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            w[k] = double(tw.w[k]);
            wq[k] = w[k] * tw.pinv;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = double(start_value(tid, k, p));
        for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int half = 8 >> l;
#pragma unroll
                for (int b = 0; b < 8; b += 2) {
                    int i0[2], k[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const int g = (b + u) / half, x = (b + u) % half;
                        i0[u] = g * 2 * half + x;
                        k[u] = (1 << l) - 1 + g;
                    }
                    if constexpr (FORM == 3) {
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const double t = mul_mod_tw(v[i0[u] + half], w[k[u]], wq[k[u]], tw.pd);
                            v[i0[u] + half] = v[i0[u]] - t;
                            v[i0[u]] = v[i0[u]] + t;
                        }
                    } else {
                        // (sched_barrier keeps the compiler from putting each
                        // chain back together: it treats the FP64 results as
                        // available at once)
                        double q[2], c[2], t[2];
#pragma unroll
                        for (int u = 0; u < 2; ++u) q[u] = fma(v[i0[u] + half], wq[k[u]], kRound);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < 2; ++u) q[u] = q[u] - kRound;
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < 2; ++u) c[u] = fma(q[u], tw.pd, -q[u]);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < 2; ++u) t[u] = fma(v[i0[u] + half], w[k[u]], -c[u]);
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < 2; ++u) t[u] = t[u] - q[u];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            v[i0[u] + half] = v[i0[u]] - t[u];
                            v[i0[u]] = v[i0[u]] + t[u];
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = fma(-(fma(v[k], tw.pinv, kRound) - kRound), tw.pd, v[k]);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double r = v[k] < 0.0 ? v[k] + tw.pd : v[k];
            out[tid * 16 + k] = uint32_t(r);
        }
    } else if constexpr (FORM == 0 || FORM == 5) {
        double nz = -0.0;
        __asm__ volatile("" : "+v"(nz));  // opaque: keeps fma(x, w, nz) an FMA
        double v[16], w[15];
#pragma unroll
        for (int k = 0; k < 15; ++k) w[k] = double(tw.w[k]);
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = double(start_value(tid, k, p));
        for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int half = 8 >> l;
#pragma unroll
                for (int g = 0; g < (1 << l); ++g)
#pragma unroll
                    for (int x = 0; x < half; ++x) {
                        const int i0 = g * 2 * half + x;
                        const double t =
                            FORM == 0 ? mul_mod_f64(v[i0 + half], w[(1 << l) - 1 + g], tw.pd, tw.pinv)
                                      : mul_mod_f64_fma(v[i0 + half], w[(1 << l) - 1 + g], tw.pd, tw.pinv, nz);
                        v[i0 + half] = v[i0] - t;
                        v[i0] = v[i0] + t;
                    }
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = fma(-(fma(v[k], tw.pinv, kRound) - kRound), tw.pd, v[k]);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double r = v[k] < 0.0 ? v[k] + tw.pd : v[k];
            out[tid * 16 + k] = uint32_t(r);
        }
    } else {
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = start_value(tid, k, p);
        for (uint32_t r = 0; r < rounds; ++r) {
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                const int half = 8 >> l;
#pragma unroll
                for (int g = 0; g < (1 << l); ++g)
#pragma unroll
                    for (int x = 0; x < half; ++x) {
                        const int i0 = g * 2 * half + x, k = (1 << l) - 1 + g;
                        const uint32_t t = FORM == 1 ? mul_mod_shoup(v[i0 + half], tw.w[k], tw.shoup[k], p)
                                                     : mul_mod_mont(v[i0 + half], tw.mont[k], p, tw.pneg_inv);
                        v[i0 + half] = sub_mod(v[i0], t, p);
                        v[i0] = add_mod(v[i0], t, p);
                    }
            }
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) out[tid * 16 + k] = v[k];
    }
}

// ---- MFMA and VALU issue overlap -------------------------------------------
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr uint32_t kOverlapThreads = 512;  // waves 0..3 MFMA, 4..7 VALU: one of each per SIMD
constexpr uint32_t kOverlapGridPerCu = 2;

// The MFMA waves run four independent v_mfma_i32_32x32x32_i8 chains
// (mfma_iters x 4 MFMAs); the VALU waves run eight independent chains of one
// operation (valu_iters x 8): kind 0 FP64 fma, 1 FP32 fma, 2 int32 24-bit
// multiply-add.  mode bit 1 runs the MFMA waves, bit 2 the VALU waves; the
// others exit at once.  Each running thread writes one word of its results.
// (MODE and KIND are template parameters so that a kernel trace names each
// combination.)
template <int MODE, int KIND>
__global__ __launch_bounds__(kOverlapThreads) void k_overlap(uint32_t mfma_iters, uint32_t valu_iters,
                                                             int32_t* __restrict__ out) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t slot = size_t(blockIdx.x) * kOverlapThreads + threadIdx.x;
    if (wave < 4) {
        if constexpr (!(MODE & 1)) return;
        const v4i a = {int(lane), int(lane) + 1, int(lane) + 2, int(lane) + 3};
        const v4i b = {int(lane) * 3, int(lane), 7, 9};
        v16i c0 = {}, c1 = {}, c2 = {}, c3 = {};
        for (uint32_t i = 0; i < mfma_iters; ++i) {
            c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c3, 0, 0, 0);
        }
        int32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) sum += c0[k] + c1[k] + c2[k] + c3[k];
        out[slot] = sum;
        return;
    }
    if constexpr (!(MODE & 2)) return;
    if constexpr (KIND == 0) {
        double x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = double(lane + k);
        for (uint32_t i = 0; i < valu_iters; ++i)
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = fma(x[k], 1.0000001, 0.5);
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += x[k];
        out[slot] = int32_t(__double2loint(sum) ^ __double2hiint(sum));
    } else if constexpr (KIND == 1) {
        float x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = float(lane + k);
        for (uint32_t i = 0; i < valu_iters; ++i)
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], 1.0000001f, 0.5f);
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += x[k];
        out[slot] = __float_as_int(sum);
    } else {
        uint32_t x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = lane + k;
        for (uint32_t i = 0; i < valu_iters; ++i)
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = __umul24(x[k], 0x9E3779u) + 12345u;
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += x[k];
        out[slot] = int32_t(sum);
    }
}

}  // namespace

extern "C" int bcrtp_issue_overlap(int mode, int kind, uint32_t mfma_iters, uint32_t valu_iters, int32_t* d_out,
                                   size_t out_words, void* stream) {
    if (mode < 1 || mode > 3 || kind < 0 || kind > 2 || !d_out) return BCRT_EINVAL;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return BCRT_ERUNTIME;
    const dim3 grid(uint32_t(cus) * kOverlapGridPerCu), blk(kOverlapThreads);
    if (out_words < size_t(grid.x) * kOverlapThreads) return BCRT_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    void (*const kernels[3][3])(uint32_t, uint32_t, int32_t*) = {
        {k_overlap<1, 0>, k_overlap<1, 1>, k_overlap<1, 2>},
        {k_overlap<2, 0>, k_overlap<2, 1>, k_overlap<2, 2>},
        {k_overlap<3, 0>, k_overlap<3, 1>, k_overlap<3, 2>}};
    hipLaunchKernelGGL(kernels[mode - 1][kind], grid, blk, 0, s, mfma_iters, valu_iters, d_out);
    return hipGetLastError() == hipSuccess ? BCRT_OK : BCRT_ERUNTIME;
}

extern "C" size_t bcrtp_issue_overlap_words() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return size_t(cus) * kOverlapGridPerCu * kOverlapThreads;
}

extern "C" int bcrtp_hbm_stream(int mode, const void* d_src, void* d_dst, size_t bytes, void* stream) {
    if (mode < 0 || mode > 2 || bytes == 0 || bytes % 16 || !d_dst || (mode != 2 && !d_src)) return BCRT_EINVAL;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return BCRT_ERUNTIME;
    const size_t n = bytes / 16;
    const dim3 grid(uint32_t(cus) * kGridPerCu), blk(kThreads);
    if (mode == 1 && n < size_t(grid.x) * kThreads) return BCRT_EINVAL;  // d_dst holds one word per thread
    hipStream_t s = static_cast<hipStream_t>(stream);
    const v4u* src = static_cast<const v4u*>(d_src);
    v4u* dst = static_cast<v4u*>(d_dst);
    if (mode == 0) hipLaunchKernelGGL(k_stream<0>, grid, blk, 0, s, src, dst, n);
    if (mode == 1) hipLaunchKernelGGL(k_stream<1>, grid, blk, 0, s, src, dst, n);
    if (mode == 2) hipLaunchKernelGGL(k_stream<2>, grid, blk, 0, s, src, dst, n);
    return hipGetLastError() == hipSuccess ? BCRT_OK : BCRT_ERUNTIME;
}

extern "C" int bcrtp_butterflies(int form, uint32_t* d_out, size_t threads, uint32_t p, const uint32_t* tw15,
                                 uint32_t rounds, uint32_t lds_bytes, void* stream) {
    if (form < 0 || form > 5 || !d_out || !tw15 || threads == 0 || threads % kThreads || p <= (1u << 30) ||
        p >= (1u << 31) || (p & 1) == 0 || threads / kThreads > 0x7FFFFFFF || lds_bytes > 65536)
        return BCRT_EINVAL;
    if (form == 3 || form == 4) {
        uint32_t k = p - 1;
        while (!(k & 1)) k >>= 1;
        if ((unsigned __int128)(uint64_t(3) * p) * k >= (unsigned __int128)1 << 53) return BCRT_EINVAL;
    }
    Twiddles tw{};
    tw.p = p;
    uint32_t inv = 1;  // p^-1 mod 2^32 by Newton iteration (p odd)
    for (int i = 0; i < 5; ++i) inv *= 2 - p * inv;
    tw.pneg_inv = 0u - inv;
    tw.pd = double(p);
    tw.pinv = 1.0 / double(p);
    for (int k = 0; k < 15; ++k) {
        if (tw15[k] >= p) return BCRT_EINVAL;
        tw.w[k] = tw15[k];
        tw.shoup[k] = uint32_t((uint64_t(tw15[k]) << 32) / p);
        tw.mont[k] = uint32_t((uint64_t(tw15[k]) << 32) % p);
    }
    const dim3 grid(uint32_t(threads / kThreads)), blk(kThreads);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (form == 0) hipLaunchKernelGGL(k_bfly<0>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    if (form == 1) hipLaunchKernelGGL(k_bfly<1>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    if (form == 2) hipLaunchKernelGGL(k_bfly<2>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    if (form == 3) hipLaunchKernelGGL(k_bfly<3>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    if (form == 4) hipLaunchKernelGGL(k_bfly<4>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    if (form == 5) hipLaunchKernelGGL(k_bfly<5>, grid, blk, lds_bytes, s, d_out, tw, rounds);
    return hipGetLastError() == hipSuccess ? BCRT_OK : BCRT_ERUNTIME;
}
