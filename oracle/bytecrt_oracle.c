/*
 * bytecrt_oracle.c -- CPU restatement of the ByteCRT exact negacyclic product.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the test suite, the
 * smoke entry and the CPU-baseline leg of bench.py load it to produce the
 * expected answer.  The product library never links, loads or calls it.
 *
 * Every function names the reference lines it restates (paths relative to the
 * reference checkout).  The reference only accepts q = 2^W - gap; this oracle
 * generalises the plan to any q with bit_length(q) == W (odd or even) and
 * replaces the near-power fold by a plain binary long division of the
 * quotient bits, which needs no property of q.  The near-power fold is kept
 * (orc_reduce_near_power) so the reference's own fixtures can be checked with
 * the reference's own reduction as well.
 *
 * Pinning (see tests/): plan constants against reference-result.json and
 * rns-plan-cpu-result.json pins and against rns-plan.hpp compiled in place
 * (oracle/_ref); CRT against the six BCRT01 fixtures; NTT against schoolbook
 * at small N; full product against Python big-integer schoolbook / sampled
 * coefficient sums and closed forms (zero, one, monomials).
 */
#include "bytecrt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* Fixed-capacity little-endian big integers (base 2^32).                  */
/* Restates the Words helpers of native-smoke/rns-plan.hpp:27-129.         */
/* ---------------------------------------------------------------------- */
#define BIG_WORDS 192 /* 6144 bits: product of up to 128 31-bit primes */

typedef struct {
    int n; /* words in use; trimmed so w[n-1] != 0, n == 0 is zero */
    uint32_t w[BIG_WORDS];
} Big;

static void big_trim(Big* a) {
    while (a->n > 0 && a->w[a->n - 1] == 0) a->n--;
}

static void big_set_words(Big* a, const uint32_t* w, int n) {
    memset(a, 0, sizeof(*a));
    memcpy(a->w, w, sizeof(uint32_t) * (size_t)n);
    a->n = n;
    big_trim(a);
}

static void big_set_u32(Big* a, uint32_t v) {
    memset(a, 0, sizeof(*a));
    a->w[0] = v;
    a->n = v ? 1 : 0;
}

/* rns-plan.hpp:34-39 */
static unsigned big_bitlen(const Big* a) {
    if (a->n == 0) return 0;
    unsigned top = 0;
    for (uint32_t t = a->w[a->n - 1]; t; t >>= 1) ++top;
    return 32u * (unsigned)(a->n - 1) + top;
}

/* rns-plan.hpp:41-47 */
static int big_cmp(const Big* a, const Big* b) {
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    for (int i = a->n; i > 0; --i)
        if (a->w[i - 1] != b->w[i - 1]) return a->w[i - 1] < b->w[i - 1] ? -1 : 1;
    return 0;
}

/* a -= b, requires a >= b (rns-plan.hpp:49-60) */
static void big_sub(Big* a, const Big* b) {
    uint64_t borrow = 0;
    for (int i = 0; i < a->n; ++i) {
        uint64_t s = (i < b->n ? (uint64_t)b->w[i] : 0) + borrow;
        uint64_t x = a->w[i];
        a->w[i] = (uint32_t)(x - s);
        borrow = x < s;
    }
    big_trim(a);
}

/* a *= s (rns-plan.hpp:62-71) */
static void big_mul_small(Big* a, uint32_t s) {
    uint64_t carry = 0;
    for (int i = 0; i < a->n; ++i) {
        uint64_t t = (uint64_t)a->w[i] * s + carry;
        a->w[i] = (uint32_t)t;
        carry = t >> 32;
    }
    if (carry) a->w[a->n++] = (uint32_t)carry;
    big_trim(a);
}

/* out = a * b (rns-plan.hpp:74-87) */
static void big_mul(Big* out, const Big* a, const Big* b) {
    Big r;
    memset(&r, 0, sizeof(r));
    for (int i = 0; i < a->n; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < b->n; ++j) {
            uint64_t t = (uint64_t)a->w[i] * b->w[j] + r.w[i + j] + carry;
            r.w[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r.w[i + b->n] = (uint32_t)carry;
    }
    r.n = a->n + b->n;
    big_trim(&r);
    *out = r;
}

/* quotient = a / d, returns remainder (rns-plan.hpp:89-101) */
static uint32_t big_div_small(Big* quo, const Big* a, uint32_t d) {
    Big q;
    memset(&q, 0, sizeof(q));
    uint64_t rem = 0;
    for (int i = a->n; i > 0; --i) {
        uint64_t num = (rem << 32) | a->w[i - 1];
        q.w[i - 1] = (uint32_t)(num / d);
        rem = num % d;
    }
    q.n = a->n;
    big_trim(&q);
    *quo = q;
    return (uint32_t)rem;
}

/* rns-plan.hpp:103-109 */
static uint32_t big_mod_small(const Big* a, uint32_t d) {
    uint64_t rem = 0;
    for (int i = a->n; i > 0; --i) rem = ((rem << 32) | a->w[i - 1]) % d;
    return (uint32_t)rem;
}

/* a <<= 1 */
static void big_shl1(Big* a) {
    uint32_t carry = 0;
    for (int i = 0; i < a->n; ++i) {
        uint32_t nc = a->w[i] >> 31;
        a->w[i] = (a->w[i] << 1) | carry;
        carry = nc;
    }
    if (carry) a->w[a->n++] = carry;
}

/* a mod q by binary long division.  The reference does the same bit-serial
 * remainder walk (rns-plan.hpp:112-129) over every bit of a; this one only
 * walks the quotient bits, which is the same division. */
static void big_mod(Big* r, const Big* a, const Big* q) {
    unsigned la = big_bitlen(a), lq = big_bitlen(q);
    Big x = *a;
    if (la >= lq) {
        for (int shift = (int)(la - lq); shift >= 0; --shift) {
            /* t = q << shift */
            Big t;
            memset(&t, 0, sizeof(t));
            int ws = shift / 32, bs = shift % 32;
            for (int i = 0; i < q->n; ++i) {
                uint64_t v = (uint64_t)q->w[i] << bs;
                t.w[i + ws] |= (uint32_t)v;
                t.w[i + ws + 1] |= (uint32_t)(v >> 32);
            }
            t.n = q->n + ws + 1;
            big_trim(&t);
            if (big_cmp(&x, &t) >= 0) big_sub(&x, &t);
        }
    }
    *r = x;
}

/* ---------------------------------------------------------------------- */
/* 32-bit modular helpers and primality (rns-plan.hpp:131-170,            */
/* reference.py:11-36).                                                    */
/* ---------------------------------------------------------------------- */
static uint32_t pow_mod(uint32_t base, uint32_t e, uint32_t m) {
    uint64_t r = 1, b = base % m;
    for (; e; e >>= 1) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
    }
    return (uint32_t)r;
}

int orc_prime32(uint32_t p) {
    static const uint32_t small[] = {2, 3, 5, 7, 11, 13, 17};
    if (p < 2) return 0;
    for (int i = 0; i < 7; ++i) {
        if (p == small[i]) return 1;
        if (p % small[i] == 0) return 0;
    }
    uint32_t d = p - 1, s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    static const uint32_t bases[] = {2, 7, 61};
    for (int i = 0; i < 3; ++i) {
        uint32_t a = bases[i];
        if (a % p == 0) continue;
        uint32_t x = pow_mod(a, d, p);
        if (x == 1 || x == p - 1) continue;
        int passed = 0;
        for (uint32_t r = 1; r < s; ++r) {
            x = (uint32_t)((uint64_t)x * x % p);
            if (x == p - 1) { passed = 1; break; }
        }
        if (!passed) return 0;
    }
    return 1;
}

/* rns-plan.hpp:163-170: smallest base a >= 2 whose (p-1)/2N power has order 2N */
static uint32_t primitive_2n_root(uint32_t p, uint32_t n) {
    uint32_t e = (p - 1) / (2 * n);
    for (uint32_t a = 2; a < p; ++a) {
        uint32_t psi = pow_mod(a, e, p);
        if (pow_mod(psi, n, p) == p - 1) return psi;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Plan.                                                                    */
/* ---------------------------------------------------------------------- */
struct OrcPlan {
    uint32_t n, w, l, pcount, kpad, terms, product_bits;
    uint32_t q[ORC_MAX_L];
    Big qbig;
    uint32_t primes[ORC_MAX_P], psi[ORC_MAX_P], omega[ORC_MAX_P], omega_inv[ORC_MAX_P];
    uint32_t psi_inv[ORC_MAX_P], n_inv[ORC_MAX_P], crt_inv[ORC_MAX_P];
    double invp[ORC_MAX_P];
    float invp_f32[ORC_MAX_P];
    /* weights[j][limb], j < pcount: P_j mod q; j == pcount: (-P) mod q */
    uint32_t* weights;   /* (pcount+1) * l */
    uint32_t* b4_weight; /* (pcount+1) * 4 * l: 256^a * weight_j mod q, [j][a][limb] */
    uint8_t* b7;         /* 7l columns * kpad, column-major (rns-plan.hpp:285-300) */
    uint8_t* b4;         /* 4l columns * kpad, column-major (rns-plan.hpp:305-322) */
};

static void set_err(char* err, int errlen, const char* msg) {
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
    }
}

/*
 * make_rns_plan restated (native-smoke/rns-plan.hpp:174-324 and
 * reference.py:39-54), generalised to any N = 2^k (2 <= N <= 2^20), any W
 * with L = ceil(W/32) and any q of exactly W bits.
 */
OrcPlan* orc_plan_create(uint32_t N, uint32_t W, const uint32_t* q, uint32_t L,
                         char* err, int errlen) {
    if (N < 2 || (N & (N - 1)) || N > (1u << 20)) {
        set_err(err, errlen, "ByteCRT oracle: N must be a power of two in [2, 2^20]");
        return NULL;
    }
    if (W < 2 || L != (W + 31) / 32 || L > ORC_MAX_L) {
        set_err(err, errlen, "ByteCRT oracle: L must equal ceil(W/32) and be <= ORC_MAX_L");
        return NULL;
    }
    OrcPlan* pl = (OrcPlan*)calloc(1, sizeof(OrcPlan));
    pl->n = N;
    pl->w = W;
    pl->l = L;
    memcpy(pl->q, q, sizeof(uint32_t) * L);
    big_set_words(&pl->qbig, q, (int)L);
    if (big_bitlen(&pl->qbig) != W) {
        set_err(err, errlen, "ByteCRT oracle: bit_length(q) must equal W");
        free(pl);
        return NULL;
    }
    /* Guard: bit_length(prod p) >= 2W + bit_length(N) + 2 (rns-plan.hpp:200-221,
     * reference.py:41-48), greedy descending primes p == 1 mod 2N in (2^30, 2^31). */
    unsigned n_bits = 0;
    for (uint32_t v = N; v; v >>= 1) ++n_bits;
    unsigned guard = 2 * W + n_bits + 2;
    uint32_t step = 2 * N;
    uint32_t p = ((1u << 31) - 2) / step * step + 1;
    Big product;
    big_set_u32(&product, 1);
    while (big_bitlen(&product) < guard) {
        while (p > (1u << 30) && !orc_prime32(p)) p -= step;
        if (p <= (1u << 30) || pl->pcount >= ORC_MAX_P) {
            set_err(err, errlen, "ByteCRT oracle: prime search exhausted");
            free(pl);
            return NULL;
        }
        pl->primes[pl->pcount++] = p;
        big_mul_small(&product, p);
        p -= step;
    }
    pl->product_bits = big_bitlen(&product);
    pl->terms = pl->pcount + 1;
    pl->kpad = (4 * pl->terms + 31) / 32 * 32;

    /* Centered CRT guard P > 4N(q-1)^2 (rns-plan.hpp:222-226, reference.py:49). */
    {
        Big qm1 = pl->qbig, one, bound;
        big_set_u32(&one, 1);
        big_sub(&qm1, &one);
        big_mul(&bound, &qm1, &qm1);
        big_mul_small(&bound, 4 * N);
        if (big_cmp(&product, &bound) <= 0) {
            set_err(err, errlen, "ByteCRT oracle: centered CRT guard failed");
            free(pl);
            return NULL;
        }
    }

    pl->weights = (uint32_t*)calloc((size_t)pl->terms * L, sizeof(uint32_t));
    for (uint32_t j = 0; j < pl->pcount; ++j) {
        uint32_t pj = pl->primes[j];
        uint32_t psi = primitive_2n_root(pj, N);
        uint32_t omega = (uint32_t)((uint64_t)psi * psi % pj);
        pl->psi[j] = psi;
        pl->psi_inv[j] = pow_mod(psi, pj - 2, pj);
        pl->omega[j] = omega;
        pl->omega_inv[j] = pow_mod(omega, pj - 2, pj);
        pl->n_inv[j] = pow_mod(N % pj, pj - 2, pj);
        Big base;
        uint32_t rem = big_div_small(&base, &product, pj);
        (void)rem;
        uint32_t base_mod_p = big_mod_small(&base, pj);
        pl->crt_inv[j] = pow_mod(base_mod_p, pj - 2, pj);
        pl->invp[j] = 1.0 / (double)pj;
        pl->invp_f32[j] = (float)pl->invp[j];
        Big wj;
        big_mod(&wj, &base, &pl->qbig);
        for (int i = 0; i < wj.n; ++i) pl->weights[(size_t)j * L + i] = wj.w[i];
    }
    /* Alpha weight (-P) mod q (rns-plan.hpp:276-283, reference.py:53). */
    {
        Big pmq;
        big_mod(&pmq, &product, &pl->qbig);
        if (pmq.n) {
            Big neg = pl->qbig;
            big_sub(&neg, &pmq);
            for (int i = 0; i < neg.n; ++i) pl->weights[(size_t)pl->pcount * L + i] = neg.w[i];
        }
    }
    /* B7: column 7*limb+d, row 4*j+a = byte_(d-a)(weight_j[limb])
     * (rns-plan.hpp:285-300, reference.py:57-71). */
    pl->b7 = (uint8_t*)calloc((size_t)7 * L * pl->kpad, 1);
    for (uint32_t limb = 0; limb < L; ++limb)
        for (uint32_t d = 0; d < 7; ++d) {
            size_t col = (size_t)(7 * limb + d) * pl->kpad;
            for (uint32_t j = 0; j < pl->terms; ++j) {
                uint32_t word = pl->weights[(size_t)j * L + limb];
                for (uint32_t a = 0; a < 4; ++a) {
                    int dd = (int)d - (int)a;
                    pl->b7[col + 4 * j + a] = (dd >= 0 && dd < 4) ? (uint8_t)(word >> (8 * dd)) : 0;
                }
            }
        }
    /* B4: weight W[j,a] = 256^a * weight_j mod q; column 4*limb+d holds byte d of
     * limb `limb` of W[j,a] at row 4j+a (rns-plan.hpp:301-322). */
    pl->b4_weight = (uint32_t*)calloc((size_t)pl->terms * 4 * L, sizeof(uint32_t));
    pl->b4 = (uint8_t*)calloc((size_t)4 * L * pl->kpad, 1);
    for (uint32_t j = 0; j < pl->terms; ++j) {
        Big sh;
        big_set_words(&sh, pl->weights + (size_t)j * L, (int)L);
        for (uint32_t a = 0; a < 4; ++a) {
            for (uint32_t limb = 0; limb < L; ++limb) {
                uint32_t word = (int)limb < sh.n ? sh.w[limb] : 0;
                pl->b4_weight[((size_t)j * 4 + a) * L + limb] = word;
                for (uint32_t d = 0; d < 4; ++d)
                    pl->b4[(size_t)(4 * limb + d) * pl->kpad + 4 * j + a] = (uint8_t)(word >> (8 * d));
            }
            if (a < 3) {
                Big t;
                big_mul_small(&sh, 256);
                big_mod(&t, &sh, &pl->qbig);
                sh = t;
            }
        }
    }
    return pl;
}

void orc_plan_free(OrcPlan* pl) {
    if (!pl) return;
    free(pl->weights);
    free(pl->b4_weight);
    free(pl->b7);
    free(pl->b4);
    free(pl);
}

uint32_t orc_plan_pcount(const OrcPlan* pl) { return pl->pcount; }
uint32_t orc_plan_kpad(const OrcPlan* pl) { return pl->kpad; }
uint32_t orc_plan_product_bits(const OrcPlan* pl) { return pl->product_bits; }

void orc_plan_tables(const OrcPlan* pl, uint32_t* primes, uint32_t* psi, uint32_t* omega,
                     uint32_t* omega_inv, uint32_t* n_inv, uint32_t* crt_inv) {
    size_t sz = sizeof(uint32_t) * pl->pcount;
    if (primes) memcpy(primes, pl->primes, sz);
    if (psi) memcpy(psi, pl->psi, sz);
    if (omega) memcpy(omega, pl->omega, sz);
    if (omega_inv) memcpy(omega_inv, pl->omega_inv, sz);
    if (n_inv) memcpy(n_inv, pl->n_inv, sz);
    if (crt_inv) memcpy(crt_inv, pl->crt_inv, sz);
}

void orc_plan_invp(const OrcPlan* pl, double* out) {
    memcpy(out, pl->invp, sizeof(double) * pl->pcount);
}

void orc_plan_weights(const OrcPlan* pl, uint32_t* out) {
    memcpy(out, pl->weights, sizeof(uint32_t) * (size_t)pl->terms * pl->l);
}

void orc_plan_b7(const OrcPlan* pl, uint8_t* out) {
    memcpy(out, pl->b7, (size_t)7 * pl->l * pl->kpad);
}

void orc_plan_b4(const OrcPlan* pl, uint8_t* out) {
    memcpy(out, pl->b4, (size_t)4 * pl->l * pl->kpad);
}

/* ---------------------------------------------------------------------- */
/* Encode: x mod p_j of an L-limb coefficient.                              */
/* ---------------------------------------------------------------------- */

/* Exact residue by Horner over limbs (the value the reference's encoders
 * produce: rns-encode.cuh:29-77, rns-encode-mma.cuh:40-52). */
void orc_encode(const OrcPlan* pl, const uint32_t* x, size_t count, uint32_t* out) {
    const uint32_t L = pl->l, P = pl->pcount;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)count; ++c) {
        const uint32_t* v = x + (size_t)c * L;
        for (uint32_t j = 0; j < P; ++j) {
            uint64_t r = 0, p = pl->primes[j];
            for (uint32_t i = L; i > 0; --i) r = ((r << 32) | v[i - 1]) % p;
            out[(size_t)j * count + (size_t)c] = (uint32_t)r;
        }
    }
}

/* The byte-power formulation of the tensor encoder: x = sum_k byte_k(x) *
 * (2^(8k) mod p), then Barrett with mu = floor((2^64-1)/p) and one
 * conditional subtraction (rns-encode-mma.cuh:15-36, 43-52).  Returns the
 * largest 4-byte dot-sum value seen (the < 2^47 bound pin). */
uint64_t orc_encode_bytepower(const OrcPlan* pl, const uint32_t* x, size_t count, uint32_t* out) {
    const uint32_t L = pl->l, P = pl->pcount, K = 4 * L;
    uint64_t largest = 0;
    uint32_t* pw = (uint32_t*)malloc(sizeof(uint32_t) * K);
    for (uint32_t j = 0; j < P; ++j) {
        uint32_t p = pl->primes[j];
        uint32_t power = 1;
        for (uint32_t k = 0; k < K; ++k) {
            pw[k] = power;
            power = (uint32_t)((uint64_t)power * 256 % p);
        }
        uint64_t mu = UINT64_MAX / p;
        for (size_t c = 0; c < count; ++c) {
            const uint8_t* bytes = (const uint8_t*)(x + c * L);
            uint64_t sums[4] = {0, 0, 0, 0};
            for (uint32_t k = 0; k < K; ++k)
                for (int d = 0; d < 4; ++d) sums[d] += (uint64_t)bytes[k] * ((pw[k] >> (8 * d)) & 255);
            uint64_t v = sums[0] + (sums[1] << 8) + (sums[2] << 16) + (sums[3] << 24);
            if (v > largest) largest = v;
            unsigned __int128 prod = (unsigned __int128)v * mu;
            uint64_t qhat = (uint64_t)(prod >> 64);
            uint64_t r = v - qhat * p;
            out[(size_t)j * count + c] = (uint32_t)(r >= p ? r - p : r);
        }
    }
    free(pw);
    return largest;
}

/* ---------------------------------------------------------------------- */
/* NTT: restates native-smoke/ntt-smoke.cu:88-129 (independent radix-2      */
/* DIT cyclic NTT with explicit bit reversal; psi twist and untwist).       */
/* ---------------------------------------------------------------------- */
static uint32_t modmul(uint32_t a, uint32_t b, uint32_t p) {
    return (uint32_t)((uint64_t)a * b % p);
}

static void cyclic_ntt(uint32_t* a, uint32_t n, uint32_t p, uint32_t root) {
    for (uint32_t i = 1, j = 0; i < n; ++i) {
        uint32_t bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) { uint32_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (uint32_t len = 2; len <= n; len <<= 1) {
        uint32_t step = pow_mod(root, n / len, p);
        for (uint32_t start = 0; start < n; start += len) {
            uint32_t w = 1;
            for (uint32_t j = 0; j < len / 2; ++j) {
                uint32_t u = a[start + j], v = modmul(a[start + j + len / 2], w, p);
                uint64_t s = (uint64_t)u + v;
                a[start + j] = (uint32_t)(s >= p ? s - p : s);
                a[start + j + len / 2] = u >= v ? u - v : (uint32_t)((uint64_t)u + p - v);
                w = modmul(w, step, p);
            }
        }
    }
}

/* c = a*b mod (X^n + 1, p) for canonical residues; psi of order 2n. */
void orc_ntt_negacyclic(uint32_t p, uint32_t psi, uint32_t n, const uint32_t* a_in,
                        const uint32_t* b_in, uint32_t* c) {
    uint32_t* a = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* b = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t omega = modmul(psi, psi, p);
    uint32_t omega_inv = pow_mod(omega, p - 2, p);
    uint32_t psi_inv = pow_mod(psi, p - 2, p);
    uint32_t n_inv = pow_mod(n % p, p - 2, p);
    uint32_t power = 1;
    for (uint32_t i = 0; i < n; ++i) {
        a[i] = modmul(a_in[i], power, p);
        b[i] = modmul(b_in[i], power, p);
        power = modmul(power, psi, p);
    }
    cyclic_ntt(a, n, p, omega);
    cyclic_ntt(b, n, p, omega);
    for (uint32_t i = 0; i < n; ++i) a[i] = modmul(a[i], b[i], p);
    cyclic_ntt(a, n, p, omega_inv);
    power = 1;
    for (uint32_t i = 0; i < n; ++i) {
        c[i] = modmul(modmul(a[i], n_inv, p), power, p);
        power = modmul(power, psi_inv, p);
    }
    free(a);
    free(b);
}

/* The ntt-smoke.cu:65-84 dense deterministic generator with boundary words. */
static uint64_t mix64(uint64_t x) {
    x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
    return x ^ (x >> 31);
}

void orc_dense_input(uint32_t n, uint32_t p, uint32_t operand, uint32_t* a) {
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t seed = ((uint64_t)p << 32) ^ (uint64_t)n ^
                        (UINT64_C(0x9e3779b97f4a7c15) * ((uint64_t)i + 1 + (uint64_t)operand * n));
        a[i] = (uint32_t)(mix64(seed) % p);
    }
    a[0] = operand ? p - 1 : 0;
    a[1] = operand ? 1 : p - 1;
    a[n / 2] = p / 2;
    a[n - 1] = operand ? p - 1 : 1;
}

/* ---------------------------------------------------------------------- */
/* CRT: centered alpha, byte-matrix sum, carry epilogue, reduction.         */
/* ---------------------------------------------------------------------- */

/* FP32 round-to-nearest alpha, as g1-crt.cuh:249-257 (each product and sum
 * rounded separately: this file is compiled with -ffp-contract=off). */
uint32_t orc_alpha_f32(const OrcPlan* pl, const uint32_t* terms) {
    float s = 0.0f;
    for (uint32_t j = 0; j < pl->pcount; ++j) {
        volatile float prod = (float)terms[j] * pl->invp_f32[j];
        s = s + prod;
    }
    return (uint32_t)lrintf(s);
}

/* FP64 alpha = round(sum t_j / p_j) (reference.py:109-112). */
uint32_t orc_alpha_f64(const OrcPlan* pl, const uint32_t* terms) {
    double s = 0.0;
    for (uint32_t j = 0; j < pl->pcount; ++j) s += (double)terms[j] * (1.0 / (double)pl->primes[j]);
    return (uint32_t)(s + 0.5);
}

/* B7 path: 7 byte-column dot sums per limb, then the carry epilogue of
 * reference.py:74-85.  Writes l+2 words of the wide sum z; returns the
 * largest dot value seen (reference-result.json largest_dot pin). */
uint32_t orc_crt_b7_wide(const OrcPlan* pl, const uint32_t* terms_alpha, uint32_t* z) {
    const uint32_t L = pl->l, K = pl->kpad;
    uint8_t a[4 * ORC_MAX_P + 64];
    memset(a, 0, sizeof(a));
    for (uint32_t j = 0; j < pl->terms; ++j)
        for (int d = 0; d < 4; ++d) a[4 * j + d] = (uint8_t)(terms_alpha[j] >> (8 * d));
    uint32_t largest = 0;
    uint64_t carry = 0;
    for (uint32_t limb = 0; limb < L; ++limb) {
        uint64_t c[7];
        for (int d = 0; d < 7; ++d) {
            const uint8_t* col = pl->b7 + (size_t)(7 * limb + d) * K;
            uint64_t s = 0;
            for (uint32_t k = 0; k < K; ++k) s += (uint64_t)a[k] * col[k];
            c[d] = s;
            if (s > largest) largest = (uint32_t)s;
        }
        uint64_t lo = c[0] + (c[1] << 8) + (c[2] << 16) + (c[3] << 24);
        uint64_t hi = c[4] + (c[5] << 8) + (c[6] << 16);
        uint64_t total = carry + lo;
        z[limb] = (uint32_t)total;
        carry = (total >> 32) + hi;
    }
    z[L] = (uint32_t)carry;
    z[L + 1] = (uint32_t)(carry >> 32);
    return largest;
}

/* B4 path: z = sum_{j,a} byte_a(term_j) * W[j,a] with W[j,a] = 256^a w_j mod q
 * (g1-crt.cuh:260-318).  Writes l+2 words. */
void orc_crt_b4_wide(const OrcPlan* pl, const uint32_t* terms_alpha, uint32_t* z) {
    const uint32_t L = pl->l;
    uint64_t acc[ORC_MAX_L + 2];
    memset(acc, 0, sizeof(acc));
    /* Column sums per (limb, byte d) then the 4-byte carry (g1-crt.cuh:304-318). */
    uint64_t carry = 0;
    for (uint32_t limb = 0; limb < L; ++limb) {
        uint64_t cs[4] = {0, 0, 0, 0};
        for (uint32_t j = 0; j < pl->terms; ++j)
            for (uint32_t a = 0; a < 4; ++a) {
                uint32_t byte = (terms_alpha[j] >> (8 * a)) & 255;
                uint32_t word = pl->b4_weight[((size_t)j * 4 + a) * L + limb];
                for (int d = 0; d < 4; ++d) cs[d] += (uint64_t)byte * ((word >> (8 * d)) & 255);
            }
        uint64_t lo = cs[0] + (cs[1] << 8) + (cs[2] << 16) + (cs[3] << 24);
        uint64_t total = carry + lo;
        z[limb] = (uint32_t)total;
        carry = total >> 32;
    }
    z[L] = (uint32_t)carry;
    z[L + 1] = (uint32_t)(carry >> 32);
    (void)acc;
}

/* Generic reduction: r = z mod q for z of l+2 words, by binary long division. */
void orc_reduce_generic(const OrcPlan* pl, const uint32_t* z, uint32_t* r) {
    Big zb, rb;
    big_set_words(&zb, z, (int)pl->l + 2);
    big_mod(&rb, &zb, &pl->qbig);
    memset(r, 0, sizeof(uint32_t) * pl->l);
    for (int i = 0; i < rb.n; ++i) r[i] = rb.w[i];
}

/* Near-power reduction for q = 2^W - gap (reference.py:88-95,
 * g1-crt.cuh:15-68).  Returns the number of folds, -1 if q is not of that
 * form. */
int orc_reduce_near_power(const OrcPlan* pl, const uint32_t* z, uint32_t* r) {
    const uint32_t W = pl->w, L = pl->l;
    Big qb = pl->qbig, pw, gap;
    big_set_u32(&pw, 1);
    for (uint32_t i = 0; i < W; ++i) big_shl1(&pw);
    if (big_cmp(&pw, &qb) <= 0) return -1;
    gap = pw;
    big_sub(&gap, &qb);
    if (gap.n > 1) return -1;
    uint32_t g = gap.n ? gap.w[0] : 0;
    if (g == 0) return -1;
    Big v;
    big_set_words(&v, z, (int)L + 2);
    int folds = 0;
    while (big_bitlen(&v) > W) {
        /* v = (v mod 2^W) + (v >> W) * gap */
        Big hi, lo;
        memset(&hi, 0, sizeof(hi));
        memset(&lo, 0, sizeof(lo));
        for (unsigned bit = 0; bit < big_bitlen(&v); ++bit) {
            uint32_t b = (v.w[bit / 32] >> (bit % 32)) & 1;
            if (bit < W) lo.w[bit / 32] |= b << (bit % 32);
            else hi.w[(bit - W) / 32] |= b << ((bit - W) % 32);
        }
        lo.n = BIG_WORDS - 1;
        hi.n = BIG_WORDS - 1;
        big_trim(&lo);
        big_trim(&hi);
        big_mul_small(&hi, g);
        /* lo += hi */
        uint64_t carry = 0;
        int n = lo.n > hi.n ? lo.n : hi.n;
        for (int i = 0; i < n; ++i) {
            uint64_t t = (uint64_t)(i < lo.n ? lo.w[i] : 0) + (i < hi.n ? hi.w[i] : 0) + carry;
            lo.w[i] = (uint32_t)t;
            carry = t >> 32;
        }
        lo.n = n;
        if (carry) lo.w[lo.n++] = (uint32_t)carry;
        big_trim(&lo);
        v = lo;
        ++folds;
    }
    if (big_cmp(&v, &qb) >= 0) big_sub(&v, &qb);
    memset(r, 0, sizeof(uint32_t) * L);
    for (int i = 0; i < v.n; ++i) r[i] = v.w[i];
    return folds;
}

/* CRT from terms: alpha (FP32 RN, as the production kernel), B4 wide sum,
 * generic reduction.  terms: count x pcount row-major; out: count x l. */
void orc_crt_terms(const OrcPlan* pl, const uint32_t* terms, size_t count, uint32_t* out) {
    const uint32_t P = pl->pcount, L = pl->l;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)count; ++c) {
        uint32_t ta[ORC_MAX_P + 1];
        uint32_t z[ORC_MAX_L + 2];
        memcpy(ta, terms + (size_t)c * P, sizeof(uint32_t) * P);
        ta[P] = orc_alpha_f32(pl, ta);
        orc_crt_b4_wide(pl, ta, z);
        orc_reduce_generic(pl, z, out + (size_t)c * L);
    }
}

/* ---------------------------------------------------------------------- */
/* Whole product: encode -> per-prime negacyclic NTT product -> CRT terms   */
/* t_j = c_j * (P/p_j)^-1 mod p_j -> alpha -> B4 sum -> mod q.              */
/* a, b, c: N x L little-endian u32 limbs, canonical [0,q).                 */
/* ---------------------------------------------------------------------- */
int orc_product(const OrcPlan* pl, const uint32_t* a, const uint32_t* b, uint32_t* c) {
    const uint32_t N = pl->n, P = pl->pcount, L = pl->l;
    uint32_t* ra = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * P);
    uint32_t* rb = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * P);
    uint32_t* rc = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N * P);
    if (!ra || !rb || !rc) { free(ra); free(rb); free(rc); return -1; }
    orc_encode(pl, a, N, ra);
    orc_encode(pl, b, N, rb);
#pragma omp parallel for schedule(dynamic, 1)
    for (long long j = 0; j < (long long)P; ++j)
        orc_ntt_negacyclic(pl->primes[j], pl->psi[j], N, ra + (size_t)j * N, rb + (size_t)j * N,
                           rc + (size_t)j * N);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)N; ++i) {
        uint32_t ta[ORC_MAX_P + 1];
        uint32_t z[ORC_MAX_L + 2];
        for (uint32_t j = 0; j < P; ++j)
            ta[j] = modmul(rc[(size_t)j * N + i], pl->crt_inv[j], pl->primes[j]);
        ta[P] = orc_alpha_f32(pl, ta);
        orc_crt_b4_wide(pl, ta, z);
        orc_reduce_generic(pl, z, c + (size_t)i * L);
    }
    free(ra);
    free(rb);
    free(rc);
    return 0;
}

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
