// Host RNS plan for the MI355X ByteCRT engine (see bytecrt_plan.hpp).
#include "bytecrt_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace bcrt {
namespace {

// ---- little-endian base-2^32 naturals -------------------------------------
using Nat = std::vector<uint32_t>;

void trim(Nat& a) {
    while (!a.empty() && a.back() == 0) a.pop_back();
}

unsigned bits(const Nat& a) {
    if (a.empty()) return 0;
    unsigned t = 0;
    for (uint32_t v = a.back(); v; v >>= 1) ++t;
    return 32u * unsigned(a.size() - 1) + t;
}

int cmp(const Nat& a, const Nat& b) {
    if (a.size() != b.size()) return a.size() < b.size() ? -1 : 1;
    for (size_t i = a.size(); i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

void sub_in(Nat& a, const Nat& b) {  // a -= b, a >= b
    uint64_t borrow = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const uint64_t s = (i < b.size() ? uint64_t(b[i]) : 0) + borrow;
        const uint64_t x = a[i];
        a[i] = uint32_t(x - s);
        borrow = x < s;
    }
    trim(a);
}

void mul_small_in(Nat& a, uint32_t m) {
    uint64_t c = 0;
    for (auto& w : a) {
        const uint64_t t = uint64_t(w) * m + c;
        w = uint32_t(t);
        c = t >> 32;
    }
    if (c) a.push_back(uint32_t(c));
    trim(a);
}

Nat mul(const Nat& a, const Nat& b) {
    Nat r(a.size() + b.size(), 0);
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t c = 0;
        for (size_t j = 0; j < b.size(); ++j) {
            const uint64_t t = uint64_t(a[i]) * b[j] + r[i + j] + c;
            r[i + j] = uint32_t(t);
            c = t >> 32;
        }
        r[i + b.size()] = uint32_t(c);
    }
    trim(r);
    return r;
}

Nat div_small(const Nat& a, uint32_t d, uint32_t& rem) {
    Nat q(a.size(), 0);
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) {
        const uint64_t num = (r << 32) | a[i];
        q[i] = uint32_t(num / d);
        r = num % d;
    }
    rem = uint32_t(r);
    trim(q);
    return q;
}

uint32_t mod_small(const Nat& a, uint32_t d) {
    uint64_t r = 0;
    for (size_t i = a.size(); i-- > 0;) r = ((r << 32) | a[i]) % d;
    return uint32_t(r);
}

Nat shl(const Nat& a, unsigned s) {
    if (a.empty()) return a;
    Nat r(a.size() + s / 32 + 1, 0);
    const unsigned ws = s / 32, bs = s % 32;
    for (size_t i = 0; i < a.size(); ++i) {
        const uint64_t v = uint64_t(a[i]) << bs;
        r[i + ws] |= uint32_t(v);
        r[i + ws + 1] |= uint32_t(v >> 32);
    }
    trim(r);
    return r;
}

// a mod m by restoring division over the quotient bits.
Nat mod(const Nat& a, const Nat& m) {
    Nat x = a;
    const unsigned la = bits(a), lm = bits(m);
    if (la < lm) return x;
    for (int s = int(la - lm); s >= 0; --s) {
        const Nat t = shl(m, unsigned(s));
        if (cmp(x, t) >= 0) sub_in(x, t);
    }
    return x;
}

// ---- 32-bit modular arithmetic ---------------------------------------------
uint32_t pow_mod(uint32_t b, uint64_t e, uint32_t m) {
    uint64_t r = 1, x = b % m;
    for (; e; e >>= 1) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
    }
    return uint32_t(r);
}

uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return uint32_t(uint64_t(a) * b % p); }

// Deterministic Miller-Rabin for 32-bit integers (bases 2, 7, 61).
bool is_prime32(uint32_t p) {
    if (p < 2) return false;
    for (uint32_t s : {2u, 3u, 5u, 7u, 11u, 13u, 17u}) {
        if (p == s) return true;
        if (p % s == 0) return false;
    }
    uint32_t d = p - 1, r = 0;
    while (!(d & 1)) { d >>= 1; ++r; }
    for (uint32_t a : {2u, 7u, 61u}) {
        uint32_t x = pow_mod(a, d, p);
        if (x == 1 || x == p - 1) continue;
        bool ok = false;
        for (uint32_t i = 1; i < r && !ok; ++i) {
            x = mulm(x, x, p);
            ok = x == p - 1;
        }
        if (!ok) return false;
    }
    return true;
}

uint32_t bitrev(uint32_t x, unsigned nbits) {
    uint32_t r = 0;
    for (unsigned i = 0; i < nbits; ++i) r |= ((x >> i) & 1u) << (nbits - 1 - i);
    return r;
}

// Balanced base-256 digits d_m in [-128, 127] of a non-negative L-word value:
// value = sum_m d_m 256^m, using 4L+1 digits.
std::vector<int8_t> balanced_digits(const uint32_t* words, uint32_t L) {
    std::vector<int8_t> d(4 * L + 1, 0);
    int carry = 0;
    for (uint32_t m = 0; m < 4 * L; ++m) {
        int v = int((words[m / 4] >> (8 * (m % 4))) & 255u) + carry;
        if (v >= 128) {
            d[m] = int8_t(v - 256);
            carry = 1;
        } else {
            d[m] = int8_t(v);
            carry = 0;
        }
    }
    d[4 * L] = int8_t(carry);
    return d;
}

uint32_t select_primes(uint32_t N, uint32_t W, std::vector<uint32_t>* primes, Nat* product) {
    unsigned n_bits = 0;
    for (uint32_t v = N; v; v >>= 1) ++n_bits;
    const unsigned guard = 2 * W + n_bits + 2;
    const uint32_t step = 2 * N;
    uint32_t p = ((uint32_t{1} << 31) - 2) / step * step + 1;
    Nat prod{1};
    uint32_t count = 0;
    while (bits(prod) < guard) {
        while (p > (uint32_t{1} << 30) && !is_prime32(p)) p -= step;
        if (p <= (uint32_t{1} << 30) || count >= kMaxPrimes)
            throw std::runtime_error("ByteCRT plan: prime search exhausted (W too large for N)");
        if (primes) primes->push_back(p);
        mul_small_in(prod, p);
        ++count;
        p -= step;
    }
    if (product) *product = prod;
    return count;
}

}  // namespace

uint32_t plan_prime_count(uint32_t N, uint32_t W) { return select_primes(N, W, nullptr, nullptr); }

uint32_t near_power_gap(uint32_t W, uint32_t L, const uint32_t* qw) {
    if (!qw || L < 3 || W < 64 || L != (W + 31) / 32) return 0;
    Nat q(qw, qw + L);
    trim(q);
    if (bits(q) != W) return 0;
    Nat gap = shl(Nat{1}, W);
    sub_in(gap, q);
    return bits(gap) <= 32 ? gap[0] : 0;  // gap >= 1: q < 2^W
}

HostPlan make_plan(uint32_t N, uint32_t W, uint32_t L, const uint32_t* qw) {
    if (N < 4096 || N > 65536 || (N & (N - 1)))
        throw std::invalid_argument("ByteCRT plan: N must be a power of two in [4096, 65536]");
    if (W < 64 || L != (W + 31) / 32 || L > kMaxLimbs)
        throw std::invalid_argument("ByteCRT plan: need 64 <= W <= 2048 and L = ceil(W/32)");
    if (!qw) throw std::invalid_argument("ByteCRT plan: q is null");
    HostPlan pl;
    Geometry& g = pl.g;
    g.N = N;
    g.W = W;
    g.L = L;
    pl.q.assign(qw, qw + L);
    Nat q(qw, qw + L);
    trim(q);
    if (bits(q) != W) throw std::invalid_argument("ByteCRT plan: bit_length(q) must equal W");

    Nat product;
    g.P = select_primes(N, W, &pl.primes, &product);
    g.product_bits = bits(product);
    {
        Nat qm1 = q;
        sub_in(qm1, Nat{1});
        Nat bound = mul(qm1, qm1);
        mul_small_in(bound, 4 * N);
        if (cmp(product, bound) <= 0) throw std::runtime_error("ByteCRT plan: centered CRT guard failed");
    }
    g.log_n = 0;
    while ((1u << g.log_n) < N) ++g.log_n;
    g.N2 = kBlockLen;
    g.N1 = N / g.N2;
    g.n1 = g.log_n - 8;
    g.Kenc = (4 * L + 31) / 32 * 32;
    g.PTenc = (g.P + 7) / 8;
    g.Kcrt = (4 * (g.P + 1) + 31) / 32 * 32;
    g.Mcrt = (4 * L + 1 + 31) / 32 * 32;
    const uint32_t P = g.P;

    pl.psi.resize(P);
    pl.crt_inverse.resize(P);
    pl.pd.resize(2 * P);
    pl.scale.resize(P);
    pl.scale_z1.resize(P);
    pl.zf16.resize(size_t(P) * 16);
    pl.invp.resize(P);
    pl.enc_off.resize(P);
    pl.enc_inv.resize(P);
    pl.weights.assign(size_t(P + 1) * L, 0);
    pl.zf.resize(size_t(P) * N);
    pl.zi.resize(size_t(P) * N);
    pl.enc_prime.assign(size_t(P) * 24, 0);
    pl.wenc.assign(size_t(g.PTenc) * 32 * g.Kenc, 0);

    std::vector<uint32_t> pw(N), pwi(N);
    pl.tw5_late = true;
    for (uint32_t j = 0; j < P; ++j) {
        const uint32_t p = pl.primes[j];
        // The middle kernel's last four inverse levels pass inputs up to
        // 192 p + 3 p / 32 to twiddle products whose quotient q times the odd
        // part k of p - 1 must stay below 2^53 (exact q (p - 1)).
        {
            uint32_t k = p - 1;
            while (!(k & 1)) k >>= 1;
            const unsigned __int128 qmax = uint64_t(192) * p + uint64_t(3) * p / 32 + 2;
            if (qmax * k >= (unsigned __int128)1 << 53) pl.tw5_late = false;
        }
        // Smallest base a >= 2 with (a^((p-1)/2N))^N == -1 (rns-plan.hpp:163-170).
        uint32_t psi = 0;
        for (uint32_t a = 2; a < p && !psi; ++a) {
            const uint32_t c = pow_mod(a, (p - 1) / (2 * N), p);
            if (pow_mod(c, N, p) == p - 1) psi = c;
        }
        if (!psi) throw std::runtime_error("ByteCRT plan: primitive root search failed");
        pl.psi[j] = psi;
        uint32_t rem = 0;
        const Nat base = div_small(product, p, rem);
        const uint32_t base_mod_p = mod_small(base, p);
        pl.crt_inverse[j] = pow_mod(base_mod_p, p - 2, p);
        const Nat wj = mod(base, q);
        for (size_t i = 0; i < wj.size(); ++i) pl.weights[size_t(j) * L + i] = wj[i];
        pl.invp[j] = 1.0 / double(p);

        pl.pd[2 * j] = double(p);
        pl.pd[2 * j + 1] = 1.0 / double(p);
        // The inverse transform leaves (N / 2) c (the middle kernel's pair
        // step, which merges the last forward level, the pointwise product
        // and the first inverse level, drops one factor 2); t_j = c crt_inverse_j.
        const uint32_t n_inv = pow_mod((N / 2) % p, p - 2, p);
        pl.scale[j] = double(mulm(n_inv, pl.crt_inverse[j], p));

        // Twiddles: zf[k] = psi^brv(k), zi[k] = psi^-brv(k) (log_n-bit reversal).
        const uint32_t psi_inv = pow_mod(psi, p - 2, p);
        uint32_t a = 1, b = 1;
        for (uint32_t e = 0; e < N; ++e) {
            pw[e] = a;
            pwi[e] = b;
            a = mulm(a, psi, p);
            b = mulm(b, psi_inv, p);
        }
        uint32_t* zf = pl.zf.data() + size_t(j) * N;
        uint32_t* zi = pl.zi.data() + size_t(j) * N;
        for (uint32_t k = 0; k < N; ++k) {
            const uint32_t e = bitrev(k, g.log_n);
            zf[k] = pw[e];
            zi[k] = pwi[e];
        }
        pl.scale_z1[j] = double(mulm(uint32_t(pl.scale[j]), zi[1], p));
        for (uint32_t k = 0; k < 16; ++k) pl.zf16[size_t(j) * 16 + k] = double(zf[k]);

        // Encode weights: r_k == 2^(8k) mod p in (-p/2, p/2], four balanced digits.
        // Row (j % 8) * 4 + d of tile j / 8; column k.
        int64_t rsum = 0;
        uint32_t power = 1;
        for (uint32_t k = 0; k < 4 * L; ++k) {
            const int64_t r = power > p / 2 ? int64_t(power) - int64_t(p) : int64_t(power);
            rsum += r;
            int64_t v = r;
            for (uint32_t d = 0; d < 4; ++d) {
                int64_t dig = ((v + 128) & 255) - 128;
                v = (v - dig) / 256;
                pl.wenc[(size_t(j / 8) * 32 + (j % 8) * 4 + d) * g.Kenc + k] = int8_t(dig);
            }
            if (v != 0) throw std::runtime_error("ByteCRT plan: encode digit overflow");
            power = uint32_t(uint64_t(power) * 256 % p);
        }
        // x == v + 128 * sum_k r_k (mod p) where v is the biased-byte dot sum,
        // |v| < 2^45; offset makes v + off non-negative and < 2^47.
        int64_t c = (128 * rsum) % int64_t(p);
        if (c < 0) c += p;
        const uint64_t m = ((uint64_t{1} << 45) + p - 1) / p;
        pl.enc_off[j] = uint64_t(c) + m * p;
        // (1/p)(1 - 2^-50) with both roundings stays below 1/p: u * enc_inv <= u/p.
        pl.enc_inv[j] = (1.0 / double(p)) * (1.0 - std::ldexp(1.0, -50));
        {
            const double offd = double(pl.enc_off[j]);  // < 2^47: exact
            const uint32_t off_lo = uint32_t(pl.enc_off[j]);
            uint8_t* rec = pl.enc_prime.data() + size_t(j) * 24;
            std::memcpy(rec, &offd, 8);
            std::memcpy(rec + 8, &pl.enc_inv[j], 8);
            std::memcpy(rec + 16, &p, 4);
            std::memcpy(rec + 20, &off_lo, 4);
        }
    }
    // Alpha weight (-P) mod q.
    {
        const Nat pm = mod(product, q);
        if (!pm.empty()) {
            Nat neg = q;
            sub_in(neg, pm);
            for (size_t i = 0; i < neg.size(); ++i) pl.weights[size_t(P) * L + i] = neg[i];
        }
    }
    // CRT weights: W[j,a] = 256^a * weight_j mod q (row k = 4j + a of the byte
    // matrix), stored transposed as balanced digits: wcrt[m][k] = digit_m(W[j,a]).
    pl.wcrt.assign(size_t(g.Mcrt) * g.Kcrt, 0);
    for (uint32_t j = 0; j <= P; ++j) {
        Nat sh(pl.weights.begin() + size_t(j) * L, pl.weights.begin() + size_t(j + 1) * L);
        trim(sh);
        for (uint32_t a = 0; a < 4; ++a) {
            std::vector<uint32_t> w(L, 0);
            for (size_t i = 0; i < sh.size(); ++i) w[i] = sh[i];
            const auto dig = balanced_digits(w.data(), L);
            for (uint32_t m = 0; m < 4 * L + 1; ++m) pl.wcrt[size_t(m) * g.Kcrt + 4 * j + a] = dig[m];
            if (a < 3) {
                mul_small_in(sh, 256);
                sh = mod(sh, q);
            }
        }
    }
    // Bias correction per digit row, folded per 4-digit limb.
    pl.corr.assign(g.Mcrt / 4, 0);
    for (uint32_t m = 0; m < g.Mcrt; ++m) {
        int64_t s = 0;
        for (uint32_t k = 0; k < g.Kcrt; ++k) s += pl.wcrt[size_t(m) * g.Kcrt + k];
        pl.corr[m / 4] += (128 * s) * (int64_t{1} << (8 * (m % 4)));
    }
    // Quotient estimate constants for the generic-q reduction.
    pl.qshift = W - 32;
    uint64_t qt = 0;
    {
        const unsigned a = pl.qshift / 32, bsh = pl.qshift % 32;
        uint64_t lo = pl.q[a];
        uint64_t hi = a + 1 < L ? pl.q[a + 1] : 0;
        qt = ((lo | (hi << 32)) >> bsh) & 0xffffffffu;
    }
    pl.qt1 = qt + 1;
    pl.inv_qt1 = 1.0 / double(qt + 1);
    pl.gap = near_power_gap(W, L, qw);
    return pl;
}

}  // namespace bcrt
