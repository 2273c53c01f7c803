"""Host model of the kernels' FP64 modular product (bytecrt_kernels.hip,
mul_mod and balanced) at the input sizes the fused encode feeds it (CPU).

mul_mod(x, w) = (h - q p) + l with h = fl(x w), l = fma(x, w, -h),
q = fma(h, 1/p, 1.5 * 2^52) - 1.5 * 2^52.  Python floats are IEEE doubles; an
FMA is modelled exactly as the correctly rounded value of the exact rational
a b + c.  The test checks, for primes in (2^30, 2^31) and inputs up to 2^48
(the fused encode's unreduced dot sums are below 2^47, and eight forward
levels add at most 8 x 0.63 p), that the result is exactly x w mod p and at
most 0.63 p in magnitude, and that balanced() returns |r| <= p/2 + 1.
The NTT kernels' five-operation twiddle product (mul_mod_tw) is modelled on
the plans' own primes within its bound |q| k < 2^53 (p - 1 = k 2^s), and shown
to fail past it.
"""
from fractions import Fraction

import numpy as np

K_ROUND = 6755399441055744.0  # 1.5 * 2^52


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def mul_mod(x: float, w: float, p: float, pinv: float) -> float:
    h = x * w
    l = fma(x, w, -h)
    q = fma(h, pinv, K_ROUND) - K_ROUND
    return fma(-q, p, h) + l


def balanced(x: float, p: float, pinv: float) -> int:
    q = fma(x, pinv, K_ROUND) - K_ROUND
    return int(fma(-q, p, x))


def test_mul_mod_exact_up_to_2_48():
    rng = np.random.default_rng(2048)
    primes = [2147352577, 1975386113, 2099838977, (1 << 30) + 3 * 2 ** 17 + 1]
    for p_int in primes:
        p, pinv = float(p_int), 1.0 / float(p_int)
        for _ in range(400):
            x_int = int(rng.integers(-(1 << 48) + 1, 1 << 48))
            w_int = int(rng.integers(0, p_int))
            r = mul_mod(float(x_int), float(w_int), p, pinv)
            assert r == int(r)
            assert (int(r) - x_int * w_int) % p_int == 0, (p_int, x_int, w_int)
            assert abs(r) <= 0.63 * p_int
            b = balanced(float(x_int), p, pinv)
            assert (b - x_int) % p_int == 0 and abs(b) <= p_int // 2 + 1
    # extremes of the range
    for p_int in primes:
        p, pinv = float(p_int), 1.0 / float(p_int)
        for x_int in ((1 << 48) - 1, -(1 << 48) + 1, (1 << 47) + 12345):
            for w_int in (p_int - 1, p_int // 2, 1):
                r = mul_mod(float(x_int), float(w_int), p, pinv)
                assert (int(r) - x_int * w_int) % p_int == 0 and abs(r) <= 0.63 * p_int


def canon_product(r: float, p: float) -> int:
    return int(r + p if r < 0.0 else r)


def test_canon_product_of_mul_mod():
    """The inverse columns write canon_product(mul_mod(x, s)): one conditional
    add of p, with no quotient step.  A mul_mod result is below p in
    magnitude (p/2 + 2^-12 p for |x| <= 2^40, the inverse columns' range,
    0.63 p up to 2^48), so the add gives the representative in [0, p)."""
    rng = np.random.default_rng(40)
    primes = [2147352577, 1975386113, (1 << 30) + 3 * 2 ** 17 + 1]
    for p_int in primes:
        p, pinv = float(p_int), 1.0 / float(p_int)
        xs = [int(v) for v in rng.integers(-(1 << 40) + 1, 1 << 40, size=300)]
        xs += [(1 << 40) - 1, -(1 << 40) + 1, 0, 1, -1, (1 << 47) + 7, -(1 << 47) - 7]
        for x_int in xs:
            for w_int in (int(rng.integers(0, p_int)), p_int - 1, 1, p_int // 2):
                c = canon_product(mul_mod(float(x_int), float(w_int), p, pinv), p)
                assert 0 <= c < p_int and c == (x_int * w_int) % p_int, (p_int, x_int, w_int)


# ---------------------------------------------------------------------------
# The NTT kernels' five-operation twiddle product (bytecrt_kernels.hip,
# mul_mod_tw): for p - 1 = k 2^s and w in [0, p), with wq = fl(w / p),
#   q = rint(x wq),  r = fma(x, w, -(q p - q)) - q  (q p - q by one FMA),
# exact while |q| k < 2^53.

def mul_mod_tw(x: float, w: float, wq: float, p: float) -> float:
    q = fma(x, wq, K_ROUND) - K_ROUND
    return fma(x, w, -fma(q, p, -q)) - q


def odd_part(n: int) -> int:
    while n % 2 == 0:
        n //= 2
    return n


def plan_primes():
    """The largest and smallest primes of the plans at N = 4,096 .. 65,536
    (p = 1 mod 2N: s >= 13 .. 17), from the oracle's plan."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import oracle
    out = []
    for N, W in ((4096, 100), (16384, 436), (32768, 868), (65536, 1747)):
        q = (1 << (W - 1)) | 1
        primes = [int(p) for p in oracle.Plan(N, W, q).primes]
        out += [(N, max(primes)), (N, min(primes))]
    return out


def test_mul_mod_tw_exact_within_its_bound():
    """Exact, with |r| <= p/2 + 2^-13 p, for every |x| whose quotient bound
    (|x| + 1) k stays below 2^53 (capped at 2^38, the kernels' largest input),
    on the plans' own primes; w * pinv is the kernels' wq."""
    rng = np.random.default_rng(55)
    for N, p_int in plan_primes():
        k = odd_part(p_int - 1)
        assert (p_int - 1) % (2 * N) == 0
        xmax = min((1 << 53) // k - 2, 1 << 38)
        p, pinv = float(p_int), 1.0 / float(p_int)
        xs = [int(v) for v in rng.integers(-xmax, xmax + 1, size=250)]
        xs += [xmax, -xmax, 0, 1, -1, p_int - 1, -(p_int - 1)]
        for x_int in xs:
            for w_int in (int(rng.integers(0, p_int)), p_int - 1, 1, 0, p_int // 2):
                w = float(w_int)
                r = mul_mod_tw(float(x_int), w, w * pinv, p)
                assert r == int(r)
                assert (int(r) - x_int * w_int) % p_int == 0, (N, p_int, x_int, w_int)
                assert abs(r) <= p_int / 2 + p_int * 2.0 ** -13


def test_mul_mod_tw_bound_is_needed():
    """Past the bound (|q| k >= 2^53) q (p - 1) is no longer exact in FP64 and
    the form fails on some inputs: the plan check and the per-kernel input
    bounds are what make it exact."""
    rng = np.random.default_rng(56)
    bad = 0
    for N, p_int in plan_primes():
        k = odd_part(p_int - 1)
        lo = (1 << 54) // k
        if lo > (1 << 45):
            continue
        p, pinv = float(p_int), 1.0 / float(p_int)
        for _ in range(200):
            x_int = int(rng.integers(lo, 4 * lo))
            w_int = int(rng.integers(p_int // 2, p_int))
            w = float(w_int)
            r = mul_mod_tw(float(x_int), w, w * pinv, p)
            if (int(r) - x_int * w_int) % p_int:
                bad += 1
    assert bad > 0
