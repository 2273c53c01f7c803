"""Host model of the kernels' FP64 modular product (bytecrt_kernels.hip,
mul_mod and balanced) at the input sizes the fused encode feeds it (CPU).

mul_mod(x, w) = (h - q p) + l with h = fl(x w), l = fma(x, w, -h),
q = fma(h, 1/p, 1.5 * 2^52) - 1.5 * 2^52.  Python floats are IEEE doubles; an
FMA is modelled exactly as the correctly rounded value of the exact rational
a b + c.  The test checks, for primes in (2^30, 2^31) and inputs up to 2^48
(the fused encode's unreduced dot sums are below 2^47, and eight forward
levels add at most 8 x 0.63 p), that the result is exactly x w mod p and at
most 0.63 p in magnitude, and that balanced() returns |r| <= p/2 + 1.
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
