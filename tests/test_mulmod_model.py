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
    (|x| + 1) k stays below 2^53 (capped at 2^39, above the kernels' largest
    input 192 p + 3 p / 32), on the plans' own primes; w * pinv is the
    kernels' wq."""
    rng = np.random.default_rng(55)
    for N, p_int in plan_primes():
        k = odd_part(p_int - 1)
        assert (p_int - 1) % (2 * N) == 0
        xmax = min((1 << 53) // k - 2, (1 << 39) - 1)
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


# ---------------------------------------------------------------------------
# The middle kernel's pair step (bytecrt_kernels.hip, mid_pair_step): the last
# forward level, the pointwise product and the first inverse level of a pair
# in one step, with w^2 taken from the level before's twiddle.

def brv(k: int, bits: int) -> int:
    return int(format(k, f"0{bits}b")[::-1], 2)


def test_twiddle_square_is_parent_twiddle():
    """z[k] = psi^brv(k) (log2 N-bit reversal) has z[2m + b]^2 = (-1)^b z[m]:
    checked for every k at N = 4,096 and on a sample at N = 65,536, with the
    oracle's psi."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import oracle
    rng = np.random.default_rng(7)
    for N, W, ks in ((4096, 100, None), (65536, 1747, 3000)):
        pl = oracle.Plan(N, W, (1 << (W - 1)) | 1)
        n = N.bit_length() - 1
        for j in (0, len(pl.primes) - 1):
            p, psi = int(pl.primes[j]), int(pl.psi[j])
            z = lambda k: pow(psi, brv(k, n), p)
            sample = range(2, N) if ks is None else (int(k) for k in rng.integers(2, N, size=ks))
            for k in sample:
                want = z(k >> 1) if k % 2 == 0 else (p - z(k >> 1)) % p
                assert z(k) * z(k) % p == want, (N, p, k)


def test_pair_step_equals_three_levels():
    """For the plan's primes and twiddles of the last level (k = N/2 + g), on
    inputs up to the pair step's bound 4.01 p: the FP64 model of the step
    returns exactly half of what the forward butterfly, the pointwise products
    and the inverse butterfly give (mod p), as its outputs (a0 b0 +/- w'
    a1 b1, (a0 + a1)(b0 + b1) - a0 b0 - a1 b1) with w' the parent twiddle and
    the sign (-1)^g; and |out0| < 2 m, |out1| < 3 m for m = p/2 + 2^-12 p."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import oracle
    rng = np.random.default_rng(71)
    N, W = 65536, 1747
    pl = oracle.Plan(N, W, (1 << (W - 1)) | 1)
    n = N.bit_length() - 1
    for j in (0, 57, len(pl.primes) - 1):
        p_int, psi = int(pl.primes[j]), int(pl.psi[j])
        p, pinv = float(p_int), 1.0 / float(p_int)
        half_inv = pow(2, p_int - 2, p_int)
        m = p_int / 2 + p_int * 2.0 ** -12
        lim = int(4.01 * p_int)
        for trial in range(120):
            g = int(rng.integers(0, N // 2))
            k = N // 2 + g
            w = pow(psi, brv(k, n), p_int)
            w_inv = pow(w, p_int - 2, p_int)
            parent = pow(psi, brv(k >> 1, n), p_int)
            if trial < 8:
                a0, a1, b0, b1 = [lim * (1 if (trial >> i) & 1 else -1) for i in range(3)] + [lim]
            else:
                a0, a1, b0, b1 = (int(v) for v in rng.integers(-lim, lim + 1, size=4))
            # the three levels, exactly
            fa0, fa1 = (a0 + w * a1) % p_int, (a0 - w * a1) % p_int
            fb0, fb1 = (b0 + w * b1) % p_int, (b0 - w * b1) % p_int
            c0, c1 = fa0 * fb0 % p_int, fa1 * fb1 % p_int
            want0, want1 = (c0 + c1) * half_inv % p_int, (c0 - c1) * w_inv * half_inv % p_int
            # the FP64 step
            c00 = mul_mod(float(a0), float(b0), p, pinv)
            c11 = mul_mod(float(a1), float(b1), p, pinv)
            cs = mul_mod(float(a0 + a1), float(b0 + b1), p, pinv)
            w2 = float(parent)
            t = mul_mod_tw(c11, w2, w2 * pinv, p)
            out0 = c00 - t if g & 1 else c00 + t
            out1 = cs - c00 - c11
            assert int(out0) % p_int == want0 and int(out1) % p_int == want1, (p_int, g, a0, a1, b0, b1)
            assert abs(out0) < 2 * m and abs(out1) < 3 * m
