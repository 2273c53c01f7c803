"""CPU tests that pin the oracle (oracle/bytecrt_oracle.c) before it is trusted.

Each check is chosen so a plausible mistake fails it: published pins of the
reference's plan, the reference's own plan header compiled in place
(oracle/_ref), the reference's six BCRT01 known-answer files, Python
big-integer arithmetic, schoolbook products and closed forms.
"""
import random
import struct

import numpy as np
import pytest

from conftest import GOLDEN


# ---------------------------------------------------------------------------
# Plan pins (reference-result.json:8-9,19-20; rns-plan-cpu-result.json:8-13;
# byte-crt-cpu-result.json:25-26,134 first/last primes)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N,W,P,kpad,bits,first,last", [
    (65536, 1747, 114, 480, 3528, 2147352577, 1975386113),
    (32768, 868, 57, 256, 1767, 2147352577, 2099838977),
])
def test_plan_pins(orc, N, W, P, kpad, bits, first, last):
    for gap in (1, 2 * N - 1, (1 << 32) - 1):
        pl = orc.Plan(N, W, (1 << W) - gap)
        assert (pl.P, pl.kpad, pl.product_bits) == (P, kpad, bits)
        assert pl.primes[0] == first and pl.primes[-1] == last
        assert np.all(np.diff(pl.primes.astype(np.int64)) < 0)
        for p, psi, om, omi, ni in zip(pl.primes, pl.psi, pl.omega, pl.omega_inv, pl.n_inv):
            p, psi = int(p), int(psi)
            assert (int(p) - 1) % (2 * N) == 0 and orc.prime32(p)
            assert pow(psi, N, p) == p - 1 and pow(psi, 2 * N, p) == 1
            assert pow(int(om), N // 2, p) == p - 1
            assert int(om) * int(omi) % p == 1 and N * int(ni) % p == 1


def test_prime32_matches_trial_division(orc):
    rng = random.Random(5)
    for _ in range(2000):
        n = rng.randrange(1 << 30, 1 << 31)
        d, is_p = 3, n % 2 == 1
        while is_p and d * d <= n:
            if n % d == 0:
                is_p = False
            d += 2
        assert orc.prime32(n) == is_p
    assert orc.prime32(2147483647) and not orc.prime32(3215031751)


def test_plan_crt_identities(orc):
    """Weights are P_j mod q and (-P) mod q; crt_inv (P_j)^-1 mod p_j (Python big ints)."""
    q = (1 << 868) - 65535
    pl = orc.Plan(32768, 868, q)
    P = 1
    for p in pl.primes:
        P *= int(p)
    w = pl.weights()
    for j, p in enumerate(pl.primes):
        Pj = P // int(p)
        assert orc.limbs_to_int(w[j]) == Pj % q
        assert int(pl.crt_inv[j]) * (Pj % int(p)) % int(p) == 1
    assert orc.limbs_to_int(w[-1]) == (-P) % q
    assert P > 4 * 32768 * (q - 1) ** 2


@pytest.mark.parametrize("N,W,gap", [(65536, 1747, 1), (65536, 1747, (1 << 32) - 1), (32768, 868, 65535)])
def test_plan_matches_reference_header(orc, N, W, gap):
    """oracle/_ref/ref_plan is native-smoke/rns-plan.hpp compiled where it lies."""
    ref = orc.ref_plan(N, W, gap)
    if ref is None:
        pytest.skip("reference checkout absent: oracle/_ref not built")
    pl = orc.Plan(N, W, (1 << W) - gap)
    assert (ref["P"], ref["kpad"], ref["L"]) == (pl.P, pl.kpad, pl.L)
    for k, mine in (("primes", pl.primes), ("psi", pl.psi), ("omega", pl.omega),
                    ("omega_inv", pl.omega_inv), ("n_inv", pl.n_inv), ("crt_inv", pl.crt_inv)):
        np.testing.assert_array_equal(ref[k], mine, err_msg=k)
    np.testing.assert_array_equal(ref["invp"], pl.invp)
    np.testing.assert_array_equal(ref["b7"], pl.b7())
    np.testing.assert_array_equal(ref["b4"], pl.b4())


# ---------------------------------------------------------------------------
# reference.py check_point restated: the seeded coefficient fixtures, the B7
# dot bound and fold counts (reference-result.json largest_dot / largest_folds)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N,W,largest", [(65536, 1747, 7423389), (32768, 868, 3824979)])
def test_b7_fixture_set_pins(orc, N, W, largest):
    seen_dot, seen_folds, checked = 0, 0, 0
    for gap in (1, 2 * N - 1, 2 * N * 124 - 1):
        q = (1 << W) - gap
        pl = orc.Plan(N, W, q)
        primes = [int(p) for p in pl.primes]
        P = 1
        for p in primes:
            P *= p
        bound = N * (q - 1) ** 2
        rng = random.Random(N + W + gap)
        for x in (0, 1, -1, bound, -bound, rng.randrange(-bound, bound + 1)):
            terms = [x % p * int(inv) % p for p, inv in zip(primes, pl.crt_inv)]
            approx = 0.0
            for t, p in zip(terms, primes):
                approx += float(t) * (1.0 / float(p))
            alpha = int(approx + 0.5)
            assert sum(t * (P // p) for t, p in zip(terms, primes)) - alpha * P == x
            ta = np.array(terms + [alpha], dtype=np.uint32)
            assert pl.alpha_f32(ta[:-1]) == alpha == pl.alpha_f64(ta[:-1])
            wide, dot = pl.crt_b7_wide(ta)
            w = pl.weights()
            assert wide == sum(int(t) * orc.limbs_to_int(wr) for t, wr in zip(ta, w))
            assert pl.crt_b4_wide(ta) % q == x % q
            red, folds = pl.reduce_near_power(wide)
            assert red == x % q == pl.reduce_generic(wide)
            seen_dot, seen_folds = max(seen_dot, dot), max(seen_folds, folds)
            checked += 1
    assert checked == 18 and seen_dot == largest and seen_folds == 2


# ---------------------------------------------------------------------------
# The reference's BCRT01 known-answer files (native-smoke/fixtures/*.bin)
# ---------------------------------------------------------------------------
def read_bcrt01(path):
    raw = path.read_bytes()
    assert raw[:8] == b"BCRT01\0\0"
    n, bits, limbs, P, kpad, m, cols, gap = struct.unpack("<8I", raw[8:40])
    off = 40
    q = np.frombuffer(raw[off:off + 4 * limbs], dtype="<u4"); off += 4 * limbs
    invp = np.frombuffer(raw[off:off + 8 * P], dtype="<f8"); off += 8 * P
    b7 = np.frombuffer(raw[off:off + cols * kpad], dtype=np.uint8).reshape(cols, kpad); off += cols * kpad
    a = np.frombuffer(raw[off:off + m * kpad], dtype=np.uint8).reshape(m, kpad); off += m * kpad
    golden = np.frombuffer(raw[off:off + 4 * m * limbs], dtype="<u4").reshape(m, limbs); off += 4 * m * limbs
    assert off == len(raw)
    terms = a[:, :4 * P].copy().view("<u4").reshape(m, P)
    return dict(n=n, bits=bits, L=limbs, P=P, kpad=kpad, m=m, gap=gap, q=q, invp=invp, b7=b7,
                terms=terms, golden=golden)


FIXTURES = sorted(GOLDEN.glob("g1-*.bin"))


def test_fixture_files_present():
    assert len(FIXTURES) == 6


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_bcrt01_fixture(orc, path):
    fx = read_bcrt01(path)
    q = orc.limbs_to_int(fx["q"])
    assert q == (1 << fx["bits"]) - fx["gap"]
    pl = orc.Plan(fx["n"], fx["bits"], q)
    assert (pl.P, pl.kpad) == (fx["P"], fx["kpad"])
    np.testing.assert_array_equal(pl.invp, fx["invp"])
    np.testing.assert_array_equal(pl.b7(), fx["b7"])
    out = pl.crt_terms(fx["terms"])
    np.testing.assert_array_equal(out, fx["golden"])
    for r in range(fx["m"]):
        ta = np.append(fx["terms"][r], pl.alpha_f32(fx["terms"][r])).astype(np.uint32)
        wide, _ = pl.crt_b7_wide(ta)
        red, folds = pl.reduce_near_power(wide)
        assert red == orc.limbs_to_int(fx["golden"][r]) and folds <= 2


# ---------------------------------------------------------------------------
# Encode
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("W", [100, 220, 868, 1747])
def test_encode_matches_python(orc, W):
    N = 4096
    q = (1 << W) - 159
    pl = orc.Plan(N, W, q)
    rng = random.Random(W)
    vals = [0, 1, q - 1, q // 2, (1 << (32 * pl.L)) - 1] + [rng.randrange(q) for _ in range(59)]
    x = orc.ints_to_rows(vals, pl.L)
    res = pl.encode(x)
    res2, largest = pl.encode_bytepower(x)
    np.testing.assert_array_equal(res, res2)
    for j, p in enumerate(pl.primes):
        assert [int(v) for v in res[j]] == [v % int(p) for v in vals]
    assert largest < 1 << 47


# ---------------------------------------------------------------------------
# NTT (ntt-smoke.cu protocol: two real primes, dense input, schoolbook)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("p", [2147352577, 2146959361])
@pytest.mark.parametrize("n", [16, 64])
def test_ntt_vs_schoolbook(orc, p, n):
    # smallest-base psi of order 2n
    psi = next(c for c in (pow(a, (p - 1) // (2 * n), p) for a in range(2, 1000)) if pow(c, n, p) == p - 1)
    a = orc.dense_input(n, p, 0)
    b = orc.dense_input(n, p, 1)
    got = orc.ntt_negacyclic(p, psi, a, b)
    want = orc.schoolbook_negacyclic([int(v) for v in a], [int(v) for v in b], p)
    assert [int(v) for v in got] == want


# ---------------------------------------------------------------------------
# Full product
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N,W,seed", [(16, 100, 1), (32, 130, 2), (64, 220, 3)])
def test_product_vs_schoolbook(orc, pkg, N, W, seed):
    q = pkg.datagen.modulus(W, seed)
    pl = orc.Plan(N, W, q)
    a = pkg.datagen.operands(N, q, seed, edges_first=True)
    b = pkg.datagen.operands(N, q, seed + 100)
    got = orc.rows_to_ints(pl.product(a, b))
    assert got == orc.schoolbook_negacyclic(orc.rows_to_ints(a), orc.rows_to_ints(b), q)


def test_product_closed_forms(orc, pkg):
    N, W = 4096, 220
    q = pkg.datagen.modulus(W, 9)
    pl = orc.Plan(N, W, q)
    b = pkg.datagen.operands(N, q, 4)
    bi = orc.rows_to_ints(b)
    zero = np.zeros_like(b)
    np.testing.assert_array_equal(pl.product(zero, b), zero)
    one = pkg.datagen.constant(N, q, 1)
    np.testing.assert_array_equal(pl.product(one, b), b)
    k = 37
    mono = np.zeros_like(b)
    mono[k] = pkg.datagen.limbs(1, pl.L)
    got = orc.rows_to_ints(pl.product(mono, b))
    want = [(-bi[i - k + N]) % q if i < k else bi[i - k] for i in range(N)]
    assert got == want
    m1 = pkg.datagen.constant(N, q, q - 1)
    res = orc.rows_to_ints(pl.product(m1, m1))
    assert res[0] == 1 and all(v == 0 for v in res[1:])


def test_product_sampled_coefficients_full_size(orc, pkg):
    """N=32768, W=868: a few c_k against Python big-integer sums."""
    N, W = 32768, 868
    q = pkg.datagen.modulus(W, 11)
    pl = orc.Plan(N, W, q)
    a = pkg.datagen.operands(N, q, 1, edges_first=True)
    b = pkg.datagen.operands(N, q, 2, edges_first=True)
    c = pl.product(a, b)
    ai, bi = orc.rows_to_ints(a), orc.rows_to_ints(b)
    for k in (0, 1, N // 2 + 3, N - 1):
        assert orc.limbs_to_int(c[k]) == orc.negacyclic_coefficient(ai, bi, k, q)
