"""Host model of the CRT kernel's reduction steps against big integers (CPU).

The CRT kernel (bytecrt_kernels.hip, crt_* functions) forms per-limb sums
s_i = 2^24 C3 + 2^16 C2 + 2^8 C1 + C0 + corr_i of its int8 GEMM rows in FP64,
estimates Q = floor(z / q) from the top four limb sums before any carry, runs
one pass of exact FP64 limb steps t = s_i - Q q_i + carry, carry = floor(t /
2^32), digit = t - carry 2^32, and fixes a wrong estimate with one add or
subtract of q.  This test replays exactly those steps (Python floats are IEEE
doubles; the steps the kernel proves exact are done with Python ints and
checked to be representable) on GEMM rows built from the plan's real weight
digits and random term bytes, and compares the digits with z mod q.  It also
forces the estimate off by one both ways to exercise the correction.
"""
import numpy as np
import pytest

TWO32 = 4294967296.0


def limbs(x, L):
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)]


def reduce_like_kernel(s, q_limbs, L, dq=0):
    assert all(float(v) == v for v in s), "limb sum not exact in FP64"
    sd = [float(v) for v in s]
    b = L - 3
    zt = ((sd[L] * TWO32 + sd[L - 1]) * TWO32 + sd[L - 2]) * TWO32 + sd[b]
    qe = (float(q_limbs[L - 1]) * TWO32 + float(q_limbs[L - 2])) * TWO32 + float(q_limbs[b])
    Q = max(int(np.floor(zt / qe)) + dq, 0)
    carry, d = 0, []
    for i in range(L):
        t = s[i] - Q * q_limbs[i] + carry
        assert abs(s[i] - Q * q_limbs[i]) < 2**53 and abs(t) < 2**53  # the FMA and the add are exact
        c = t >> 32
        d.append(t - (c << 32))
        carry = c
    top = s[L] + carry
    sub, add = top > 0, top < 0
    if top == 0:
        sub = True
        for i in range(L - 1, -1, -1):
            if d[i] != q_limbs[i]:
                sub = d[i] > q_limbs[i]
                break
    fixed = sub or add
    if fixed:
        cy = 0
        for i in range(L):
            v = d[i] + (q_limbs[i] if add else -q_limbs[i]) + cy
            cy = v >> 32
            d[i] = v & 0xFFFFFFFF
    return d, fixed


def gemm_rows(pkg, N, W, q):
    pl = pkg.bytecrt.plan_host(N, W, q)
    L, P = pl["L"], pl["P"]
    wts = [sum(int(w) << (32 * i) for i, w in enumerate(pl["weights"][j])) for j in range(P + 1)]
    Wk = []
    for j in range(P + 1):
        x = wts[j]
        for _ in range(4):
            Wk.append(x)
            x = (x * 256) % q
    D = np.zeros((4 * L + 1, len(Wk)), dtype=np.int64)  # balanced base-256 digits of each W[j,a]
    for k, x in enumerate(Wk):
        carry = 0
        for m in range(4 * L):
            v = ((x >> (8 * m)) & 255) + carry
            D[m, k], carry = (v - 256, 1) if v >= 128 else (v, 0)
        D[4 * L, k] = carry
    corr = [0] * (L + 1)
    for m in range(4 * L + 1):
        corr[m // 4] += 128 * int(D[m].sum()) * (1 << (8 * (m % 4)))
    return L, Wk, D, corr


@pytest.mark.parametrize("N,W", [(65536, 1747), (32768, 868), (4096, 65), (8192, 97), (65536, 1955)])
def test_reduction_matches_big_integers(pkg, N, W):
    q = pkg.datagen.modulus(W, W + 7)
    L, Wk, D, corr = gemm_rows(pkg, N, W, q)
    ql = limbs(q, L)
    rng = np.random.default_rng(W)
    B = rng.integers(0, 256, size=(len(Wk), 24), dtype=np.int64)
    B[:, 0], B[:, 1] = 255, 0
    acc = D @ (B - 128)
    assert np.abs(acc).max() < 2**23
    for c in range(B.shape[1]):
        a = [int(v) for v in acc[:, c]] + [0, 0, 0]
        s = [a[4 * i] + 256 * a[4 * i + 1] + 65536 * a[4 * i + 2] + 16777216 * a[4 * i + 3] + corr[i]
             for i in range(L + 1)]
        assert max(abs(v) for v in s) < 2**49
        z = sum(si << (32 * i) for i, si in enumerate(s))
        assert z == sum(int(B[k, c]) * Wk[k] for k in range(len(Wk))) and 0 <= z < 2 ** (W + 18)
        for dq in (0, -1, 1):
            d, fixed = reduce_like_kernel(s, ql, L, dq)
            assert sum(v << (32 * i) for i, v in enumerate(d)) == z % q, (N, W, c, dq)
            if dq == 1 or (dq == -1 and z >= q):
                assert fixed
