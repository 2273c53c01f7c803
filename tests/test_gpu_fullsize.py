"""GPU parity at the headline size: B=128 products at N=65,536, W=1,747 (L=55,
P=114), device-resident, one hipGraph launch with the tuned pipelining -- the
exact configuration the benchmark times.  The oracle is too slow for all 128
products, so the batch is checked through properties that do not depend on
its size:

- planted closed forms: a = 0 gives 0, a = 1 gives b, a = X^k gives the
  signed negacyclic rotation of b;
- each sampled product equals its own B=1 run (other launch shapes: no
  pipelining, the single-product grids);
- one product (with q-1 coefficients planted) equals the oracle bit for bit;
- sampled coefficients of another equal the negacyclic convolution sum
  c_k = sum_{i<=k} a_i b_{k-i} - sum_{i>k} a_i b_{N+k-i} (mod q), computed
  directly with Python integers;
- a second replay of the graph gives the same batch.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gu():
    import gpu_util
    return gpu_util


def random_batch(B: int, N: int, L: int, W: int, seed: int) -> torch.Tensor:
    """[B][N][L] int32 limbs of values uniform in [0, 2^(W-1)) (below q, whose
    bit W-1 is set), generated on the device."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    x = torch.randint(0, 1 << 32, (B, N, L), generator=g, device="cuda", dtype=torch.int64)
    x[..., L - 1] &= (1 << (W - 1 - 32 * (L - 1))) - 1
    return torch.where(x >= (1 << 31), x - (1 << 32), x).to(torch.int32).contiguous()


def test_headline_batch_properties(pkg, orc, gu):
    N, W, B = 65536, 1747, 128
    dg = pkg.datagen
    q = dg.modulus(W, 2026)
    e = pkg.Engine(N, W, q, max_batch=B)
    L = e.L
    assert (L, e.P) == (55, 114)
    e.apply_tuning(B)

    a = random_batch(B, N, L, W, 501)
    b = random_batch(B, N, L, W, 502)
    qm1 = torch.from_numpy(dg.limbs(q - 1, L).view(np.int32)).cuda()
    a[:, ::97] = qm1  # the largest canonical coefficient, in every product
    b[:, ::89] = qm1
    k = 40000
    a[1].zero_()
    a[2].zero_()
    a[2, 0] = torch.from_numpy(dg.limbs(1, L).view(np.int32)).cuda()
    a[3].zero_()
    a[3, k] = torch.from_numpy(dg.limbs(1, L).view(np.int32)).cuda()
    torch.cuda.synchronize()

    c = gu.empty_dev((B, N, L))
    e.run_device(a, b, c, B, graph=True)
    c2 = gu.empty_dev((B, N, L))
    e.run_device(a, b, c2, B, graph=True)
    torch.cuda.synchronize()
    assert torch.equal(c, c2), "graph replays differ"

    host = lambda t: gu.from_dev(t).reshape(N, L)
    assert not host(c[1]).any()
    np.testing.assert_array_equal(host(c[2]), host(b[2]))
    b3 = orc.rows_to_ints(host(b[3]))
    got = orc.rows_to_ints(host(c[3]))
    assert got == [(-b3[i - k + N]) % q if i < k else b3[i - k] for i in range(N)]

    for i in (0, 64, 127):
        one = gu.empty_dev((N, L))
        e.run_device(a[i].contiguous(), b[i].contiguous(), one, 1)
        np.testing.assert_array_equal(host(one), host(c[i]), err_msg=f"product {i}: B=1 differs from the batch")

    np.testing.assert_array_equal(host(c[127]), orc.Plan(N, W, q).product(host(a[127]), host(b[127])))

    ai = orc.rows_to_ints(host(a[0]))
    bi = orc.rows_to_ints(host(b[0]))
    ci = orc.rows_to_ints(host(c[0]))
    for kk in (0, 1, 32767, 65535):
        s = sum(ai[i] * bi[kk - i] for i in range(kk + 1)) - sum(ai[i] * bi[N + kk - i] for i in range(kk + 1, N))
        assert ci[kk] == s % q, f"coefficient {kk}"


def _rotation(b_ints, k, q, N):
    """b * X^k mod (X^N + 1, q) for b given as Python integers."""
    return [(-b_ints[i - k + N]) % q if i < k else b_ints[i - k] for i in range(N)]


def test_uneven_batch_config(pkg, orc, gu):
    """Uneven batch at its stated shape: ONE hipGraph launch of a mix holding
    96 products at N=32,768, W=868 (57 primes) and 32 at N=65,536, W=1,747
    (114 primes), three seeded moduli per geometry, operands uniform in [0, q)
    with 10% edge coefficients (0, 1, q-1, floor(q/2), q-floor(q/2)).  Every
    product equals its own engine's B=1 run; each modulus group has a = 0,
    a = 1 and a = X^k planted (0, b and the signed rotation of b) and one
    product checked against the oracle (six oracle products)."""
    dg = pkg.datagen
    spec = [(32768, 868, (32, 32, 32)), (65536, 1747, (11, 11, 10))]
    groups, host = [], []
    for N, W, batches in spec:
        for m, B in enumerate(batches):
            q = dg.modulus(W, 4100 + 10 * m)
            e = pkg.Engine(N, W, q, max_batch=B)
            assert (e.L, e.P) == ((28, 57) if N == 32768 else (55, 114))
            base = 100000 * m + N
            a = np.stack([dg.operands(N, q, base + i, edge_fraction=0.1) for i in range(B)])
            b = np.stack([dg.operands(N, q, base + 500 + i, edge_fraction=0.1) for i in range(B)])
            k = N // 3 + 17 * m
            a[0] = 0
            a[1] = dg.constant(N, q, 1)
            a[2] = 0
            a[2, k] = dg.limbs(1, e.L)
            da, db = gu.to_dev(a).view(B, N, e.L), gu.to_dev(b).view(B, N, e.L)
            out = gu.empty_dev((B, N, e.L))
            groups.append((e, da, db, out, B))
            host.append((e, q, k, a, b, da, db, out))
    assert sum(g[4] for g in groups) == 128
    mix = pkg.Mix(groups)
    mix.run(graph=True)
    gu.sync()
    first = [out.clone() for *_, out in host]
    for gi, (e, q, k, a, b, da, db, out) in enumerate(host):
        B, N, L = a.shape[0], e.N, e.L
        alone = gu.empty_dev((B, N, L))
        for i in range(B):
            e.run_device(da[i], db[i], alone[i], 1)
        gu.sync()
        if not torch.equal(alone, out):
            bad = [i for i in range(B) if not torch.equal(alone[i], out[i])]
            raise AssertionError(f"group {gi} (N={N}): products {bad} differ from their B=1 runs")
        got = gu.from_dev(out).reshape(B, N, L)
        assert not got[0].any(), f"N={N}: a = 0 must give 0"
        np.testing.assert_array_equal(got[1], b[1], err_msg=f"N={N}: a = 1 must give b")
        assert orc.rows_to_ints(got[2]) == _rotation(orc.rows_to_ints(b[2]), k, q, N), f"N={N}: a = X^{k}"
        np.testing.assert_array_equal(got[3], orc.Plan(N, e.W, q).product(a[3], b[3]),
                                      err_msg=f"N={N}: product 3 differs from the oracle")
    # a second replay of the same graph gives the same batch
    for *_, out in host:
        out.zero_()
    gu.sync()
    mix.run(graph=True)
    gu.sync()
    for (*_, out), want in zip(host, first):
        assert torch.equal(out, want), "mix graph replays differ"
    # the same batch as persistent work-queue launches, classes on two streams
    mix.set_schedule(1, 2, pkg.Mix.PERSISTENT)
    for _ in range(2):
        for *_, out in host:
            out.zero_()
        gu.sync()
        mix.run(graph=True)
        gu.sync()
        for (*_, out), want in zip(host, first):
            assert torch.equal(out, want), "persistent mix differs"
    mix.close()


SERVING_GEOMS = [(8192, 220), (16384, 436), (32768, 868), (65536, 1747)]


def test_serving_poisson_config(pkg, orc, gu):
    """Serving at its stated shape: a Poisson request stream (exponential
    inter-arrival times at 400 requests/s) with N drawn uniformly from
    {8,192; 16,384; 32,768; 65,536} and W = {220; 436; 868; 1,747}, operands
    uniform in [0, q), into a server whose budget holds at most four N=65,536
    requests in its arena, so requests are admitted over many waves while
    others are on the GPU.  Every result equals the product computed alone on
    its own engine, and one per geometry equals the oracle."""
    import time
    dg = pkg.datagen
    rng = np.random.default_rng(np.random.SeedSequence([0x5E4E, 5]))
    n_req, qps = 48, 400.0
    geo = rng.integers(0, 4, size=n_req)
    for g in range(4):  # every geometry at least twice
        geo[2 * g], geo[2 * g + 1] = g, g
    rng.shuffle(geo)
    qs = {g: dg.modulus(SERVING_GEOMS[g][1], 5100 + g) for g in range(4)}
    reqs = []
    for i, g in enumerate(geo):
        N, W = SERVING_GEOMS[g]
        reqs.append((int(g), N, W, qs[g], dg.operands(N, qs[g], 7000 + i), dg.operands(N, qs[g], 8000 + i)))
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n_req))
    srv = pkg.Server(hbm_budget_bytes=560 << 20, max_wave=16)  # arena 420 MiB: four N=65,536 requests
    st0 = srv.stats()
    assert st0["arena_bytes"] < 5 * (3 * 65536 * 55 * 4 + 2 * 114 * 65536 * 4)
    tickets = []
    t0 = time.perf_counter()
    for (g, N, W, q, a, b), t in zip(reqs, arrivals):
        delay = t0 + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(N, W, q, a, b))
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    assert st["completed"] == n_req and st["waves"] >= 8, st
    assert st["plans"] == 4 and st["evictions"] == 0, st
    lat = {g: [] for g in range(4)}
    engines = {g: pkg.Engine(*SERVING_GEOMS[g], qs[g]) for g in range(4)}
    checked = set()
    for (g, N, W, q, a, b), (c, latency) in zip(reqs, results):
        lat[g].append(latency)
        e = engines[g]
        out = gu.empty_dev(N * e.L)
        e.run_device(gu.to_dev(a), gu.to_dev(b), out, 1)
        np.testing.assert_array_equal(c, gu.from_dev(out).reshape(N, e.L), err_msg=f"N={N}: served != alone")
        if g not in checked:
            np.testing.assert_array_equal(c, orc.Plan(N, W, q).product(a, b), err_msg=f"N={N}: served != oracle")
            checked.add(g)
    assert checked == {0, 1, 2, 3}
    for g in range(4):
        ms = np.array(lat[g]) * 1e3
        print(f"N={SERVING_GEOMS[g][0]}: {len(ms)} requests, p50 {np.percentile(ms, 50):.2f} ms, "
              f"p99 {np.percentile(ms, 99):.2f} ms")
    srv.close()


def test_server_table_budget_below_plans(pkg, orc, gu):
    """The server's plan tables count against its HBM budget: twelve plans
    (four moduli at each of N = 8,192, 16,384 and 32,768) whose tables
    together exceed the whole budget are served exactly, the table share is
    never exceeded, and idle plans are evicted (least recently used first)."""
    dg = pkg.datagen
    geoms = [(8192, 220), (16384, 436), (32768, 868)]
    budget = 64 << 20
    plans, reqs = {}, []
    for gi, (N, W) in enumerate(geoms):
        for m in range(4):
            q = dg.modulus(W, 5300 + 10 * gi + m)
            plans[(N, W, q)] = pkg.Engine(N, W, q)
    table_sum = sum(2 * e.P * e.N * 4 for e in plans.values())
    assert table_sum > budget
    keys = list(plans)
    for i in range(30):
        N, W, q = keys[(7 * i) % len(keys)]
        reqs.append((N, W, q, dg.operands(N, q, 9000 + i, edge_fraction=0.05), dg.operands(N, q, 9500 + i)))
    srv = pkg.Server(hbm_budget_bytes=budget, max_wave=8)
    tickets = [srv.submit(N, W, q, a, b) for N, W, q, a, b in reqs]
    results = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["evictions"] > 0 and st["table_bytes"] <= st["table_cap"] == budget // 4, st
    seen = set()
    for (N, W, q, a, b), c in zip(reqs, results):
        e = plans[(N, W, q)]
        out = gu.empty_dev(N * e.L)
        e.run_device(gu.to_dev(a), gu.to_dev(b), out, 1)
        np.testing.assert_array_equal(c, gu.from_dev(out).reshape(N, e.L), err_msg=f"N={N}")
        if N not in seen:
            np.testing.assert_array_equal(c, orc.Plan(N, W, q).product(a, b), err_msg=f"N={N} vs oracle")
            seen.add(N)
    # a plan whose tables alone exceed the table share is refused before it is built
    q = dg.modulus(1747, 5)
    big = np.zeros((65536, 55), dtype=np.uint32)
    with pytest.raises(ValueError, match="budget"):
        srv.submit(65536, 1747, q, big, big)
    srv.close()
