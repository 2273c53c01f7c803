"""Uneven batches and serving: mixed-geometry launches (bcrt_mix_*) and the
continuous-batching server (bcrt_server_*) against per-engine results.

Every product of a mixed launch or a server wave must equal the same product
run alone on its own engine (whose parity with the oracle the other GPU tests
pin), and a sample is checked against the oracle directly.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (N, W, modulus seed): two moduli share a geometry so groups differ by q only
MIX = [(8192, 220, 1), (16384, 436, 2), (4096, 100, 3), (8192, 220, 4)]


@pytest.fixture(scope="module")
def gu():
    import gpu_util
    return gpu_util


def alone(pkg, gu, eng, a, b):
    out = gu.empty_dev(eng.N * eng.L)
    eng.run_device(gu.to_dev(a), gu.to_dev(b), out, 1)
    return gu.from_dev(out).reshape(eng.N, eng.L)


def test_mix_uneven_batches(pkg, orc, gu):
    batches = [3, 1, 2, 2]
    groups, host = [], []
    for (N, W, seed), B in zip(MIX, batches):
        q = pkg.datagen.modulus(W, 100 + seed)
        e = pkg.Engine(N, W, q, max_batch=B)
        a = np.stack([pkg.datagen.operands(N, q, 10 * seed + i, edge_fraction=0.05) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 10 * seed + 5 + i) for i in range(B)])
        out = gu.empty_dev(B * N * e.L)
        groups.append((e, gu.to_dev(a), gu.to_dev(b), out, B))
        host.append((e, q, a, b, out))
    mix = pkg.Mix(groups)
    mix.run()
    gu.sync()
    first = []
    for e, q, a, b, out in host:
        got = gu.from_dev(out).reshape(-1, e.N, e.L)
        for i in range(got.shape[0]):
            np.testing.assert_array_equal(got[i], alone(pkg, gu, e, a[i], b[i]), err_msg=f"N={e.N} i={i}")
        first.append(got)
    e, q, a, b, _ = host[1]
    np.testing.assert_array_equal(first[1][0], orc.Plan(e.N, e.W, q).product(a[0], b[0]))
    for _ in range(3):
        for *_, out in host:
            out.zero_()
        gu.sync()
        mix.run(graph=True)
        gu.sync()
        for (e, *_rest, out), want in zip(host, first):
            np.testing.assert_array_equal(gu.from_dev(out).reshape(-1, e.N, e.L), want)
    mix.close()


def test_mix_fused_encode_class(pkg, gu):
    """Config-4 shape in small: two N=65,536 groups under different moduli
    (their class launches the fused encode + forward-column kernel) beside an
    N=32,768 group (separate kernels); eager and graph runs equal each product
    computed alone."""
    spec = [(65536, 1747, 2, 501), (65536, 1747, 1, 502), (32768, 868, 2, 503)]
    groups, host = [], []
    for N, W, B, seed in spec:
        q = pkg.datagen.modulus(W, seed)
        e = pkg.Engine(N, W, q, max_batch=B)
        a = np.stack([pkg.datagen.operands(N, q, 10 * seed + i, edge_fraction=0.1) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 10 * seed + 5 + i) for i in range(B)])
        out = gu.empty_dev(B * N * e.L)
        groups.append((e, gu.to_dev(a), gu.to_dev(b), out, B))
        host.append((e, a, b, out))
    want = [np.stack([alone(pkg, gu, e, a[i], b[i]) for i in range(a.shape[0])]) for e, a, b, _ in host]
    mix = pkg.Mix(groups)
    for graph in (False, True, True):
        for *_, out in host:
            out.zero_()
        gu.sync()
        mix.run(graph=graph)
        gu.sync()
        for (e, _a, _b, out), w in zip(host, want):
            np.testing.assert_array_equal(gu.from_dev(out).reshape(-1, e.N, e.L), w, err_msg=f"N={e.N}")
    mix.close()


def test_mix_schedules(pkg, orc, gu):
    """26 products of three geometries under every schedule shape: one launch
    sequence per class on the caller's stream (the default), classes dealt
    into chunks on worker streams (each group split across chunks), chunks
    sharing fewer streams, every class merged into the same launches, and
    persistent work-queue launches (alone, chunked over streams, merged);
    eager and graph replays equal each product computed alone."""
    spec = [(8192, 220, 12), (4096, 100, 9), (16384, 436, 5)]
    groups, host = [], []
    for k, (N, W, B) in enumerate(spec):
        q = pkg.datagen.modulus(W, 300 + k)
        e = pkg.Engine(N, W, q, max_batch=B)
        a = np.stack([pkg.datagen.operands(N, q, 3000 + 50 * k + i, edge_fraction=0.1) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 4000 + 50 * k + i) for i in range(B)])
        out = gu.empty_dev(B * N * e.L)
        groups.append((e, gu.to_dev(a), gu.to_dev(b), out, B))
        host.append((e, q, a, b, out))
    mix = pkg.Mix(groups)
    assert mix.schedule == (1, 1, 0) and mix.chunks == 3  # one launch sequence per geometry class
    want = []
    for e, q, a, b, out in host:
        want.append(np.stack([alone(pkg, gu, e, a[i], b[i]) for i in range(a.shape[0])]))
    # (chunks per class, streams, flags) -> launch sequences: 5 products cap the N=16,384 class
    merge, pq = pkg.Mix.MERGE_CLASSES, pkg.Mix.PERSISTENT
    for sched, launches in (((1, 1, 0), 3), ((3, 3, 0), 9), ((4, 2, 0), 12), ((2, 1, 0), 6), ((1, 1, merge), 1),
                            ((3, 2, merge), 3), ((1, 1, pq), 3), ((2, 3, pq), 6), ((2, 2, merge | pq), 2)):
        mix.set_schedule(*sched)
        assert mix.chunks == launches, sched
        for graph in (False, True, True):
            for *_, out in host:
                out.zero_()
            gu.sync()
            mix.run(graph=graph)
            gu.sync()
            for (e, *_r, out), w in zip(host, want):
                np.testing.assert_array_equal(gu.from_dev(out).reshape(-1, e.N, e.L), w,
                                              err_msg=f"N={e.N} schedule={sched} graph={graph}")
    for bad in ((0, 1, 0), (17, 1, 0), (1, 0, 0), (1, 17, 0), (1, 1, 4)):
        with pytest.raises(ValueError):
            mix.set_schedule(*bad)
    e, q, a, b, _ = host[2]
    np.testing.assert_array_equal(want[2][4], orc.Plan(e.N, e.W, q).product(a[4], b[4]))
    mix.close()


def test_tune_mix_exact(pkg, gu):
    """The mix tuner times every candidate schedule, checks each one's
    products bit for bit and leaves the mix on the fastest; the tuned mix
    still equals each product computed alone."""
    from importlib import import_module
    tune = import_module("bytecrt_amd.tune")
    spec = [(8192, 220, 6, 311), (16384, 436, 3, 312), (8192, 220, 2, 313)]
    groups, host = [], []
    for N, W, B, seed in spec:
        q = pkg.datagen.modulus(W, seed)
        e = pkg.Engine(N, W, q, max_batch=B)
        a = np.stack([pkg.datagen.operands(N, q, 10 * seed + i, edge_fraction=0.1) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 10 * seed + 5 + i) for i in range(B)])
        out = gu.empty_dev(B * N * e.L)
        groups.append((e, gu.to_dev(a), gu.to_dev(b), out, B))
        host.append((e, a, b, out))
    mix = pkg.Mix(groups)
    res = tune.tune_mix(mix, [g[3] for g in groups], reps=2)
    assert len(res["rows"]) == len(tune.MIX_CANDIDATES)
    best = res["best"]
    assert mix.schedule == (best["chunks"], best["streams"], best["flags"])
    for *_, out in host:
        out.zero_()
    gu.sync()
    mix.run(graph=True)
    gu.sync()
    for e, a, b, out in host:
        got = gu.from_dev(out).reshape(-1, e.N, e.L)
        for i in range(got.shape[0]):
            np.testing.assert_array_equal(got[i], alone(pkg, gu, e, a[i], b[i]), err_msg=f"N={e.N} i={i}")
    mix.close()


def test_mix_rejects_bad_groups(pkg, gu):
    q = pkg.datagen.modulus(100, 1)
    e = pkg.Engine(4096, 100, q, max_batch=1)
    a = gu.empty_dev(4096 * e.L)
    with pytest.raises(ValueError):
        pkg.Mix([(e, a, a, a, 2)])  # batch above the engine's max_batch
    with pytest.raises(ValueError):
        pkg.Mix([(e, a, a, a, 1), (e, a, a, a, 1)])  # one engine twice


def test_server_alone_vs_mix(pkg, orc, gu):
    """Requests of three geometries submitted from several threads under a
    budget that holds only a few products per wave: each result equals the
    product computed alone."""
    cases = []
    engines = {}
    for k in range(18):
        N, W, seed = MIX[k % len(MIX)]
        q = pkg.datagen.modulus(W, 100 + seed)
        a = pkg.datagen.operands(N, q, 1000 + k, edge_fraction=0.05)
        b = pkg.datagen.operands(N, q, 2000 + k)
        cases.append((N, W, q, a, b))
        engines.setdefault((N, W, q), pkg.Engine(N, W, q))
    srv = pkg.Server(hbm_budget_bytes=16 << 20, max_wave=8)
    tickets = [None] * len(cases)

    def submit(idx):
        for i in idx:
            N, W, q, a, b = cases[i]
            tickets[i] = srv.submit(N, W, q, a, b)

    threads = [threading.Thread(target=submit, args=(range(t, len(cases), 3),)) for t in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    assert st["completed"] == len(cases)
    assert 1 <= st["plans"] <= len(engines) and st["table_bytes"] <= st["table_cap"]
    assert st["waves"] >= 3 and 1 <= st["max_wave"] <= 8
    for (N, W, q, a, b), (c, lat) in zip(cases, results):
        assert lat > 0
        np.testing.assert_array_equal(c, alone(pkg, gu, engines[(N, W, q)], a, b), err_msg=f"N={N} W={W}")
    N, W, q, a, b = cases[1]
    np.testing.assert_array_equal(results[1][0], orc.Plan(N, W, q).product(a, b))
    # a product larger than the whole budget is refused at submit
    q = pkg.datagen.modulus(1747, 5)
    big = np.zeros((65536, 55), dtype=np.uint32)
    with pytest.raises(ValueError, match="budget"):
        srv.submit(65536, 1747, q, big, big)
    # one request alone afterwards still matches
    N, W, q, a, b = cases[0]
    c, _ = srv.wait(srv.submit(N, W, q, a, b))
    np.testing.assert_array_equal(c, results[0][0])
    srv.close()


def test_mix_outlives_closed_engine(pkg, gu):
    """A mix holds its engines: closing an engine while a mix names it defers
    the engine's destruction, the mix keeps computing the same products
    (eager and graph), and destroying the mix then frees the engine."""
    spec = [(8192, 220, 3), (4096, 100, 2)]
    groups, host = [], []
    for N, W, B in spec:
        q = pkg.datagen.modulus(W, 900 + N % 11)
        e = pkg.Engine(N, W, q, max_batch=B)
        a = np.stack([pkg.datagen.operands(N, q, 40 + i) for i in range(B)])
        b = np.stack([pkg.datagen.operands(N, q, 50 + i) for i in range(B)])
        out = gu.empty_dev(B * N * e.L)
        groups.append((e, gu.to_dev(a), gu.to_dev(b), out, B))
        host.append((e, a, b, out))
    want = [np.stack([alone(pkg, gu, e, a[i], b[i]) for i in range(a.shape[0])]) for e, a, b, _ in host]
    mix = pkg.Mix(groups)
    for e, *_ in host:
        e.close()  # deferred: the mix still names the engine
    for graph in (False, True, True):
        for *_, out in host:
            out.zero_()
        gu.sync()
        mix.run(graph=graph)
        gu.sync()
        for (e, _a, _b, out), w in zip(host, want):
            np.testing.assert_array_equal(gu.from_dev(out).reshape(w.shape), w)
    mix.close()  # frees both engines now
