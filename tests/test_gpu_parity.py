"""GPU parity: every stage and the whole product against the oracle, bit-exact.

All calls go through the C ABI (libbytecrt.so via ctypes); torch only holds
device memory.
"""
import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle import read_bcrt01

pytestmark = pytest.mark.gpu

GEOMS = [(4096, 100), (8192, 220), (16384, 436), (32768, 868), (65536, 1747)]


@pytest.fixture(scope="module")
def gu():
    import gpu_util
    return gpu_util


def test_mfma_layout(pkg, gu):
    rng = np.random.default_rng(1)
    a = rng.integers(-128, 128, size=(32, 32), dtype=np.int8)
    b = rng.integers(-128, 128, size=(32, 32), dtype=np.int8)
    a[0, :] = -128
    b[0, :] = -128
    da, db = gu.to_dev(a), gu.to_dev(b)
    dc = gu.empty_dev((32, 32))
    pkg.bytecrt.mfma_probe(da, db, dc)
    got = gu.from_dev(dc, np.int32)
    want = a.astype(np.int32) @ b.astype(np.int32).T
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("N,W", GEOMS)
def test_encode_stage(pkg, orc, gu, N, W):
    q = pkg.datagen.modulus(W, N)
    e = pkg.Engine(N, W, q)
    a = pkg.datagen.operands(N, q, 1, edges_first=True)
    b = pkg.datagen.operands(N, q, 2, edge_fraction=0.1)
    a[-1] = (1 << 32) - 1  # all-ones limbs: the largest byte pattern
    planes = gu.empty_dev(e.planes_words(1))
    e.run_stages(pkg.bytecrt.STAGE_ENCODE, a=gu.to_dev(a), b=gu.to_dev(b), planes=planes)
    got = gu.from_dev(planes).reshape(2, e.P, N)
    pl = orc.Plan(N, W, q)
    for op, x in enumerate((a, b)):
        want = gu.to_planes(pl.encode(x), e.N1)
        np.testing.assert_array_equal(got[op], want, err_msg=f"operand {op}")


@pytest.mark.parametrize("tw6", [False, True], ids=["tw5", "tw6"])
@pytest.mark.parametrize("N,W", [(4096, 100), (16384, 436), (32768, 868), (65536, 1747)])
def test_ntt_stages_dense_protocol(pkg, orc, gu, N, W, tw6):
    """ntt-smoke.cu protocol: dense splitmix inputs with boundary words per prime;
    the three NTT stages leave t_j = (a*b mod (X^N+1, p_j)) (P/p_j)^-1 mod p_j.
    Both twiddle-product forms: the five-operation one (the default; at N <=
    32,768 the middle kernel's last four inverse levels keep mul_mod, its
    plan check fails there) and mul_mod throughout (FLAG_MULMOD_TW6)."""
    q = pkg.datagen.modulus(W, 3)
    e = pkg.Engine(N, W, q)
    if tw6:
        e.set_kernel_flags(e.FLAG_MULMOD_TW6)
    pl = orc.Plan(N, W, q)
    nat = np.zeros((2, e.P, N), dtype=np.uint32)
    want = np.zeros((e.P, N), dtype=np.uint32)
    for j, p in enumerate(pl.primes):
        p = int(p)
        nat[0, j] = orc.dense_input(N, p, 0)
        nat[1, j] = orc.dense_input(N, p, 1)
        c = orc.ntt_negacyclic(p, int(pl.psi[j]), nat[0, j], nat[1, j]).astype(np.uint64)
        want[j] = (c * int(pl.crt_inv[j]) % p).astype(np.uint32)
    planes = gu.to_dev(gu.to_planes(nat, e.N1).reshape(-1))
    e.run_stages(pkg.bytecrt.STAGE_FWD_COLS | pkg.bytecrt.STAGE_MIDDLE | pkg.bytecrt.STAGE_INV_COLS,
                 planes=planes)
    got = gu.from_planes(gu.from_dev(planes).reshape(2, e.P, N)[0], e.N1)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} mismatches, first (prime, coef) = {bad[0].tolist()}"


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("g1-*.bin")), ids=lambda p: p.stem)
def test_crt_stage_bcrt01(pkg, orc, gu, path):
    fx = read_bcrt01(path)
    N, W = fx["n"], fx["bits"]
    q = orc.limbs_to_int(fx["q"])
    e = pkg.Engine(N, W, q)
    assert e.P == fx["P"] and e.Kcrt == fx["kpad"]
    pos = np.zeros((2, e.P, N), dtype=np.uint32)
    pos[0, :, :fx["m"]] = fx["terms"].T  # plane positions f = 0..15
    planes = gu.to_dev(pos.reshape(-1))
    nat = gu.natural_of_position(N, e.N1)
    rest = np.ones(N, dtype=bool)
    rest[nat[:fx["m"]]] = False
    # every fixture modulus is 2^W - gap: the near-power reduction, then the
    # generic-q steps on the same terms
    for flags in (0, e.FLAG_GENERIC_Q):
        e.set_kernel_flags(flags)
        out = gu.empty_dev(N * e.L)
        e.run_stages(pkg.bytecrt.STAGE_CRT, c=out, planes=planes)
        got = gu.from_dev(out).reshape(N, e.L)
        np.testing.assert_array_equal(got[nat[:fx["m"]]], fx["golden"])
        assert not got[rest].any()


@pytest.mark.parametrize("N,W,gap", [(8192, 220, 1), (8192, 220, (1 << 32) - 1), (32768, 868, 2 * 32768 - 1),
                                     (65536, 1747, 0x9E3779B9)])
def test_crt_stage_near_power_against_generic(pkg, orc, gu, N, W, gap):
    """CRT of random canonical terms t_j < p_j under q = 2^W - gap, with the
    extremes (all t_j = 0, all t_j = p_j - 1, one term nonzero) planted: the
    near-power reduction equals the generic-q steps everywhere, and the
    oracle's CRT on sampled positions."""
    q = (1 << W) - gap
    e = pkg.Engine(N, W, q)
    primes = e.plan_tables()["primes"].astype(np.uint64)
    rng = np.random.default_rng(gap % 1000003)
    terms = (rng.integers(0, 1 << 62, size=(N, e.P), dtype=np.uint64) % primes).astype(np.uint32)
    terms[0] = 0
    terms[1] = (primes - 1).astype(np.uint32)
    terms[2] = 0
    terms[2, e.P // 2] = 1
    pos = np.zeros((2, e.P, N), dtype=np.uint32)
    pos[0] = terms.T  # plane position f holds row f
    planes = gu.to_dev(pos.reshape(-1))
    outs = []
    for flags in (0, e.FLAG_GENERIC_Q):
        e.set_kernel_flags(flags)
        out = gu.empty_dev(N * e.L)
        e.run_stages(pkg.bytecrt.STAGE_CRT, c=out, planes=planes)
        outs.append(gu.from_dev(out).reshape(N, e.L))
    np.testing.assert_array_equal(outs[0], outs[1])
    nat = gu.natural_of_position(N, e.N1)
    sample = np.concatenate([np.arange(3), rng.choice(np.arange(3, N), size=253, replace=False)])
    want = orc.Plan(N, W, q).crt_terms(terms[sample])
    np.testing.assert_array_equal(outs[0][nat[sample]], want)


def product_case(pkg, orc, gu, N, W, q, seed, edge_fraction=0.0):
    e = pkg.Engine(N, W, q)
    a = pkg.datagen.operands(N, q, seed, edges_first=True, edge_fraction=edge_fraction)
    b = pkg.datagen.operands(N, q, seed + 1, edge_fraction=edge_fraction)
    out = gu.empty_dev(N * e.L)
    e.run_device(gu.to_dev(a), gu.to_dev(b), out, 1)
    got = gu.from_dev(out).reshape(N, e.L)
    want = orc.Plan(N, W, q).product(a, b)
    bad = np.argwhere((got != want).any(axis=1))
    assert bad.size == 0, f"{len(bad)} coefficient mismatches, first {bad[:4].ravel().tolist()}"
    return a, b, got


@pytest.mark.parametrize("N,W", GEOMS)
def test_product_random_modulus(pkg, orc, gu, N, W):
    product_case(pkg, orc, gu, N, W, pkg.datagen.modulus(W, 17), seed=5, edge_fraction=0.1)


@pytest.mark.parametrize("N,W", [(4096, 1000), (8192, 500), (16384, 129), (16384, 1200),
                                 (32768, 1500), (65536, 640)])
def test_batch_off_diagonal_geometries(pkg, orc, gu, N, W):
    """Geometries away from the configs' W ~ N / 37.5 line (other encode and
    CRT k-step counts, fused and separate first stages): a batch of two
    through the device path, each product against the oracle."""
    B = 2
    q = pkg.datagen.modulus(W, 31)
    e = pkg.Engine(N, W, q, max_batch=B)
    a = np.stack([pkg.datagen.operands(N, q, 1400 + i, edges_first=(i == 0), edge_fraction=0.1) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 1500 + i, edge_fraction=0.1) for i in range(B)])
    out = gu.empty_dev(B * N * e.L)
    e.run_device(gu.to_dev(a), gu.to_dev(b), out, B)
    got = gu.from_dev(out).reshape(B, N, e.L)
    plan = orc.Plan(N, W, q)
    for i in range(B):
        want = plan.product(a[i], b[i])
        bad = np.argwhere((got[i] != want).any(axis=1))
        assert bad.size == 0, f"product {i}: {len(bad)} mismatches, first {bad[:4].ravel().tolist()}"


@pytest.mark.parametrize("N,W", [(8192, 220), (32768, 868), (65536, 1747)])
def test_twiddle_forms_agree(pkg, orc, gu, N, W):
    """The NTT kernels' five-operation twiddle products (the default) and the
    six-operation TwoProduct form (FLAG_MULMOD_TW6) give the same products
    (a batch of two with 10% edge coefficients and all-(q-1) rows planted),
    both equal to the oracle."""
    B = 2
    q = pkg.datagen.modulus(W, 41)
    e = pkg.Engine(N, W, q, max_batch=B)
    a = np.stack([pkg.datagen.operands(N, q, 1600 + i, edges_first=(i == 0), edge_fraction=0.1) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 1700 + i, edge_fraction=0.1) for i in range(B)])
    qm1 = orc.int_to_limbs(q - 1, e.L)
    a[1, :64] = qm1
    b[1, :64] = qm1
    da, db = gu.to_dev(a), gu.to_dev(b)
    got = {}
    for flags in (0, e.FLAG_MULMOD_TW6):
        e.set_kernel_flags(flags)
        out = gu.empty_dev(B * N * e.L)
        e.run_device(da, db, out, B)
        got[flags] = gu.from_dev(out).reshape(B, N, e.L)
    np.testing.assert_array_equal(got[0], got[e.FLAG_MULMOD_TW6])
    plan = orc.Plan(N, W, q)
    for i in range(B):
        want = plan.product(a[i], b[i])
        bad = np.argwhere((got[0][i] != want).any(axis=1))
        assert bad.size == 0, f"product {i}: {len(bad)} mismatches, first {bad[:4].ravel().tolist()}"


@pytest.mark.parametrize("N,W", [(32768, 70), (65536, 100), (65536, 1955)])
def test_product_extreme_widths(pkg, orc, gu, N, W):
    """Large N with the narrowest and widest moduli: few primes (a single
    encode weight tile, one prime pair in the fused kernel) and the most
    (127 primes, the widest CRT GEMM)."""
    product_case(pkg, orc, gu, N, W, pkg.datagen.modulus(W, 23), seed=7, edge_fraction=0.1)


@pytest.mark.parametrize("gap", [1, (1 << 32) - 1])
def test_product_near_power_modulus(pkg, orc, gu, gap):
    product_case(pkg, orc, gu, 65536, 1747, (1 << 1747) - gap, seed=9)


@pytest.mark.parametrize("N,W", [(16384, 436), (65536, 1747)])
def test_product_closed_forms(pkg, orc, gu, N, W):
    """Products whose result has a closed form.  At N=65,536 the single product
    takes the fused encode (512 workgroups) and its partly filled last group of
    16 primes (P=114)."""
    q = pkg.datagen.modulus(W, 21)
    e = pkg.Engine(N, W, q)
    dg = pkg.datagen
    b = dg.operands(N, q, 3)
    bi = orc.rows_to_ints(b)

    def run(a, bb):
        out = gu.empty_dev(N * e.L)
        e.run_device(gu.to_dev(a), gu.to_dev(bb), out, 1)
        return gu.from_dev(out).reshape(N, e.L)

    zero = np.zeros_like(b)
    np.testing.assert_array_equal(run(zero, b), zero)
    np.testing.assert_array_equal(run(dg.constant(N, q, 1), b), b)
    k = 1234
    mono = np.zeros_like(b)
    mono[k] = dg.limbs(1, e.L)
    got = orc.rows_to_ints(run(mono, b))
    assert got == [(-bi[i - k + N]) % q if i < k else bi[i - k] for i in range(N)]
    m1 = dg.constant(N, q, q - 1)
    res = orc.rows_to_ints(run(m1, m1))
    assert res[0] == 1 and not any(res[1:])
    # all coefficients q-1 times all q-1: c_k = (k+1) - (N-1-k) = 2k + 2 - N (mod q)
    f1 = dg.filled(N, q, q - 1)
    res = orc.rows_to_ints(run(f1, f1))
    assert res == [(2 * k + 2 - N) % q for k in range(N)]
    # the largest canonical magnitudes: h = floor(q/2) as a constant and dense
    h = q // 2
    hc = dg.constant(N, q, h)
    res = orc.rows_to_ints(run(hc, hc))
    assert res[0] == h * h % q and not any(res[1:])
    fh = dg.filled(N, q, h)
    res = orc.rows_to_ints(run(fh, fh))
    assert res == [h * h * (2 * k + 2 - N) % q for k in range(N)]


def test_batch_graph_and_whole_call(pkg, orc, gu):
    """Each product of a batch equals its B=1 result; graph replays equal eager;
    the host whole-call path equals the device path."""
    N, W, B = 32768, 868, 4
    q = pkg.datagen.modulus(W, 33)
    e = pkg.Engine(N, W, q, max_batch=B)
    a = np.stack([pkg.datagen.operands(N, q, 40 + i, edge_fraction=0.05) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 80 + i) for i in range(B)])
    da, db = gu.to_dev(a), gu.to_dev(b)
    out = gu.empty_dev(B * N * e.L)
    e.run_device(da, db, out, B)
    batched = gu.from_dev(out).reshape(B, N, e.L)
    for i in range(B):
        one = gu.empty_dev(N * e.L)
        e.run_device(gu.to_dev(a[i]), gu.to_dev(b[i]), one, 1)
        np.testing.assert_array_equal(gu.from_dev(one).reshape(N, e.L), batched[i])
    np.testing.assert_array_equal(batched[1], orc.Plan(N, W, q).product(a[1], b[1]))
    for _ in range(3):
        out.zero_()
        e.run_device(da, db, out, B, graph=True)
        np.testing.assert_array_equal(gu.from_dev(out).reshape(B, N, e.L), batched)
    whole = e.run(a, b)
    np.testing.assert_array_equal(whole, batched)
    # pipelined: chunks of one product over three worker streams, eager and graph
    e.set_pipeline(1, 3)
    for graph in (False, True, True):
        out.zero_()
        e.run_device(da, db, out, B, graph=graph)
        np.testing.assert_array_equal(gu.from_dev(out).reshape(B, N, e.L), batched)
    e.set_pipeline(3, 2)
    out.zero_()
    e.run_device(da, db, out, B, graph=True)
    np.testing.assert_array_equal(gu.from_dev(out).reshape(B, N, e.L), batched)


@pytest.mark.skip(reason="pipeline tuner switched off: the test process died with SIGSEGV on the host (no GPU "
                         "fault) inside the graph capture after a candidate's eager run, in one run out of two "
                         "after a fix of event reuse; cause not established, so the case stays off the GPU")
def test_tune_pipeline_headline(pkg, gu):
    """The pipeline tuner at the benchmark's geometry (B=128, N=65,536,
    W=1,747): every candidate is captured right after its eager run, every
    candidate's batch equals the first bit for bit, the engine is left on the
    fastest, and graph replays after the tuning give the same batch."""
    import importlib
    import torch
    tune = importlib.import_module("bytecrt_amd.tune")
    N, W, B = 65536, 1747, 128
    q = pkg.datagen.modulus(W, 55)
    e = pkg.Engine(N, W, q, max_batch=B)
    g = torch.Generator(device="cuda")
    g.manual_seed(55)
    qtop = q >> (32 * (e.L - 1))
    x = torch.randint(0, 1 << 32, (2, B, N, e.L), generator=g, device="cuda", dtype=torch.int64)
    x[..., e.L - 1] %= qtop  # below q: the top limb stays below q's
    x = torch.where(x >= (1 << 31), x - (1 << 32), x).to(torch.int32)
    da, db = x[0].contiguous(), x[1].contiguous()
    del x
    out = torch.empty_like(da)
    res = tune.tune_pipeline(e, da, db, out, B, candidates=((0, 1), (32, 2), (16, 4)), reps=2)
    assert {(r["chunk"], r["streams"]) for r in res["rows"]} == {(128, 1), (32, 2), (16, 4)}
    assert e.pipeline == (res["best"]["chunk"], res["best"]["streams"])
    want = out.clone()
    for _ in range(2):
        out.zero_()
        e.run_device(da, db, out, B, graph=True)
        gu.sync()
        assert torch.equal(out, want)
    one = torch.empty_like(out[0])
    e.run_device(da[77].contiguous(), db[77].contiguous(), one, 1)
    gu.sync()
    assert torch.equal(one, want[77])


@pytest.mark.skip(reason="pipeline tuner switched off: the test process died with SIGSEGV on the host (no GPU "
                         "fault) inside the graph capture after a candidate's eager run, in one run out of two "
                         "after a fix of event reuse; cause not established, so the case stays off the GPU")
def test_tune_pipeline_exact(pkg, gu, tmp_path):
    """The autotuner times its candidates through the graph path, keeps every
    result bit-exact, leaves the engine on the fastest, and a fresh engine
    that applies the recorded table computes the same products."""
    import importlib
    import torch
    tune = importlib.import_module("bytecrt_amd.tune")
    N, W, B = 32768, 868, 12
    q = pkg.datagen.modulus(W, 77)
    a = np.stack([pkg.datagen.operands(N, q, 700 + i) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 800 + i) for i in range(B)])
    e = pkg.Engine(N, W, q, max_batch=B)
    da, db = gu.to_dev(a).view(B, N, -1), gu.to_dev(b).view(B, N, -1)
    out = torch.empty_like(da)
    e.set_pipeline(B, 1)
    e.run_device(da, db, out, B)
    gu.sync()
    plain = out.clone()
    res = tune.tune_pipeline(e, da, db, out, B, candidates=((0, 1), (8, 2), (4, 4), (4, 2)), reps=2)
    assert {(r["chunk"], r["streams"]) for r in res["rows"]} == {(12, 1), (8, 2), (4, 4), (4, 2)}
    assert e.pipeline == (res["best"]["chunk"], res["best"]["streams"])
    assert torch.equal(out, plain)
    path = tmp_path / "table.json"
    tune.record(N, W, B, res, device="test", path=path)
    f = pkg.Engine(N, W, q, max_batch=B)
    assert f.apply_tuning(B, tune.load_table(path)) and f.pipeline == e.pipeline
    out2 = torch.zeros_like(out)
    f.run_device(da, db, out2, B, graph=True)
    gu.sync()
    assert torch.equal(out2, plain)


@pytest.mark.parametrize("N,W", [(65536, 1747), (65536, 1955), (65536, 300), (32768, 868), (16384, 436)])
def test_fused_encode_forward(pkg, gu, N, W):
    """For N >= 16,384 the product path runs encode and the forward column
    levels as one kernel (a workgroup per 256 coefficients of whole columns);
    its residue planes are congruent, word for word modulo each prime, to
    those of the separate encode and forward-column kernels, and every word
    is a balanced residue |r| <= p/2 + 1 (N1 = 256, 128, 64; encode k-step
    counts 7, 8, 2, 4).  The fused kernel feeds the unreduced encode sums
    (< 2^47) into the transform, so a word may be the other balanced
    representative when the value sits at +-p/2."""
    B = 2
    q = pkg.datagen.modulus(W, W + 11)
    e = pkg.Engine(N, W, q, max_batch=B)
    a = np.stack([pkg.datagen.operands(N, q, 60 + i, edges_first=(i == 0)) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 70 + i, edge_fraction=0.1) for i in range(B)])
    a[0, -1] = (1 << 32) - 1
    da, db = gu.to_dev(a), gu.to_dev(b)
    sep, fused = gu.empty_dev(e.planes_words(B)), gu.empty_dev(e.planes_words(B))
    e.run_stages(pkg.bytecrt.STAGE_ENCODE, a=da, b=db, planes=sep, batch=B)
    e.run_stages(2, planes=sep, batch=B)
    e.run_stages(3, a=da, b=db, planes=fused, batch=B)
    got = gu.from_dev(fused, np.int32).reshape(B, 2, e.P, N).astype(np.int64)
    want = gu.from_dev(sep, np.int32).reshape(B, 2, e.P, N).astype(np.int64)
    p = pkg.bytecrt.plan_host(N, W, q)["primes"].astype(np.int64)[None, None, :, None]
    assert (np.abs(got) <= p // 2 + 1).all() and (np.abs(want) <= p // 2 + 1).all()
    bad = np.argwhere((got - want) % p != 0)
    assert bad.size == 0, f"{len(bad)} words differ mod p, first {bad[:4].tolist()}"
    same = np.mean(got == want)
    assert same > 0.999, same  # other representatives only at the +-p/2 boundary


@pytest.mark.parametrize("N,W", [(4096, 100), (65536, 1747)])
def test_encode_tiles_against_oracle(pkg, orc, gu, N, W):
    """Every encode tile shape the tuner may pick (2^ts_log s by 128 >> ts_log
    r) writes the oracle's residue planes."""
    q = pkg.datagen.modulus(W, N + 3)
    e = pkg.Engine(N, W, q)
    a = pkg.datagen.operands(N, q, 31, edges_first=True)
    b = pkg.datagen.operands(N, q, 32, edge_fraction=0.1)
    pl = orc.Plan(N, W, q)
    want = [gu.to_planes(pl.encode(x), e.N1) for x in (a, b)]
    da, db = gu.to_dev(a), gu.to_dev(b)
    tiles = tuple(range(0, min(7, e.N1.bit_length() - 1) + 1))
    assert len(tiles) == (5 if N == 4096 else 8)
    for ts in tiles:
        e.set_encode_tile(ts)
        planes = gu.empty_dev(e.planes_words(1))
        e.run_stages(pkg.bytecrt.STAGE_ENCODE, a=da, b=db, planes=planes)
        got = gu.from_dev(planes).reshape(2, e.P, N)
        for op in (0, 1):
            np.testing.assert_array_equal(got[op], want[op], err_msg=f"ts_log {ts} operand {op}")
    with pytest.raises(ValueError):
        e.set_encode_tile(tiles[-1] + 1)


def test_tune_encode_tile_exact(pkg, gu, tmp_path):
    """The encode-tile tuner checks every shape's product, leaves the engine on
    the fastest, and a fresh engine applying the recorded table computes the
    same products."""
    import importlib
    import torch
    tune = importlib.import_module("bytecrt_amd.tune")
    N, W, B = 32768, 868, 4
    q = pkg.datagen.modulus(W, 78)
    a = np.stack([pkg.datagen.operands(N, q, 900 + i) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 950 + i) for i in range(B)])
    e = pkg.Engine(N, W, q, max_batch=B)
    da, db = gu.to_dev(a).view(B, N, -1), gu.to_dev(b).view(B, N, -1)
    out = torch.empty_like(da)
    e.run_device(da, db, out, B)
    gu.sync()
    plain = out.clone()
    res = tune.tune_encode_tile(e, da, db, out, B, reps=1)
    assert [r["ts_log"] for r in res["rows"]] == list(range(8))
    assert e.encode_tile == res["best"]["ts_log"]
    assert torch.equal(out, plain)
    path = tmp_path / "table.json"
    tune.record(N, W, B, None, device="test", path=path, encode=res)
    f = pkg.Engine(N, W, q, max_batch=B)
    assert f.apply_tuning(B, tune.load_table(path)) and f.encode_tile == e.encode_tile
    out2 = torch.zeros_like(out)
    f.run_device(da, db, out2, B, graph=True)
    gu.sync()
    assert torch.equal(out2, plain)


@pytest.mark.parametrize("N,W,B", [(16384, 436, 1), (32768, 868, 3)])
def test_tune_fusion_exact(pkg, gu, tmp_path, N, W, B):
    """The fusion tuner runs encode + forward columns fused and as two kernels,
    keeps the product bit-exact either way, leaves the engine on the faster,
    and a fresh engine applying the recorded choice computes the same
    products through the graph path."""
    import importlib
    import torch
    tune = importlib.import_module("bytecrt_amd.tune")
    q = pkg.datagen.modulus(W, 79)
    a = np.stack([pkg.datagen.operands(N, q, 1200 + i, edge_fraction=0.1) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 1300 + i) for i in range(B)])
    e = pkg.Engine(N, W, q, max_batch=B)
    da, db = gu.to_dev(a).view(B, N, -1), gu.to_dev(b).view(B, N, -1)
    out = torch.empty_like(da)
    e.run_device(da, db, out, B)
    gu.sync()
    plain = out.clone()
    res = tune.tune_fusion(e, da, db, out, B, reps=2)
    assert [r["structure"] for r in res["rows"]] == ["fused", "separate"]
    assert e.fuse_min_tiles == res["best"]["fuse_min_tiles"]
    assert torch.equal(out, plain)
    path = tmp_path / "table.json"
    tune.record(N, W, B, None, device="test", path=path, fusion=res)
    f = pkg.Engine(N, W, q, max_batch=B)
    assert f.apply_tuning(B, tune.load_table(path)) and f.fuse_min_tiles == e.fuse_min_tiles
    out2 = torch.zeros_like(out)
    f.run_device(da, db, out2, B, graph=True)
    gu.sync()
    assert torch.equal(out2, plain)


def test_whole_call_multi_chunk(pkg, gu):
    """The host whole call at N=65,536 with B=5: three chunks through the
    staging ring, several pieces per copy, equal to the device-resident run."""
    N, W, B = 65536, 1747, 5
    q = pkg.datagen.modulus(W, 44)
    e = pkg.Engine(N, W, q, max_batch=B)
    a = np.stack([pkg.datagen.operands(N, q, 140 + i, edge_fraction=0.05) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 180 + i) for i in range(B)])
    out = gu.empty_dev(B * N * e.L)
    e.run_device(gu.to_dev(a), gu.to_dev(b), out, B)
    want = gu.from_dev(out).reshape(B, N, e.L)
    for _ in range(2):
        np.testing.assert_array_equal(e.run(a, b), want)
    np.testing.assert_array_equal(e.run(a[3], b[3]), want[3])


def test_fherma_mirror_roundtrip(pkg, orc, gu):
    bc = pkg.bytecrt
    N, W, B = 8192, 220, 2
    q = pkg.datagen.modulus(W, 8)
    L = (W + 31) // 32
    st = bc.fherma_init(bc.Point(B=B, N=N, W=W, L=L, q=bc.Tensor(shape=[L], data=bc.limbs_of(q, L))))
    a = np.stack([pkg.datagen.operands(N, q, i) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 10 + i) for i in range(B)])
    out = bc.fherma_run(st, bc.Inputs(a=bc.Tensor([B, N, L], a.reshape(-1)), b=bc.Tensor([B, N, L], b.reshape(-1))))
    assert out.c.shape == [B, N, L]
    c = out.c.data.reshape(B, N, L)
    pl = orc.Plan(N, W, q)
    for i in range(B):
        np.testing.assert_array_equal(c[i], pl.product(a[i], b[i]))
    with pytest.raises(RuntimeError, match="coefficient count"):
        bc.fherma_run(st, bc.Inputs(a=bc.Tensor([1], a.reshape(-1)[:5]), b=bc.Tensor([1], b.reshape(-1)[:5])))
    bc.fherma_free(st)


@pytest.mark.parametrize("N,W,B", [(65536, 1747, 2), (16384, 436, 3), (8192, 220, 1)])
def test_tune_kernel_flags_exact(pkg, orc, gu, tmp_path, N, W, B):
    """Every kernel-flag setting (XCD tile renumbering, the fused encode's
    MFMA priority and weight staging, the CRT's occupancy and weight
    look-ahead build, the separate encode's MFMA chains per wave) gives the
    same products, bit for bit;
    the tuner leaves the engine on the fastest, records it, and a fresh engine
    that applies the table computes the same products (eager and graph).
    Tables built at N = 8,192 exercise the separate encode kernel and the
    plain grids under the flags too."""
    import importlib
    import torch
    tune = importlib.import_module("bytecrt_amd.tune")
    q = pkg.datagen.modulus(W, W + 5)
    a = np.stack([pkg.datagen.operands(N, q, 30 + i, edge_fraction=0.1) for i in range(B)])
    b = np.stack([pkg.datagen.operands(N, q, 40 + i) for i in range(B)])
    e = pkg.Engine(N, W, q, max_batch=B)
    da, db = gu.to_dev(a).view(B, N, -1), gu.to_dev(b).view(B, N, -1)
    out = torch.empty_like(da)
    res = tune.tune_kernel_flags(e, da, db, out, B, reps=1)
    assert [r["kernel_flags"] for r in res["rows"]] == list(range(32))
    assert e.kernel_flags == res["best"]["kernel_flags"]
    got = gu.from_dev(out).reshape(B, N, e.L)
    np.testing.assert_array_equal(got[0], orc.Plan(N, W, q).product(a[0], b[0]))
    path = tmp_path / "table.json"
    tune.record(N, W, B, None, device="test", path=path, flags=res)
    f = pkg.Engine(N, W, q, max_batch=B)
    assert f.apply_tuning(B, tune.load_table(path)) and f.kernel_flags == e.kernel_flags
    for graph in (False, True):
        out2 = torch.zeros_like(out)
        f.run_device(da, db, out2, B, graph=graph)
        gu.sync()
        assert torch.equal(out2, out)
