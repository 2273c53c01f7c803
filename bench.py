"""Exact negacyclic products/s on one MI355X: B=128 products at N=65,536,
W=1,747 (L=55, 114 primes), batched (one hipGraph launch per step) vs plain
(one product per call), with p50/p99 call latency.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One JSON line on stdout.  Inputs are generated on the device (seeded,
uniform in [0, q) by rejection) and are resident in HBM before timing.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent

N, W, B = 65536, 1747, 128
HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-calls", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-whole-call", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-qps", type=float, default=1000.0)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-warmup-seconds", type=float, default=2.0,
                    help="untimed Poisson stream before the serving leg's timed window")
    ap.add_argument("--serve-budget-gib", type=float, default=4.0)
    ap.add_argument("--only-stage", choices=["encode", "fwd_cols", "encode_fwd", "middle", "inv_cols", "crt", "product",
                                             "hbm_copy", "hbm_read", "hbm_write", "bfly_fp64", "bfly_shoup32",
                                             "bfly_montgomery32", "bfly_fp64_tw5", "bfly_fp64_tw5_pairs", "bfly_fp64_fma_product", "overlap"],
                    help="profiling mode: run one stage (or the whole product, or one probe) --reps times")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--sweep", action="store_true",
                    help="autotune the batch pipelining, write the pipeline table and exit")
    ap.add_argument("--sweep-geometries", default="65536:1747,32768:868,16384:436,8192:220",
                    help="N:W pairs --sweep tunes (B products each)")
    ap.add_argument("--sweep-out", default="pipeline_table.json",
                    help="where --sweep writes its table (copy it over fherma-bytecrt_amd/pipeline_table.json to use it)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="profiling mode: device-to-device copy of this many MiB, --reps times")
    return ap.parse_args()


STAGE_MASK = {"encode": 1, "fwd_cols": 2, "encode_fwd": 3, "middle": 4, "inv_cols": 8, "crt": 16, "product": 31}


def probe_tw15(p: int, psi: int):
    """15 twiddles for the butterfly probes: psi^1 .. psi^15 mod p."""
    return [pow(psi, k, p) for k in range(1, 16)]


def profile_mode(args, torch):
    """Work for rocprofv3 runs: nothing but the named kernels after setup."""
    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as ge
    pkg = ge.load_package()
    probes = pkg.probes
    if args.copy_probe_mib or args.only_stage.startswith("hbm_"):
        nbytes = (args.copy_probe_mib or 2048) * (1 << 20)
        mode = 0 if args.copy_probe_mib else {"hbm_copy": 0, "hbm_read": 1, "hbm_write": 2}[args.only_stage]
        src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.reps):
            probes.hbm_stream(mode, src, dst, nbytes)
        torch.cuda.synchronize()
        print(json.dumps({"hbm_probe": probes.MODES[mode], "bytes": nbytes, "reps": args.reps}), flush=True)
        return
    if args.only_stage == "overlap":
        # every (mode, kind) of the MFMA/VALU overlap probe, reps launches each
        words = int(probes.load_library().bcrtp_issue_overlap_words())
        out = torch.zeros(words, dtype=torch.int32, device="cuda")
        for kind in probes.KINDS:
            for mode in (1, 2, 3):
                for _ in range(args.reps):
                    probes.issue_overlap(mode, kind, 4000, 16000, out)
        torch.cuda.synchronize()
        print(json.dumps({"overlap_probe": list(probes.KINDS.values()), "mfma_iters": 4000, "valu_iters": 16000,
                          "reps": args.reps}), flush=True)
        return
    q = pkg.datagen.modulus(W, 2026)
    if args.only_stage.startswith("bfly_"):
        pl = pkg.bytecrt.plan_host(N, W, q)
        p, psi = int(pl["primes"][0]), int(pl["psi"][0])
        form = {"bfly_fp64": 0, "bfly_shoup32": 1, "bfly_montgomery32": 2, "bfly_fp64_tw5": 3,
                "bfly_fp64_tw5_pairs": 4, "bfly_fp64_fma_product": 5}[args.only_stage]
        threads, rounds = 1 << 20, 64
        out = torch.empty(threads * 16, dtype=torch.int32, device="cuda")
        for _ in range(args.reps):
            probes.butterflies(form, out, threads, p, probe_tw15(p, psi), rounds)
        torch.cuda.synchronize()
        print(json.dumps({"butterfly_probe": probes.FORMS[form], "butterflies": threads * rounds * 32,
                          "reps": args.reps}), flush=True)
        return
    eng = pkg.Engine(N, W, q, max_batch=B)
    a = device_operands(torch, B, N, q, 11)
    b = device_operands(torch, B, N, q, 12)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    eng.run_stages(31, a=a, b=b, c=c, batch=B, stream=stream)
    for _ in range(args.reps):
        eng.run_stages(STAGE_MASK[args.only_stage], a=a, b=b, c=c, batch=B, stream=stream)
    torch.cuda.synchronize()
    print(json.dumps({"stage": args.only_stage, "reps": args.reps, "batch": B}), flush=True)


def device_operands(torch, batch: int, n: int, q: int, seed: int):
    """[batch][n][L] int32 (u32 bit patterns) uniform in [0, q), on the GPU."""
    L = (q.bit_length() + 31) // 32
    w = q.bit_length()
    qv = torch.tensor([(q >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=torch.int64, device="cuda")
    top_mask = (1 << (w - 32 * (L - 1))) - 1
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.empty((batch * n, L), dtype=torch.int64, device="cuda")
    todo = torch.arange(batch * n, device="cuda")
    while todo.numel():
        x = torch.randint(0, 1 << 32, (todo.numel(), L), generator=g, device="cuda", dtype=torch.int64)
        x[:, L - 1] &= top_mask
        diff = x != qv
        # highest differing limb decides x < q
        idx = (L - 1) - torch.argmax(diff.flip(1).to(torch.int8), dim=1)
        lt = diff.any(dim=1) & (x.gather(1, idx[:, None]).squeeze(1) < qv[idx])
        out[todo[lt]] = x[lt]
        todo = todo[~lt]
    out = torch.where(out >= (1 << 31), out - (1 << 32), out).to(torch.int32)
    return out.reshape(batch, n, L).contiguous()


def measured_traffic(stage: str):
    """HBM bytes per B=128 dispatch of a stage kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated), or None."""
    path = ROOT / "profiles" / "traffic_b128_n65536_w1747.txt"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == "k_" + stage:
            return int(f[1]) + int(f[2])
    return None


def stage_bytes(P: int, L: int):
    """Algorithmic HBM bytes per product for each kernel (compulsory reads+writes)."""
    R = N * P * 4          # one residue plane set for one operand
    io = N * L * 4         # one coefficient tensor
    return {
        "encode": 2 * io + 2 * R,
        "fwd_cols": 4 * R,
        "encode_fwd": 2 * io + 2 * R,
        "middle": 3 * R,
        "inv_cols": 2 * R,
        "crt": R + io,
    }


def stage_ops(P: int, L: int, Kenc: int, Kcrt: int, Mcrt: int):
    """Algorithmic int8 MFMA ops (2 per MAC) per product for the GEMM kernels."""
    enc = 2 * 2 * N * (4 * L) * (4 * P)
    return {"encode": enc, "encode_fwd": enc, "crt": 2 * N * (4 * (P + 1)) * (4 * L + 1)}


def whole_call(eng, a, b, c_dev, torch):
    """The reference's timing boundary: host buffers in, host buffers out
    (pack, H2D, all GPU work, D2H, unpack), B=128 per call and B=1 per call."""
    ah = a.cpu().numpy().view(np.uint32)
    bh = b.cpu().numpy().view(np.uint32)
    ref = c_dev.cpu().numpy().view(np.uint32)
    eng.run(ah, bh)  # warm-up: allocates staging once
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        ch = eng.run(ah, bh)
        times.append(time.perf_counter() - t0)
    exact = bool(np.array_equal(ch, ref))
    one = []
    for i in range(16):
        t0 = time.perf_counter()
        eng.run(ah[i], bh[i])
        one.append(time.perf_counter() - t0)
    ms = float(np.median(times) * 1e3)
    ms1 = float(np.median(one) * 1e3)
    out = {"boundary": "whole call: host arrays in and out (PCIe and host copies included)",
           "batch_ms_median": round(ms, 3), "batch_products_per_s": round(B / (ms / 1e3), 1),
           "single_ms_median": round(ms1, 4), "single_products_per_s": round(1e3 / ms1, 1),
           "batched_vs_plain": round((B / ms) / (1.0 / ms1), 3), "exact_vs_device": exact}
    out["fherma_run"] = fherma_run_timing()
    return out


def fherma_run_timing():
    """fherma_run through include/fherma.h at the reference harness's timing
    boundary (std::vector operands in, owned std::vector out), measured by the
    compiled fherma_call_bench in a child process: B=128 per call, and B=1
    per call (the plain one-product-per-call form)."""
    import subprocess
    exe = ROOT / "fherma-bytecrt_amd" / "lib" / "fherma_call_bench"
    res = {}
    for batch, reps in ((B, 7), (1, 41)):
        r = subprocess.run([str(exe), str(batch), str(N), str(W), "7", str(reps)], capture_output=True, text=True,
                           timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"fherma_call_bench B={batch} failed: {r.stderr.strip()[-400:]}")
        res[batch] = json.loads(r.stdout.strip().splitlines()[-1])
    bms, ms1 = res[B]["median_ms"], res[1]["median_ms"]
    return {"boundary": "fherma_run call (native-batched/main.cpp:182-187): pack, H2D, GPU, D2H, owned vector",
            "B128_ms_median": bms, "B128_ms_min": res[B]["min_ms"], "B128_products_per_s": round(B / bms * 1e3, 1),
            "B1_ms_median": ms1, "B1_products_per_s": round(1e3 / ms1, 1),
            "plain_serial_B128_ms": round(B * ms1, 2), "batched_vs_plain": round(B * ms1 / bms, 3),
            "init_s": res[B]["init_s"], "exact": bool(res[B]["exact"] and res[1]["exact"]),
            "reference_rtx_pro_6000": {"B128_ms": 89.648935, "B1_ms": 1.674401, "plain_serial_B128_ms": 291.519151}}


def with_edges(torch, x, q: int, fraction: float, seed: int):
    """Replace `fraction` of the coefficients of x ([..., N, L] int32 limbs) by
    edge values {0, 1, q-1, floor(q/2), q-floor(q/2)}."""
    L = x.shape[-1]
    vals = [0, 1, q - 1, q // 2, q - q // 2]
    table = torch.tensor([[(v >> (32 * i)) & 0xFFFFFFFF for i in range(L)] for v in vals], dtype=torch.int64,
                         device=x.device)
    table = torch.where(table >= (1 << 31), table - (1 << 32), table).to(torch.int32)
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    flat = x.reshape(-1, L)
    mask = torch.rand(flat.shape[0], generator=g, device=x.device) < fraction
    pick = torch.randint(0, len(vals), (flat.shape[0],), generator=g, device=x.device)
    flat[mask] = table[pick[mask]]
    return x


UNEVEN = [(32768, 868, 32), (32768, 868, 32), (32768, 868, 32), (65536, 1747, 11), (65536, 1747, 11),
          (65536, 1747, 10)]


def uneven_leg(pkg, torch, steps: int):
    """One launch per stage over 96 products at N=32,768/W=868 and 32 at
    N=65,536/W=1,747, three seeded moduli per geometry, 10% edge coefficients;
    against the same groups launched one after another."""
    groups, engines = [], []
    for k, (n, w, cnt) in enumerate(UNEVEN):
        q = pkg.datagen.modulus(w, 400 + k)
        e = pkg.Engine(n, w, q, max_batch=cnt)
        a = with_edges(torch, device_operands(torch, cnt, n, q, 500 + k), q, 0.1, 600 + k)
        b = with_edges(torch, device_operands(torch, cnt, n, q, 700 + k), q, 0.1, 800 + k)
        c = torch.empty_like(a)
        groups.append((e, a, b, c, cnt))
        engines.append(e)
    torch.cuda.synchronize()
    mix = pkg.Mix(groups)
    stream = torch.cuda.Stream()

    run_mixed = lambda: mix.run(stream=stream, graph=True)
    run_separate = lambda: [g[0].run_device(g[1], g[2], g[3], g[4], stream=stream, graph=True) for g in groups]
    run_mixed()
    torch.cuda.synchronize()
    got = [g[3].clone() for g in groups]
    for g in groups:
        g[3].zero_()
    run_separate()
    torch.cuda.synchronize()
    exact = all(bool(torch.equal(g[3], r)) for g, r in zip(groups, got))

    # the two forms timed round robin (alternating which goes first), medians:
    # timed one after the other, the later form would also gain from a clock
    # still rising after the work before
    times = ([], [])
    for _ in range(2):
        run_mixed()
        run_separate()
    torch.cuda.synchronize()
    for r in range(6):
        for i in ((0, 1) if r % 2 == 0 else (1, 0)):
            fn = run_mixed if i == 0 else run_separate
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            for _ in range(steps):
                fn()
            e.record(stream)
            torch.cuda.synchronize()
            times[i].append(s.elapsed_time(e) / steps)
    mixed_ms, separate_ms = statistics.median(times[0]), statistics.median(times[1])
    one = torch.empty_like(groups[3][1][0])
    groups[3][0].run_device(groups[3][1][5], groups[3][2][5], one, 1, stream=stream)
    torch.cuda.synchronize()
    exact = exact and bool(torch.equal(one, got[3][5]))
    # every candidate schedule, each checked bit for bit by the tuner
    import bytecrt_amd.tune as tune
    sweep = tune.tune_mix(mix, [g[3] for g in groups], reps=steps, stream=stream)
    exact = exact and all(bool(torch.equal(g[3], r)) for g, r in zip(groups, got))
    total = sum(g[4] for g in groups)
    mix.close()
    return {"workload": "96 x (N=32768, W=868) + 32 x (N=65536, W=1747), 3 moduli per geometry, "
                        "10% edge coefficients, one launch per stage and geometry class (hipGraph)",
            "products": total, "mixed_ms": round(mixed_ms, 3),
            "mixed_products_per_s": round(total / mixed_ms * 1e3, 1),
            "separate_launches_ms": round(separate_ms, 3), "exact": exact,
            "schedule_sweep": sweep["rows"]}


SERVE_GEOMS = [(8192, 220), (16384, 436), (32768, 868), (65536, 1747)]


def serving_leg(pkg, qps: float, seconds: float, budget_gib: float, warm_seconds: float = 2.0):
    """Poisson arrivals at a fixed rate, N drawn uniformly from SERVE_GEOMS,
    through the continuous-batching server; every result checked against the
    64-bit xxh3 digest of its reference product (a full compare of every
    result on the client's one waiter thread cannot keep up with the arrivals).
    The stream runs warm_seconds before the timed `seconds`: in its first
    second or so a fresh server process sees a transient of 20-30 ms
    latencies (host copies slowing for about 0.2 s; the cause is not
    established), so steady-state p50/p99 come from the timed window and the
    warm-up window's p99 is reported as cold_p99_ms."""
    import threading
    try:
        import xxhash
        digest = xxhash.xxh3_64_intdigest
    except ImportError:  # exact compare through a hash of the raw bytes
        import hashlib
        digest = lambda x: hashlib.blake2b(x.tobytes(), digest_size=8).digest()  # noqa: E731
    rng = np.random.default_rng(2026)
    pools = []
    for k, (n, w) in enumerate(SERVE_GEOMS):
        q = pkg.datagen.modulus(w, 900 + k)
        e = pkg.Engine(n, w, q, max_batch=1)
        entries = []
        for i in range(3):
            a = pkg.datagen.operands(n, q, 1000 + 10 * k + i)
            b = pkg.datagen.operands(n, q, 2000 + 10 * k + i)
            ref = e.run(a, b)
            entries.append((a, b, digest(np.ascontiguousarray(ref))))
        e.close()
        pools.append((n, w, q, entries))
    count = max(1, int(qps * (seconds + warm_seconds)))
    gaps = rng.exponential(1.0 / qps, size=count)
    kinds = rng.integers(0, len(SERVE_GEOMS), size=count)
    picks = rng.integers(0, 3, size=count)
    srv = pkg.Server(hbm_budget_bytes=int(budget_gib * (1 << 30)), max_wave=128)
    # warm the plan cache (table upload is a one-time cost per geometry)
    for n, w, q, entries in pools:
        srv.wait(srv.submit(n, w, q, entries[0][0], entries[0][1]))
    tickets = [None] * count
    done_at = [0.0] * count
    sub_at = [0.0] * count
    lat = [0.0] * count
    ok = [False] * count
    submitted = threading.Semaphore(0)

    # result buffers are recycled per geometry, as a serving client would, and
    # written once up front (untouched pages would fault inside the server's
    # copy into them)
    free = [[np.full((n, (w + 31) // 32), 0, dtype=np.uint32) for _ in range(48)] for n, w in SERVE_GEOMS]
    free_lock = threading.Lock()

    def waiter():
        for i in range(count):
            submitted.acquire()
            c, lat[i] = srv.wait(tickets[i])
            done_at[i] = time.perf_counter()
            ok[i] = digest(c) == pools[kinds[i]][3][picks[i]][2]
            with free_lock:
                free[kinds[i]].append(c)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for i in range(count):
        due += gaps[i]
        now = time.perf_counter()
        if due > now:
            time.sleep(due - now)
        n, w, q, entries = pools[kinds[i]]
        a, b, _ = entries[picks[i]]
        with free_lock:
            c = free[kinds[i]].pop() if free[kinds[i]] else np.full((n, (w + 31) // 32), 0, dtype=np.uint32)
        sub_at[i] = time.perf_counter() - t0
        tickets[i] = srv.submit(n, w, q, a, b, c)
        submitted.release()
    th.join()
    st = srv.stats()
    srv.close()
    timed = np.array(sub_at) >= warm_seconds
    span = max(done_at) - (t0 + warm_seconds)
    all_ms = np.array(lat) * 1e3
    lat_ms = all_ms[timed]
    kinds_t = kinds[timed]
    per = {}
    for k, (n, w) in enumerate(SERVE_GEOMS):
        sel = lat_ms[kinds_t == k]
        if sel.size:
            per[f"N{n}"] = {"requests": int(sel.size), "p50_ms": round(float(np.percentile(sel, 50)), 3),
                            "p99_ms": round(float(np.percentile(sel, 99)), 3)}
    return {"workload": "Poisson arrivals, N uniform over {8192, 16384, 32768, 65536} with W {220, 436, 868, "
                        "1747}, host operands in and results out (whole call per request)",
            "offered_qps": qps, "requests": int(timed.sum()), "hbm_budget_gib": budget_gib,
            "products_per_s": round(int(timed.sum()) / span, 1),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3), "per_geometry": per,
            "warm_up": {"seconds": warm_seconds, "requests": int((~timed).sum()),
                        "cold_p99_ms": round(float(np.percentile(all_ms[~timed], 99)), 3) if (~timed).any() else None},
            "waves": st["waves"] - len(SERVE_GEOMS), "max_wave": st["max_wave"], "exact": all(ok)}


def cpu_threads():
    """Threads for the CPU baseline: the CPU share this process is given
    (OMP_NUM_THREADS where the host sets it, which is how a GPU host states
    one GPU's share of its cores), else every CPU of the affinity mask."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    share = int(env) if env.isdigit() and int(env) > 0 else affinity
    return min(share, affinity), affinity


def cpu_model():
    cpu = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def cpu_products(pkg, n: int, w: int, q: int, seconds: float, threads: int):
    """The oracle's full product (C restatement of the reference algorithm,
    OpenMP) on seeded operands of one geometry, repeated for about `seconds`."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    oracle.set_threads(threads)
    pl = oracle.Plan(n, w, q)
    a = pkg.datagen.operands(n, q, 1001)
    b = pkg.datagen.operands(n, q, 1002)
    done, t0 = 0, time.perf_counter()
    while True:
        pl.product(a, b)
        done += 1
        el = time.perf_counter() - t0
        if el >= seconds or done >= 256:
            break
    return done, el


def cpu_baseline(pkg, q, seconds: float):
    threads, affinity = cpu_threads()
    done, el = cpu_products(pkg, N, W, q, seconds, threads)
    n1, w1 = 32768, 868
    q1 = pkg.datagen.modulus(w1, 2026)
    done1, el1 = cpu_products(pkg, n1, w1, q1, max(seconds / 3, 2.0), threads)
    cpu = cpu_model()
    return {"value": round(done / el, 3), "unit": "products/s", "cores": threads, "kind": "port",
            "affinity_cpus": affinity,
            "sample": f"{done} full products N={N} W={W} (oracle C restatement of the reference algorithm, "
                      f"OpenMP, {threads} threads) in {el:.1f} s on {cpu}",
            "config1": {"N": n1, "W": w1, "value": round(done1 / el1, 3), "unit": "products/s",
                        "sample": f"{done1} full products in {el1:.1f} s, same threads"}}


def sweep_mode(args, torch):
    """Autotune the encode tile shape, the fusion of encode with the forward
    columns, the kernel flags and then the batch pipelining for B products at
    each geometry of --sweep-geometries (the bench geometry and the other
    configs' N, W; every candidate's product is checked bit for bit against
    the first) and record the choices in a pipeline table (copy it into the
    package to make it the default)."""
    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as ge
    pkg = ge.load_package()
    import bytecrt_amd.tune as tune
    out = Path(args.sweep_out)
    out.parent.mkdir(parents=True, exist_ok=True)
    for geo in args.sweep_geometries.split(","):
        n, w = (int(x) for x in geo.split(":"))
        sweep_geometry(args, torch, pkg, tune, n, w, B, out)


def sweep_geometry(args, torch, pkg, tune, n: int, w: int, batch: int, out: Path):
    """Tune the encode tile, the fusion and the kernel flags (and the
    pipelining, while its tuner is on) for ``batch`` products at (n, w) and
    record them in the table at ``out``."""
    q = pkg.datagen.modulus(w, 2026)
    eng = pkg.Engine(n, w, q, max_batch=batch)
    a = device_operands(torch, batch, n, q, 11)
    b = device_operands(torch, batch, n, q, 12)
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    tag = {"N": n, "W": w, "B": batch}
    enc = tune.tune_encode_tile(eng, a, b, c, batch, reps=args.reps, stream=stream)
    for r in enc["rows"]:
        print(json.dumps({**tag, **r}), flush=True)
    fus = tune.tune_fusion(eng, a, b, c, batch, reps=args.reps, stream=stream)
    for r in fus["rows"]:
        print(json.dumps({**tag, **r}), flush=True)
    kfl = tune.tune_kernel_flags(eng, a, b, c, batch, reps=args.reps, stream=stream)
    for r in kfl["rows"]:
        print(json.dumps({**tag, **r}), flush=True)
    try:
        res = tune.tune_pipeline(eng, a, b, c, batch, reps=args.reps, stream=stream)
    except RuntimeError as x:  # the pipeline tuner is switched off
        print(json.dumps({"pipeline": str(x)}), flush=True)
        res = None
    for r in (res or {}).get("rows", []):
        print(json.dumps({**tag, **r}), flush=True)
    tune.record(n, w, batch, res, device=torch.cuda.get_device_name(0), path=out, encode=enc, fusion=fus, flags=kfl)
    print(json.dumps({**tag, "best": res and res["best"], "encode_tile": enc["best"], "fusion": fus["best"],
                      "kernel_flags": kfl["best"], "table": str(out)}), flush=True)
    del eng, a, b, c
    torch.cuda.empty_cache()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if args.only_stage or args.copy_probe_mib:
        profile_mode(args, torch)
        return
    if args.sweep:
        sweep_mode(args, torch)
        return
    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as ge
    pkg = ge.load_package()
    q = pkg.datagen.modulus(W, 2026)
    eng = pkg.Engine(N, W, q, max_batch=B)
    L, P = eng.L, eng.P
    a = device_operands(torch, B, N, q, 11)
    b = device_operands(torch, B, N, q, 12)
    c = torch.empty_like(a)
    c1 = torch.empty_like(a[0])
    torch.cuda.synchronize()
    # A dedicated stream: events and kernels are ordered on the same queue.
    stream = torch.cuda.Stream()

    # -- plain: one product per call, each call waits for its result ----------
    for i in range(3):
        eng.run_device(a[i], b[i], c1, 1, stream=stream)
    torch.cuda.synchronize()
    lat = []
    for i in range(args.plain_calls):
        t0 = time.perf_counter()
        eng.run_device(a[i % B], b[i % B], c1, 1, stream=stream)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    lat_ms = np.array(lat) * 1e3
    plain_pps = 1000.0 / lat_ms.mean()
    # exactness spot check: batched product 7 equals its plain result
    eng.run_device(a[7], b[7], c1, 1, stream=stream)

    # -- batched: one graph launch per step, pipelining from the tuned table ---
    tuned = eng.apply_tuning(B)
    pipeline = {"chunk": eng.pipeline[0], "streams": eng.pipeline[1], "encode_tile_ts_log": eng.encode_tile,
                "fuse_min_tiles": eng.fuse_min_tiles, "kernel_flags": eng.kernel_flags,
                "source": "table" if tuned else "default"}
    for _ in range(max(args.warmup, 1)):
        eng.run_device(a, b, c, B, stream=stream, graph=True)
    torch.cuda.synchronize()
    exact = bool(torch.equal(c[7], c1))
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in zip(starts, ends):
        s.record(stream)
        eng.run_device(a, b, c, B, stream=stream, graph=True)
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = B / (ms_per_step / 1e3)
    # -- batched call latency: >= 100 graph replays, each waited for ----------
    lat_b = []
    for _ in range(max(100, args.steps)):
        t1 = time.perf_counter()
        eng.run_device(a, b, c, B, stream=stream, graph=True)
        stream.synchronize()
        lat_b.append(time.perf_counter() - t1)
    batch_ms = np.array(lat_b) * 1e3

    # -- per-kernel durations on the same stream (hip events) ----------------
    # The product path's kernels (encode and forward columns fused), then the
    # separate encode and forward-column kernels for comparison.
    names = ["encode_fwd", "middle", "inv_cols", "crt"]
    masks = [3, 4, 8, 16]
    unfused_names = ["encode", "fwd_cols"]
    reps = 3
    dur = {n: 0.0 for n in names + unfused_names}
    for _ in range(reps):
        for n, m in zip(names + unfused_names, masks + [1, 2]):
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            eng.run_stages(m, a=a, b=b, c=c, batch=B, stream=stream)
            e.record(stream)
            torch.cuda.synchronize()
            dur[n] += s.elapsed_time(e) / reps
    sb = stage_bytes(P, L)
    so = stage_ops(P, L, eng.Kenc, eng.Kcrt, eng.Mcrt)
    stages = {}
    for n in names + unfused_names:
        gbs = sb[n] * B / (dur[n] * 1e-3) / 1e9
        st = {"ms": round(dur[n], 4), "hbm_GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        if n in so:
            tops = so[n] * B / (dur[n] * 1e-3) / 1e12
            st.update(int8_TOPs=round(tops, 1), mfma_frac=round(tops / INT8_PEAK_TOPS, 4))
        stages[n] = st
    unfused = {n: stages.pop(n) for n in unfused_names}
    dom = max(names, key=lambda n: dur[n])
    if dom in so and stages[dom]["mfma_frac"] > stages[dom]["hbm_frac"]:
        roof = {"bound": "mfma", "achieved": stages[dom]["int8_TOPs"], "peak": INT8_PEAK_TOPS,
                "unit": "TOP/s", "frac": stages[dom]["mfma_frac"], "traffic": None, "kernel": dom}
    else:
        roof = {"bound": "hbm", "achieved": stages[dom]["hbm_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": stages[dom]["hbm_frac"], "traffic": measured_traffic(dom), "kernel": dom,
                "algorithmic_bytes_per_launch": sb[dom] * B}
    # Measured ceilings: hand-written HBM stream kernels and the butterfly
    # rate of each exact modular-multiplication form (probes library).
    pt = eng.plan_tables()
    p0, psi0 = int(pt["primes"][0]), int(pt["psi"][0])
    ceilings = {"hbm": pkg.probes.measure_hbm(torch),
                "butterflies": pkg.probes.measure_butterflies(torch, p0, probe_tw15(p0, psi0)),
                "issue_overlap": pkg.probes.measure_overlap(torch)}
    if roof["bound"] == "hbm":
        roof["practical_peak"] = ceilings["hbm"]["copy_GBps"]
        roof["practical_frac"] = round(roof["achieved"] / roof["practical_peak"], 4)
    # NTT butterflies per batch (2 forward + 1 inverse transform per prime,
    # N/2 log2 N each) at the measured rates of the forms the kernels use: the
    # fused encode's forward column levels (n1 per operand) the six-operation
    # form, the middle and inverse-column levels the five-operation form
    # (three and four waves per SIMD, the kernels' occupancy).  The middle
    # kernel's pair step (last forward level of both operands, pointwise
    # product, first inverse level: 28 FP64 operations per pair) counts as
    # four butterflies of seven operations, where those three levels had three
    bf = ceilings["butterflies"]
    n1 = (N // 256).bit_length() - 1
    cols_bfly = 2 * P * (N // 2) * n1 * B
    rest_bfly = 3 * P * (N // 2) * 16 * B - cols_bfly + P * (N // 2) * B
    ceilings["ntt_fp64_floor_ms"] = round((cols_bfly / bf["fp64_4w_Tbfly_per_s"] +
                                           rest_bfly / bf["fp64_tw5_4w_Tbfly_per_s"]) / 1e12 * 1e3, 3)
    # NTT-only kernels of the product path (the forward column levels run
    # inside the fused encode kernel)
    ntt_ms = dur["middle"] + dur["inv_cols"]
    ntt_bytes = (sb["middle"] + sb["inv_cols"]) * B
    ntt_gbs = ntt_bytes / (ntt_ms * 1e-3) / 1e9

    whole = None if args.no_whole_call else whole_call(eng, a, b, c, torch)
    uneven = None if args.no_uneven else uneven_leg(pkg, torch, max(3, args.steps // 2))
    serving = None if args.no_serving else serving_leg(pkg, args.serve_qps, args.serve_seconds,
                                                       args.serve_budget_gib, args.serve_warmup_seconds)
    cpu = None if args.no_cpu else cpu_baseline(pkg, q, args.cpu_seconds)
    line = {
        "metric": "exact products/sec, B=128 N=65,536 W=1,747 (batched vs plain); p50/p99 latency",
        "value": round(value, 2),
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/int8-mfma (exact integer)",
        "data": "synthetic: seeded random odd q of W bits, operands uniform in [0,q), device-resident",
        "config": {"workload": "batched homogeneous exact negacyclic product", "B": B, "N": N, "W": W,
                   "L": L, "primes": P, "boundary": "device-resident (inputs in HBM before timing)"},
        "plain": {"products_per_s": round(plain_pps, 2), "calls": args.plain_calls,
                  "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
                  "p99_ms": round(float(np.percentile(lat_ms, 99)), 4)},
        "batched": {"products_per_s": round(value, 2), "replays": int(batch_ms.size),
                    "p50_ms": round(float(np.percentile(batch_ms, 50)), 4),
                    "p99_ms": round(float(np.percentile(batch_ms, 99)), 4)},
        "batched_vs_plain": round(value / plain_pps, 3),
        "pipeline": pipeline,
        "exact_spot_check": exact,
        "stages": stages,
        "unfused_stages": unfused,
        "ntt_hbm_frac": round(ntt_gbs / HBM_PEAK_GBS, 4),
        "ntt_practical_frac": round(ntt_gbs / ceilings["hbm"]["copy_GBps"], 4),
        "roofline": roof,
        "ceilings": ceilings,
        "whole_call": whole,
        "uneven": uneven,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
