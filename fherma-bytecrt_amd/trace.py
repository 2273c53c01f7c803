"""Tracing for the ByteCRT engine: a rocprofv3 wrapper and a decoder.

    python fherma-bytecrt_amd/trace.py run --out DIR -- python3 bench.py --only-stage product --reps 3
    python fherma-bytecrt_amd/trace.py run --out DIR --pmc FETCH_SIZE -- python3 bench.py ...
    python fherma-bytecrt_amd/trace.py decode DIR [--markdown FILE] [--scale FETCH_SIZE=2.0]

``run`` executes the program under rocprofv3 with kernel, memory-copy and
marker (ROCTx) traces and per-kernel statistics, CSV output, and BCRT_TRACE=1
so the library emits its host ranges (whole calls, operand packing, waits,
unpacking, server waves).  With ``--pmc`` it collects those counters instead
(a counter run carries no traces).  ``decode`` turns the CSV files into:

* per-kernel statistics (calls, mean / min / max duration, share of GPU time);
* per-lane activity -- compute (kernels), H2D and D2H copies, and each host
  range name -- with busy time, first-to-last coverage, idle gaps (sum, largest)
  and the union of the GPU lanes;
* for every whole-call range (bcrt:whole_call / bcrt:fherma_run) the same lane
  figures inside that call, which show which resource bounds it;
* with ``--pmc`` output, mean counter values per kernel (optionally scaled by a
  calibration factor per counter).

Everything is plain Python over the CSV files; no GPU is needed to decode.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import subprocess
import sys
from collections import defaultdict
from pathlib import Path

# ---------------------------------------------------------------------------
# Interval arithmetic
# ---------------------------------------------------------------------------


def union(intervals):
    """Sorted disjoint union of [start, end) intervals."""
    out = []
    for s, e in sorted(intervals):
        if e <= s:
            continue
        if out and s <= out[-1][1]:
            out[-1][1] = max(out[-1][1], e)
        else:
            out.append([s, e])
    return [tuple(x) for x in out]


def clip(intervals, lo, hi):
    return [(max(s, lo), min(e, hi)) for s, e in intervals if e > lo and s < hi]


def lane_stats(intervals, lo=None, hi=None):
    """Busy time, coverage and idle gaps of a lane (ns), optionally inside [lo, hi)."""
    if lo is not None:
        intervals = clip(intervals, lo, hi)
    u = union(intervals)
    if not u:
        return {"count": 0, "busy_ns": 0, "coverage_ns": 0, "idle_ns": 0, "max_gap_ns": 0}
    busy = sum(e - s for s, e in u)
    first, last = u[0][0], u[-1][1]
    gaps = [b[0] - a[1] for a, b in zip(u, u[1:])]
    out = {"count": len(intervals), "busy_ns": busy, "coverage_ns": last - first, "idle_ns": sum(gaps),
           "max_gap_ns": max(gaps) if gaps else 0}
    if lo is not None:
        out["head_idle_ns"] = first - lo
        out["tail_idle_ns"] = hi - last
    return out


# ---------------------------------------------------------------------------
# CSV loading
# ---------------------------------------------------------------------------


def _find(d: Path, suffix: str):
    hits = sorted(d.rglob(f"*{suffix}"))
    return hits[0] if hits else None


def _rows(path):
    if path is None:
        return []
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def short_kernel(name: str) -> str:
    """k_name<args> (integer and bool template arguments kept, spaces dropped)."""
    m = re.search(r"(k_\w+(<[\w, ]+>)?)", name)
    return m.group(1).replace(" ", "") if m else name.split("(")[0][-60:]


def load(d: Path):
    kernels = [(short_kernel(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
               for r in _rows(_find(d, "kernel_trace.csv"))]
    copies = []
    for r in _rows(_find(d, "memory_copy_trace.csv")):
        direction = r.get("Direction") or r.get("Operation") or r.get("Kind", "")
        lane = "h2d" if "HOST_TO_DEVICE" in direction else "d2h" if "DEVICE_TO_HOST" in direction else "d2d"
        copies.append((lane, int(r["Start_Timestamp"]), int(r["End_Timestamp"]), int(r.get("Bytes", 0) or 0)))
    markers = []
    for r in _rows(_find(d, "marker_api_trace.csv")):
        name = r.get("Function") or r.get("Message") or ""
        if not name.startswith("bcrt:"):
            continue
        markers.append((name, int(r.get("Thread_Id", 0) or 0), int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    counters = _rows(_find(d, "counter_collection.csv"))
    return kernels, copies, markers, counters


# ---------------------------------------------------------------------------
# Decoding
# ---------------------------------------------------------------------------


def decode(d: Path, scale: dict | None = None) -> dict:
    kernels, copies, markers, counters = load(d)
    out: dict = {"dir": str(d)}
    if kernels:
        per = defaultdict(list)
        for n, s, e in kernels:
            per[n].append(e - s)
        total = sum(sum(v) for v in per.values())
        out["kernels"] = {n: {"calls": len(v), "mean_us": sum(v) / len(v) / 1e3, "min_us": min(v) / 1e3,
                              "max_us": max(v) / 1e3, "share": sum(v) / total}
                          for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))}
    lanes = {"compute": [(s, e) for _, s, e in kernels]}
    for lane in ("h2d", "d2h", "d2d"):
        iv = [(s, e) for ln, s, e, _ in copies if ln == lane]
        if iv:
            lanes[lane] = iv
    if any(b for *_, b in copies):  # older rocprofv3 versions report no sizes
        out["copy_bytes"] = {lane: sum(b for ln, _, _, b in copies if ln == lane) for lane in ("h2d", "d2h")}
    host = defaultdict(list)
    for name, _, s, e in markers:
        host[name].append((s, e))
    gpu = lanes["compute"] + lanes.get("h2d", []) + lanes.get("d2h", [])
    out["lanes"] = {k: lane_stats(v) for k, v in lanes.items()}
    out["lanes"]["gpu_union"] = lane_stats(gpu)
    out["host_ranges"] = {k: lane_stats(v) for k, v in sorted(host.items())}
    calls = sorted(host.get("bcrt:fherma_run", []) or host.get("bcrt:whole_call", []))
    out["calls"] = []
    for lo, hi in calls:
        c = {"wall_ns": hi - lo}
        for k, v in lanes.items():
            c[k] = lane_stats(v, lo, hi)
        c["gpu_union"] = lane_stats(gpu, lo, hi)
        for name in ("bcrt:pack", "bcrt:unpack", "bcrt:wait", "bcrt:slot_wait"):
            if name in host:
                c[name] = lane_stats(host[name], lo, hi)
        out["calls"].append(c)
    if counters:
        per = defaultdict(lambda: defaultdict(list))
        for r in counters:
            per[short_kernel(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        scale = scale or {}
        out["counters"] = {k: {c: sum(v) / len(v) * scale.get(c, 1.0) for c, v in cs.items()}
                           for k, cs in per.items()}
    return out


def markdown(res: dict) -> str:
    lines = [f"Decoded from `{res['dir']}`.", ""]
    if "kernels" in res:
        lines += ["| kernel | calls | mean us | min us | max us | share |", "|---|---:|---:|---:|---:|---:|"]
        for n, k in res["kernels"].items():
            lines.append(f"| {n} | {k['calls']} | {k['mean_us']:.1f} | {k['min_us']:.1f} | {k['max_us']:.1f} | "
                         f"{100 * k['share']:.1f}% |")
        lines.append("")
    lines += ["| lane | intervals | busy ms | coverage ms | idle ms | largest gap us |",
              "|---|---:|---:|---:|---:|---:|"]
    for n, s in list(res["lanes"].items()) + list(res["host_ranges"].items()):
        lines.append(f"| {n} | {s['count']} | {s['busy_ns'] / 1e6:.3f} | {s['coverage_ns'] / 1e6:.3f} | "
                     f"{s['idle_ns'] / 1e6:.3f} | {s['max_gap_ns'] / 1e3:.1f} |")
    if res["calls"]:
        lines += ["", "Per whole call (busy ms inside the call):", "",
                  "| call | wall ms | compute | h2d | d2h | gpu union | pack | unpack | wait |",
                  "|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for i, c in enumerate(res["calls"]):
            def b(k):
                return f"{c[k]['busy_ns'] / 1e6:.2f}" if k in c else "-"
            lines.append(f"| {i} | {c['wall_ns'] / 1e6:.2f} | {b('compute')} | {b('h2d')} | {b('d2h')} | "
                         f"{b('gpu_union')} | {b('bcrt:pack')} | {b('bcrt:unpack')} | {b('bcrt:wait')} |")
    if "counters" in res:
        names = sorted({c for v in res["counters"].values() for c in v})
        lines += ["", "| kernel | " + " | ".join(names) + " |", "|---|" + "---:|" * len(names)]
        for k, v in res["counters"].items():
            lines.append(f"| {k} | " + " | ".join(f"{v.get(c, 0):.4g}" for c in names) + " |")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------------------
# rocprofv3 wrapper
# ---------------------------------------------------------------------------


def run(out: Path, program: list[str], pmc: list[str] | None, timeout: float) -> int:
    out = out.resolve()
    out.mkdir(parents=True, exist_ok=True)
    env = dict(os.environ, BCRT_TRACE="1", TMPDIR="/tmp")
    cmd = ["rocprofv3"]
    if pmc:
        cmd += ["--pmc", *pmc]
    else:
        cmd += ["--kernel-trace", "--memory-copy-trace", "--marker-trace", "--stats"]
    # rocprofv3 runs from /tmp: arguments naming existing files become absolute.
    program = [str(Path(a).resolve()) if not a.startswith("-") and Path(a).exists() else a for a in program]
    cmd += ["--output-format", "csv", "-d", str(out), "-o", "trace", "--", *program]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, env=env, cwd="/tmp", timeout=timeout)
    return r.returncode


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--out", required=True, type=Path)
    r.add_argument("--pmc", nargs="+")
    r.add_argument("--timeout", type=float, default=300.0)
    r.add_argument("--markdown", type=Path)
    r.add_argument("program", nargs=argparse.REMAINDER)
    dd = sub.add_parser("decode")
    dd.add_argument("dir", type=Path)
    dd.add_argument("--markdown", type=Path)
    dd.add_argument("--scale", nargs="*", default=[], help="COUNTER=factor calibration")
    args = ap.parse_args(argv)
    if args.cmd == "run":
        prog = args.program[1:] if args.program[:1] == ["--"] else args.program
        if not prog:
            ap.error("run needs a program after --")
        rc = run(args.out, prog, args.pmc, args.timeout)
        if rc != 0:
            return rc
        res = decode(args.out)
    else:
        scale = {k: float(v) for k, v in (s.split("=", 1) for s in args.scale)}
        res = decode(args.dir, scale)
    text = markdown(res)
    if args.markdown:
        args.markdown.write_text(text)
    print(json.dumps({k: v for k, v in res.items() if k != "calls"} | {"calls": res["calls"][:8]}, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
