"""Build recipe for the gfx950 ByteCRT library (hipcc, no cmake).

    python fherma-bytecrt_amd/build.py

Outputs (git-ignored, shipped in-tree to the GPU host):
    fherma-bytecrt_amd/lib/libbytecrt.so      C ABI (include/bytecrt.h) + fherma.h wrapper
    fherma-bytecrt_amd/lib/libbytecrt_probes.so  measurement probes (include/bytecrt_probes.h)
    fherma-bytecrt_amd/lib/bytecrt_solution   point-directory runner (reference wire format)
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["bytecrt_kernels.hip", "bytecrt_noalloc.hip", "bytecrt_plan.cpp", "bytecrt_io.cpp", "bytecrt_trace.cpp",
           "bytecrt_engine.cpp", "bytecrt_server.cpp", "fherma_solve.cpp"]
# harness_main.cpp is linked separately into lib/bytecrt_solution.

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
# -ffp-contract=off: the exact FP64 modular arithmetic relies on every product
# and sum being rounded where written (explicit fma() calls only).
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=off"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False) -> Path:
    LIB.mkdir(exist_ok=True)
    jobs = []
    for src in SOURCES:
        s = CSRC / src
        o = LIB / (s.stem + ".o")
        cmd = [HIPCC, *COMMON, "-c", str(s), "-o", str(o)]
        if s.suffix == ".hip":
            cmd[1:1] = DEVICE
        else:
            cmd[1:1] = [f"--offload-arch={ARCH}", "-x", "hip"] if src in ("bytecrt_engine.cpp", "bytecrt_server.cpp") else []
        jobs.append((cmd, str(o)))
    # Independent translation units: compile them side by side.
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        list(pool.map(_run, [cmd for cmd, _ in jobs]))
    objs = [o for _, o in jobs]
    out = LIB / "libbytecrt.so"
    tmp = LIB / "libbytecrt.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
          "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    # Measurement probes (include/bytecrt_probes.h): a library of their own.
    probes = LIB / "libbytecrt_probes.so"
    _run([HIPCC, *DEVICE, *COMMON, "-shared", str(CSRC / "bytecrt_probes.hip"), "-o", str(LIB / "libbytecrt_probes.so.tmp")])
    os.replace(LIB / "libbytecrt_probes.so.tmp", probes)
    # The harness runner (reference wire format) links the library by rpath.
    # So does the whole-call timer (fherma_run at the harness boundary).
    for exe, src in (("bytecrt_solution", "harness_main.cpp"), ("fherma_call_bench", "fherma_call_bench.cpp")):
        _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(CSRC / src),
              f"-L{LIB}", "-lbytecrt", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", str(LIB / exe)])
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
