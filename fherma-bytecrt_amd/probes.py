"""Measurement probes of include/bytecrt_probes.h over ctypes.

``measure_hbm`` times hand-written HBM stream kernels (copy, read, write) --
the practical ceiling the product's kernels are compared with, besides the
published 8 TB/s.  ``measure_butterflies`` times the same radix-16 butterfly
rounds in the three exact modular-multiplication forms a 31-bit-prime NTT can
use on gfx950 (FP64 as the product kernels use it, 32-bit Shoup, 32-bit
Montgomery) and checks that all three produce the same residues.  Timing uses
HIP events on one torch stream; torch only holds device memory.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libbytecrt_probes.so"
EXPORTED = ("bcrtp_hbm_stream", "bcrtp_butterflies", "bcrtp_issue_overlap", "bcrtp_issue_overlap_words")
KINDS = {0: "fp64_fma", 1: "fp32_fma", 2: "int24_mad"}
FORMS = {0: "fp64", 1: "shoup32", 2: "montgomery32", 3: "fp64_tw5", 4: "fp64_tw5_pairs", 5: "fp64_fma_product"}
# Dynamic LDS per 256-thread workgroup that holds a CU to four of them: four
# waves per SIMD, the occupancy of the middle NTT kernel.
LDS_FOUR_WAVES = 40 * 1024
MODES = {0: "copy", 1: "read", 2: "write"}
_LIB = None


def load_library():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"ByteCRT probe library missing: {LIB_PATH} (run __graft_entry__.build())")
        lib = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        lib.bcrtp_hbm_stream.restype = ctypes.c_int
        lib.bcrtp_hbm_stream.argtypes = [ctypes.c_int, vp, vp, ctypes.c_size_t, vp]
        lib.bcrtp_butterflies.restype = ctypes.c_int
        lib.bcrtp_butterflies.argtypes = [ctypes.c_int, vp, ctypes.c_size_t, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32, vp]
        lib.bcrtp_issue_overlap.restype = ctypes.c_int
        lib.bcrtp_issue_overlap.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, vp,
                                            ctypes.c_size_t, vp]
        lib.bcrtp_issue_overlap_words.restype = ctypes.c_size_t
        lib.bcrtp_issue_overlap_words.argtypes = []
        _LIB = lib
    return _LIB


def hbm_stream(mode: int, src, dst, nbytes: int, stream=None):
    rc = load_library().bcrtp_hbm_stream(mode, src.data_ptr() if src is not None else None, dst.data_ptr(),
                                        nbytes, stream.cuda_stream if stream is not None else None)
    if rc != 0:
        raise RuntimeError(f"ByteCRT HBM probe mode {mode} failed ({rc})")


def butterflies(form: int, out, threads: int, p: int, tw15, rounds: int, stream=None, lds_bytes: int = 0):
    arr = (ctypes.c_uint32 * 15)(*[int(w) for w in tw15])
    rc = load_library().bcrtp_butterflies(form, out.data_ptr(), threads, p, arr, rounds, lds_bytes,
                                         stream.cuda_stream if stream is not None else None)
    if rc != 0:
        raise RuntimeError(f"ByteCRT butterfly probe form {form} failed ({rc})")


def issue_overlap(mode: int, kind: int, mfma_iters: int, valu_iters: int, out, stream=None):
    rc = load_library().bcrtp_issue_overlap(mode, kind, mfma_iters, valu_iters, out.data_ptr(), out.numel(),
                                           stream.cuda_stream if stream is not None else None)
    if rc != 0:
        raise RuntimeError(f"ByteCRT overlap probe mode {mode} kind {kind} failed ({rc})")


def _timed(torch, stream, fn, reps: int) -> float:
    """Median milliseconds of fn() over reps launches (HIP events on stream)."""
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def measure_hbm(torch, nbytes: int = 2 << 30, reps: int = 5) -> dict:
    """GB/s of the copy (read + write bytes), read and write stream kernels."""
    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.Stream()
    out = {"bytes": nbytes}
    for mode, name in MODES.items():
        ms = _timed(torch, stream, lambda: hbm_stream(mode, src, dst, nbytes, stream), reps)
        moved = nbytes * (2 if mode == 0 else 1)
        out[f"{name}_GBps"] = round(moved / (ms * 1e-3) / 1e9, 1)
    hbm_stream(0, src, dst, nbytes, stream)  # the write probe ran last: copy again and check it
    torch.cuda.synchronize()
    out["copy_exact"] = bool(torch.equal(dst, src))
    del src, dst
    return out


def measure_butterflies(torch, p: int, tw15, threads: int = 1 << 20, rounds: int = 64, reps: int = 5,
                        warmup_s: float = 0.5) -> dict:
    """Butterflies per second (x 1e12) of each form, with as many waves per
    SIMD as registers allow and (``*_4w``) with four, and whether all forms
    produced the same residues.  Every variant runs round robin, first for
    ``warmup_s`` untimed, then ``reps`` timed rounds in alternating order,
    and each rate is the median of its rounds: timing the forms one after
    the other let the clock, still rising after the work before, favour
    whichever came later (by up to 20%).  The five-operation forms need
    p - 1 = k 2^s with 3 p k < 2^53 (true for the primes of N = 65,536); for
    other primes they are left out."""
    import time

    stream = torch.cuda.Stream()
    res = {"threads": threads, "rounds": rounds, "p": p}
    k = p - 1
    while k % 2 == 0:
        k //= 2
    variants = [(form, lds) for form in FORMS if not (form in (3, 4) and 3 * p * k >= 1 << 53)
                for lds in (0, LDS_FOUR_WAVES)]
    outs = {v: torch.empty(threads * 16, dtype=torch.int32, device="cuda") for v in variants}

    def launch(v):
        butterflies(v[0], outs[v], threads, p, tw15, rounds, stream, v[1])

    t0 = time.perf_counter()
    while True:
        for v in variants:
            launch(v)
        stream.synchronize()
        if time.perf_counter() - t0 >= warmup_s:
            break
    times = {v: [] for v in variants}
    for r in range(reps):
        for v in (variants if r % 2 == 0 else variants[::-1]):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            launch(v)
            b.record(stream)
            b.synchronize()
            times[v].append(a.elapsed_time(b))
    for (form, lds), ts in times.items():
        ts.sort()
        ms = ts[len(ts) // 2]
        res[f"{FORMS[form]}{'_4w' if lds else ''}_Tbfly_per_s"] = round(threads * rounds * 32 / (ms * 1e-3) / 1e12, 3)
    torch.cuda.synchronize()
    ref = outs[(0, 0)]
    res["forms_agree"] = all(torch.equal(ref, o) for o in outs.values())
    return res


def measure_overlap(torch, mfma_iters: int = 4000, valu_iters=None, reps: int = 5) -> dict:
    """Milliseconds of MFMA waves alone, VALU waves alone and both together on
    the same SIMDs, per VALU kind, and overlap = (alone sum - both) / smaller
    alone (1: the smaller hides completely, 0: they take turns).  Checks that
    running together changes no wave's results."""
    valu_iters = valu_iters or {0: 16000, 1: 16000, 2: 16000}
    words = int(load_library().bcrtp_issue_overlap_words())
    stream = torch.cuda.Stream()
    res = {"mfma_iters": mfma_iters, "valu_iters": valu_iters}
    agree = True
    for kind, name in KINDS.items():
        outs, ms = {}, {}
        for mode in (1, 2, 3):
            out = torch.zeros(words, dtype=torch.int32, device="cuda")
            ms[mode] = _timed(torch, stream,
                              lambda: issue_overlap(mode, kind, mfma_iters, valu_iters[kind], out, stream), reps)
            outs[mode] = out.view(-1, 512)
        torch.cuda.synchronize()
        agree &= bool(torch.equal(outs[3][:, :256], outs[1][:, :256]) and
                      torch.equal(outs[3][:, 256:], outs[2][:, 256:]))
        small = min(ms[1], ms[2])
        res[name] = {"mfma_ms": round(ms[1], 4), "valu_ms": round(ms[2], 4), "both_ms": round(ms[3], 4),
                     "overlap": round((ms[1] + ms[2] - ms[3]) / small, 3)}
    res["results_agree"] = agree
    return res
