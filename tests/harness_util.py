"""Writes point directories in the reference harness's wire format
(native-batched/main.cpp:3-15): manifest.json, point/q.bin, cases/NNNNNN/{a,b}.bin."""
import json
from pathlib import Path

import numpy as np


def write_point(root: Path, B: int, N: int, W: int, q: int, cases):
    """cases: list of (a, b) arrays shaped [B][N][L] u32."""
    L = (W + 31) // 32
    root.mkdir(parents=True, exist_ok=True)
    (root / "manifest.json").write_text(json.dumps({"B": B, "N": N, "W": W, "L": L, "cases": len(cases)}))
    (root / "point").mkdir(exist_ok=True)
    qw = np.array([(q >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype="<u4")
    (root / "point" / "q.bin").write_bytes(qw.tobytes())
    for i, (a, b) in enumerate(cases):
        d = root / "cases" / f"{i:06d}"
        d.mkdir(parents=True, exist_ok=True)
        (d / "a.bin").write_bytes(np.ascontiguousarray(a, dtype="<u4").tobytes())
        (d / "b.bin").write_bytes(np.ascontiguousarray(b, dtype="<u4").tobytes())


def read_answer(root: Path, i: int, B: int, N: int, L: int) -> np.ndarray:
    raw = (root / "out" / f"{i:06d}" / "c.bin").read_bytes()
    return np.frombuffer(raw, dtype="<u4").reshape(B, N, L)


def read_results(root: Path) -> dict:
    return json.loads((root / "out" / "results.json").read_text())
