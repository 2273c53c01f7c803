"""Seeded synthetic inputs for the exact negacyclic product.

* ``modulus(W, seed)``: W random bits, then top bit and bit 0 set (odd q of
  exactly W bits).  ``near_power_modulus(W, gap)`` gives the reference's
  q = 2^W - gap form.
* ``operands(N, q, seed, ...)``: N coefficients uniform in [0, q) by rejection
  on W-bit draws, as [N][L] little-endian u32 limbs, with optional edge rows
  (0, 1, q-1, floor(q/2), q - floor(q/2)).
"""
from __future__ import annotations

import numpy as np


def limbs(x: int, L: int) -> np.ndarray:
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


def modulus(W: int, seed: int) -> int:
    rng = np.random.default_rng(np.random.SeedSequence([0x5157, W, seed]))
    L = (W + 31) // 32
    words = rng.integers(0, 1 << 32, size=L, dtype=np.uint64)
    q = int.from_bytes(words.astype("<u4").tobytes(), "little") & ((1 << W) - 1)
    return q | 1 | (1 << (W - 1))


def near_power_modulus(W: int, gap: int) -> int:
    return (1 << W) - gap


def _less_than(x: np.ndarray, qv: np.ndarray) -> np.ndarray:
    """Row-wise x < q for [n][L] limb arrays (limb L-1 most significant)."""
    diff = x != qv[None, :]
    any_diff = diff.any(axis=1)
    top = x.shape[1] - 1 - np.argmax(diff[:, ::-1], axis=1)
    rows = np.arange(x.shape[0])
    return any_diff & (x[rows, top] < qv[top])


def operands(N: int, q: int, seed: int, edge_fraction: float = 0.0, edges_first: bool = False) -> np.ndarray:
    """[N][L] u32 coefficients uniform in [0, q)."""
    W = q.bit_length()
    L = (W + 31) // 32
    qv = limbs(q, L)
    top_mask = np.uint32((1 << (W - 32 * (L - 1))) - 1)
    rng = np.random.default_rng(np.random.SeedSequence([0xC0EF, W, N, seed]))
    out = np.empty((N, L), dtype=np.uint32)
    todo = np.arange(N)
    while todo.size:
        draw = rng.integers(0, 1 << 32, size=(todo.size, L), dtype=np.uint64).astype(np.uint32)
        draw[:, L - 1] &= top_mask
        ok = _less_than(draw, qv)
        out[todo[ok]] = draw[ok]
        todo = todo[~ok]
    edge_vals = [0, 1, q - 1, q // 2, q - q // 2]
    if edge_fraction > 0:
        k = int(round(edge_fraction * N))
        idx = rng.choice(N, size=k, replace=False)
        pick = rng.integers(0, len(edge_vals), size=k)
        rows = np.stack([limbs(v, L) for v in edge_vals])
        out[idx] = rows[pick]
    if edges_first:
        for i, v in enumerate(edge_vals[: min(N, len(edge_vals))]):
            out[i] = limbs(v, L)
    return out


def constant(N: int, q: int, value: int) -> np.ndarray:
    L = (q.bit_length() + 31) // 32
    out = np.zeros((N, L), dtype=np.uint32)
    out[0] = limbs(value % q, L)
    return out


def filled(N: int, q: int, value: int) -> np.ndarray:
    L = (q.bit_length() + 31) // 32
    return np.tile(limbs(value % q, L), (N, 1))
