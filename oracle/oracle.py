"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker.  The product path
(fherma-bytecrt_amd) never imports it.

The heavy lifting is ``bytecrt_oracle.c`` (ctypes).  This file adds pure-Python
big-integer helpers that are independent of the C restatement: limb packing,
schoolbook negacyclic products for small N and single-coefficient sums for
large N (the c_k formula of a*b mod (X^N+1)).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "liboracle.so"
REF_PLAN = HERE / "_ref" / "ref_plan"

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    """Compile the oracle (and oracle/_ref when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _lib():
    global _LIB
    try:
        return _LIB
    except NameError:
        pass
    if not LIB_PATH.exists():
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(str(LIB_PATH))
    lib.orc_plan_create.restype = ctypes.c_void_p
    lib.orc_plan_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32,
                                    ctypes.c_char_p, ctypes.c_int]
    lib.orc_plan_free.argtypes = [ctypes.c_void_p]
    for name in ("orc_plan_pcount", "orc_plan_kpad", "orc_plan_product_bits"):
        getattr(lib, name).restype = ctypes.c_uint32
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    lib.orc_plan_tables.argtypes = [ctypes.c_void_p] + [_u32p] * 6
    lib.orc_plan_invp.argtypes = [ctypes.c_void_p, _f64p]
    lib.orc_plan_weights.argtypes = [ctypes.c_void_p, _u32p]
    lib.orc_plan_b7.argtypes = [ctypes.c_void_p, _u8p]
    lib.orc_plan_b4.argtypes = [ctypes.c_void_p, _u8p]
    lib.orc_prime32.restype = ctypes.c_int
    lib.orc_prime32.argtypes = [ctypes.c_uint32]
    lib.orc_encode.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_size_t, _u32p]
    lib.orc_encode_bytepower.restype = ctypes.c_uint64
    lib.orc_encode_bytepower.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_size_t, _u32p]
    lib.orc_ntt_negacyclic.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       _u32p, _u32p, _u32p]
    lib.orc_dense_input.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p]
    lib.orc_alpha_f32.restype = ctypes.c_uint32
    lib.orc_alpha_f32.argtypes = [ctypes.c_void_p, _u32p]
    lib.orc_alpha_f64.restype = ctypes.c_uint32
    lib.orc_alpha_f64.argtypes = [ctypes.c_void_p, _u32p]
    lib.orc_crt_b7_wide.restype = ctypes.c_uint32
    lib.orc_crt_b7_wide.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    lib.orc_crt_b4_wide.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    lib.orc_reduce_generic.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    lib.orc_reduce_near_power.restype = ctypes.c_int
    lib.orc_reduce_near_power.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    lib.orc_crt_terms.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_size_t, _u32p]
    lib.orc_product.restype = ctypes.c_int
    lib.orc_product.argtypes = [ctypes.c_void_p, _u32p, _u32p, _u32p]
    lib.orc_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    _LIB = lib  # noqa: PLW0603
    return lib


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_u32p)


# --------------------------------------------------------------------------
# Limb helpers (pure Python / numpy, independent of the C code)
# --------------------------------------------------------------------------
def int_to_limbs(x: int, L: int) -> np.ndarray:
    assert 0 <= x < 1 << (32 * L)
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


def limbs_to_int(v) -> int:
    return int.from_bytes(np.ascontiguousarray(v, dtype="<u4").tobytes(), "little")


def rows_to_ints(a: np.ndarray) -> list[int]:
    a = np.ascontiguousarray(a, dtype="<u4")
    L = a.shape[-1]
    raw = a.reshape(-1, L).tobytes()
    return [int.from_bytes(raw[4 * L * i:4 * L * (i + 1)], "little") for i in range(a.size // L)]


def ints_to_rows(xs, L: int) -> np.ndarray:
    out = bytearray()
    for x in xs:
        out += int(x).to_bytes(4 * L, "little")
    return np.frombuffer(bytes(out), dtype="<u4").reshape(len(xs), L).astype(np.uint32)


def schoolbook_negacyclic(a: list[int], b: list[int], q: int) -> list[int]:
    """c = a*b mod (X^N + 1, q) by the O(N^2) definition."""
    n = len(a)
    c = [0] * n
    for i, ai in enumerate(a):
        if not ai:
            continue
        for j, bj in enumerate(b):
            k = i + j
            if k < n:
                c[k] += ai * bj
            else:
                c[k - n] -= ai * bj
    return [v % q for v in c]


def negacyclic_coefficient(a: list[int], b: list[int], k: int, q: int) -> int:
    """c_k = sum_{i<=k} a_i b_{k-i} - sum_{i>k} a_i b_{N+k-i}  (mod q)."""
    n = len(a)
    s = 0
    for i in range(k + 1):
        s += a[i] * b[k - i]
    for i in range(k + 1, n):
        s -= a[i] * b[n + k - i]
    return s % q


# --------------------------------------------------------------------------
# Plan and compute wrappers
# --------------------------------------------------------------------------
class Plan:
    def __init__(self, N: int, W: int, q: int):
        self.N, self.W, self.q = N, W, q
        self.L = (W + 31) // 32
        lib = _lib()
        qw = int_to_limbs(q, self.L)
        err = ctypes.create_string_buffer(256)
        h = lib.orc_plan_create(N, W, _p32(qw), self.L, err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self._h = ctypes.c_void_p(h)
        self.P = lib.orc_plan_pcount(self._h)
        self.kpad = lib.orc_plan_kpad(self._h)
        self.product_bits = lib.orc_plan_product_bits(self._h)
        P = self.P
        arrs = [np.zeros(P, dtype=np.uint32) for _ in range(6)]
        lib.orc_plan_tables(self._h, *[_p32(x) for x in arrs])
        (self.primes, self.psi, self.omega, self.omega_inv, self.n_inv, self.crt_inv) = arrs
        self.invp = np.zeros(P, dtype=np.float64)
        lib.orc_plan_invp(self._h, self.invp.ctypes.data_as(_f64p))

    def __del__(self):
        try:
            _lib().orc_plan_free(self._h)
        except Exception:  # noqa: BLE001
            pass

    def weights(self) -> np.ndarray:
        w = np.zeros((self.P + 1, self.L), dtype=np.uint32)
        _lib().orc_plan_weights(self._h, _p32(w))
        return w

    def b7(self) -> np.ndarray:
        b = np.zeros(7 * self.L * self.kpad, dtype=np.uint8)
        _lib().orc_plan_b7(self._h, b.ctypes.data_as(_u8p))
        return b.reshape(7 * self.L, self.kpad)

    def b4(self) -> np.ndarray:
        b = np.zeros(4 * self.L * self.kpad, dtype=np.uint8)
        _lib().orc_plan_b4(self._h, b.ctypes.data_as(_u8p))
        return b.reshape(4 * self.L, self.kpad)

    def encode(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, self.L)
        out = np.zeros((self.P, x.shape[0]), dtype=np.uint32)
        _lib().orc_encode(self._h, _p32(x), x.shape[0], _p32(out))
        return out

    def encode_bytepower(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.uint32).reshape(-1, self.L)
        out = np.zeros((self.P, x.shape[0]), dtype=np.uint32)
        largest = _lib().orc_encode_bytepower(self._h, _p32(x), x.shape[0], _p32(out))
        return out, int(largest)

    def alpha_f32(self, terms: np.ndarray) -> int:
        t = np.ascontiguousarray(terms, dtype=np.uint32)
        return int(_lib().orc_alpha_f32(self._h, _p32(t)))

    def alpha_f64(self, terms: np.ndarray) -> int:
        t = np.ascontiguousarray(terms, dtype=np.uint32)
        return int(_lib().orc_alpha_f64(self._h, _p32(t)))

    def crt_b7_wide(self, terms_alpha: np.ndarray):
        t = np.ascontiguousarray(terms_alpha, dtype=np.uint32)
        z = np.zeros(self.L + 2, dtype=np.uint32)
        largest = _lib().orc_crt_b7_wide(self._h, _p32(t), _p32(z))
        return limbs_to_int(z), int(largest)

    def crt_b4_wide(self, terms_alpha: np.ndarray) -> int:
        t = np.ascontiguousarray(terms_alpha, dtype=np.uint32)
        z = np.zeros(self.L + 2, dtype=np.uint32)
        _lib().orc_crt_b4_wide(self._h, _p32(t), _p32(z))
        return limbs_to_int(z)

    def reduce_generic(self, z: int) -> int:
        zw = int_to_limbs(z, self.L + 2)
        r = np.zeros(self.L, dtype=np.uint32)
        _lib().orc_reduce_generic(self._h, _p32(zw), _p32(r))
        return limbs_to_int(r)

    def reduce_near_power(self, z: int):
        zw = int_to_limbs(z, self.L + 2)
        r = np.zeros(self.L, dtype=np.uint32)
        folds = _lib().orc_reduce_near_power(self._h, _p32(zw), _p32(r))
        return limbs_to_int(r), folds

    def crt_terms(self, terms: np.ndarray) -> np.ndarray:
        t = np.ascontiguousarray(terms, dtype=np.uint32).reshape(-1, self.P)
        out = np.zeros((t.shape[0], self.L), dtype=np.uint32)
        _lib().orc_crt_terms(self._h, _p32(t), t.shape[0], _p32(out))
        return out

    def product(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32).reshape(self.N, self.L)
        b = np.ascontiguousarray(b, dtype=np.uint32).reshape(self.N, self.L)
        c = np.zeros((self.N, self.L), dtype=np.uint32)
        if _lib().orc_product(self._h, _p32(a), _p32(b), _p32(c)) != 0:
            raise MemoryError("oracle product allocation failed")
        return c


def prime32(p: int) -> bool:
    return bool(_lib().orc_prime32(p))


def ntt_negacyclic(p: int, psi: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    c = np.zeros_like(a)
    _lib().orc_ntt_negacyclic(p, psi, a.size, _p32(a), _p32(b), _p32(c))
    return c


def dense_input(n: int, p: int, operand: int) -> np.ndarray:
    a = np.zeros(n, dtype=np.uint32)
    _lib().orc_dense_input(n, p, operand, _p32(a))
    return a


def set_threads(n: int) -> None:
    _lib().orc_set_threads(int(n))


def threads() -> int:
    return int(_lib().orc_threads())


def ref_plan(N: int, W: int, gap: int):
    """Plan tables from the reference's rns-plan.hpp compiled in place
    (oracle/_ref/ref_plan).  Returns None when that binary was not built."""
    if not REF_PLAN.exists():
        return None
    raw = subprocess.run([str(REF_PLAN), str(N), str(W), str(gap)], check=True,
                         capture_output=True).stdout
    P, kpad, L = np.frombuffer(raw[:12], dtype="<u4")
    off = 12
    out = {}
    for name in ("primes", "psi", "omega", "omega_inv", "n_inv", "crt_inv"):
        out[name] = np.frombuffer(raw[off:off + 4 * P], dtype="<u4").copy()
        off += 4 * P
    out["invp"] = np.frombuffer(raw[off:off + 8 * P], dtype="<f8").copy()
    off += 8 * P
    out["b7"] = np.frombuffer(raw[off:off + 7 * L * kpad], dtype=np.uint8).reshape(7 * L, kpad).copy()
    off += 7 * L * kpad
    out["b4"] = np.frombuffer(raw[off:off + 4 * L * kpad], dtype=np.uint8).reshape(4 * L, kpad).copy()
    out.update(P=int(P), kpad=int(kpad), L=int(L))
    return out


if os.environ.get("BYTECRT_ORACLE_THREADS"):
    try:
        set_threads(int(os.environ["BYTECRT_ORACLE_THREADS"]))
    except RuntimeError:
        pass
