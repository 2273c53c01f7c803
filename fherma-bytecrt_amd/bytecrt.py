"""Python host mirror of the ByteCRT boundary over the gfx950 C ABI.

The reference boundary is three C++ calls (native-batched/fherma.h:47-49):
``fherma_init(Point)``, ``fherma_run(state, Inputs) -> Outputs`` and
``fherma_free(state)``.  This module keeps those names, argument meanings and
error behaviour (``RuntimeError`` where the reference throws
``std::runtime_error``, ``ValueError`` for ``std::invalid_argument``; messages
start with "ByteCRT") over numpy arrays, and adds the device-resident entry
points of include/bytecrt.h for torch tensors already in HBM.

There is no CPU fallback: if libbytecrt.so is missing or no GPU is present,
every compute call raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libbytecrt.so"

BCRT_OK, BCRT_EINVAL, BCRT_ERUNTIME = 0, 1, 2
STAGE_ENCODE, STAGE_FWD_COLS, STAGE_MIDDLE, STAGE_INV_COLS, STAGE_CRT = 1, 2, 4, 8, 16
STAGE_ALL = 31

EXPORTED = (
    "bcrt_create", "bcrt_destroy", "bcrt_run", "bcrt_run_sink", "bcrt_run_device", "bcrt_run_device_graph",
    "bcrt_synchronize", "bcrt_last_error", "bcrt_geometry", "bcrt_plan_tables",
    "bcrt_run_stages", "bcrt_planes_words", "bcrt_mfma_probe", "bcrt_plan_host", "bcrt_set_pipeline",
    "bcrt_set_encode_tile", "bcrt_set_fusion",
    "bcrt_mix_create", "bcrt_mix_run", "bcrt_mix_last_error", "bcrt_mix_destroy", "bcrt_mix_set_schedule",
    "bcrt_mix_chunks",
    "bcrt_server_create", "bcrt_server_submit", "bcrt_server_wait", "bcrt_server_stats",
    "bcrt_server_last_error", "bcrt_server_destroy", "bcrt_run_device_noalloc", "bcrt_set_kernel_flags",
    "bcrt_near_power_gap",
)

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_LIB = None


class _Group(ctypes.Structure):
    """struct bcrt_group of bytecrt.h."""
    _fields_ = [("engine", ctypes.c_void_p), ("d_a", ctypes.c_void_p), ("d_b", ctypes.c_void_p),
                ("d_c", ctypes.c_void_p), ("batch", ctypes.c_uint32)]


def load_library():
    """Load libbytecrt.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(f"ByteCRT native library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = ctypes.c_void_p
    lib.bcrt_create.restype = ctypes.c_int
    lib.bcrt_create.argtypes = [ctypes.POINTER(vp), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_uint32, _u32p, ctypes.c_char_p, ctypes.c_size_t]
    lib.bcrt_destroy.argtypes = [vp]
    lib.bcrt_run.restype = ctypes.c_int
    lib.bcrt_run.argtypes = [vp, _u32p, _u32p, _u32p, ctypes.c_uint32]
    for name in ("bcrt_run_device", "bcrt_run_device_graph", "bcrt_run_device_noalloc"):
        f = getattr(lib, name)
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, vp, vp, ctypes.c_uint32, vp]
    lib.bcrt_synchronize.restype = ctypes.c_int
    lib.bcrt_synchronize.argtypes = [vp]
    lib.bcrt_last_error.restype = ctypes.c_char_p
    lib.bcrt_last_error.argtypes = [vp]
    lib.bcrt_geometry.restype = ctypes.c_int
    lib.bcrt_geometry.argtypes = [vp, _u32p]
    lib.bcrt_plan_tables.restype = ctypes.c_int
    lib.bcrt_plan_tables.argtypes = [vp, _u32p, _u32p, _u32p, _u32p]
    lib.bcrt_run_stages.restype = ctypes.c_int
    lib.bcrt_run_stages.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, vp, ctypes.c_uint32, vp]
    lib.bcrt_planes_words.restype = ctypes.c_size_t
    lib.bcrt_planes_words.argtypes = [vp, ctypes.c_uint32]
    lib.bcrt_mfma_probe.restype = ctypes.c_int
    lib.bcrt_mfma_probe.argtypes = [vp, vp, vp]
    lib.bcrt_set_pipeline.restype = ctypes.c_int
    lib.bcrt_set_pipeline.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    lib.bcrt_set_encode_tile.restype = ctypes.c_int
    lib.bcrt_set_encode_tile.argtypes = [vp, ctypes.c_uint32]
    lib.bcrt_set_fusion.restype = ctypes.c_int
    lib.bcrt_set_fusion.argtypes = [vp, ctypes.c_uint32]
    lib.bcrt_set_kernel_flags.restype = ctypes.c_int
    lib.bcrt_set_kernel_flags.argtypes = [vp, ctypes.c_uint32]
    lib.bcrt_near_power_gap.restype = ctypes.c_int
    lib.bcrt_near_power_gap.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]
    lib.bcrt_plan_host.restype = ctypes.c_int
    lib.bcrt_plan_host.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                   _u32p, _u32p, _u32p, _u32p, ctypes.c_char_p, ctypes.c_size_t]
    lib.bcrt_mix_create.restype = ctypes.c_int
    lib.bcrt_mix_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(_Group), ctypes.c_uint32,
                                    ctypes.c_char_p, ctypes.c_size_t]
    lib.bcrt_mix_run.restype = ctypes.c_int
    lib.bcrt_mix_run.argtypes = [vp, vp, ctypes.c_int]
    lib.bcrt_mix_last_error.restype = ctypes.c_char_p
    lib.bcrt_mix_last_error.argtypes = [vp]
    lib.bcrt_mix_destroy.argtypes = [vp]
    lib.bcrt_mix_set_schedule.restype = ctypes.c_int
    lib.bcrt_mix_set_schedule.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    lib.bcrt_mix_chunks.restype = ctypes.c_int
    lib.bcrt_mix_chunks.argtypes = [vp]
    lib.bcrt_server_create.restype = ctypes.c_int
    lib.bcrt_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_uint32,
                                       ctypes.c_char_p, ctypes.c_size_t]
    lib.bcrt_server_submit.restype = ctypes.c_int
    lib.bcrt_server_submit.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p,
                                       _u32p, _u32p, _u64p]
    lib.bcrt_server_wait.restype = ctypes.c_int
    lib.bcrt_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.bcrt_server_stats.restype = ctypes.c_int
    lib.bcrt_server_stats.argtypes = [vp, _u64p]
    lib.bcrt_server_last_error.restype = ctypes.c_char_p
    lib.bcrt_server_last_error.argtypes = [vp]
    lib.bcrt_server_destroy.argtypes = [vp]
    _LIB = lib
    return lib


def _raise(code: int, msg: str):
    if code == BCRT_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def limbs_of(q: int, L: int) -> np.ndarray:
    return np.array([(q >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


def _dptr(t) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return t
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("ByteCRT device entry points need contiguous GPU tensors")
    return t.data_ptr()


def _stream_handle(stream) -> int | None:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Engine:
    """One plan (N, W, q) with device workspace for up to ``max_batch`` products."""

    def __init__(self, N: int, W: int, q: int, max_batch: int = 1):
        lib = load_library()
        self.N, self.W, self.q, self.max_batch = N, W, q, max_batch
        self.L = (W + 31) // 32
        qw = limbs_of(q, self.L)
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.bcrt_create(ctypes.byref(handle), max_batch, N, W, self.L,
                             qw.ctypes.data_as(_u32p), err, 512)
        if rc != BCRT_OK:
            _raise(rc, err.value.decode())
        self._h = handle
        self.pipeline = (32, 2)  # the library's default (chunk, streams)
        self.encode_tile = min(5, (N // 256).bit_length() - 1)  # the library's default ts_log
        self.fuse_min_tiles = 256  # the library's default fusion threshold
        self.kernel_flags = 0  # the library's default kernel-level choices
        g = np.zeros(10, dtype=np.uint32)
        lib.bcrt_geometry(self._h, g.ctypes.data_as(_u32p))
        (self.N, self.W, self.L, self.P, self.N1, self.N2, self.Kenc, self.Kcrt, self.Mcrt,
         self.product_bits) = (int(v) for v in g)

    def close(self):
        if getattr(self, "_h", None):
            load_library().bcrt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def _check(self, rc: int):
        if rc != BCRT_OK:
            _raise(rc, load_library().bcrt_last_error(self._h).decode())

    # -- plan ------------------------------------------------------------------
    def plan_tables(self):
        P, L = self.P, self.L
        primes, psi, crt = (np.zeros(P, dtype=np.uint32) for _ in range(3))
        weights = np.zeros((P + 1, L), dtype=np.uint32)
        self._check(load_library().bcrt_plan_tables(
            self._h, primes.ctypes.data_as(_u32p), psi.ctypes.data_as(_u32p),
            crt.ctypes.data_as(_u32p), weights.ctypes.data_as(_u32p)))
        return dict(primes=primes, psi=psi, crt_inverse=crt, weights=weights)

    def planes_words(self, batch: int) -> int:
        return int(load_library().bcrt_planes_words(self._h, batch))

    # -- compute ---------------------------------------------------------------
    def run(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """Whole call on host arrays [B][N][L] (or [N][L])."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if a.shape != b.shape or a.size % (self.N * self.L):
            raise RuntimeError("ByteCRT input coefficient count")
        batch = a.size // (self.N * self.L)
        c = np.empty_like(a)
        self._check(load_library().bcrt_run(self._h, a.ctypes.data_as(_u32p), b.ctypes.data_as(_u32p),
                                            c.ctypes.data_as(_u32p), batch))
        return c

    def run_device(self, a, b, c, batch: int, stream=None, graph: bool = False):
        """Enqueue on device tensors (no sync).  graph=True replays a captured hipGraph."""
        f = load_library().bcrt_run_device_graph if graph else load_library().bcrt_run_device
        self._check(f(self._h, _dptr(a), _dptr(b), _dptr(c), batch, _stream_handle(stream)))

    def run_device_noalloc(self, a, b, c, batch: int, stream=None):
        """Enqueue on device tensors with the residue planes in the library's
        fixed 64 MiB device scratch array: nothing is allocated (no sync)."""
        self._check(load_library().bcrt_run_device_noalloc(self._h, _dptr(a), _dptr(b), _dptr(c), batch,
                                                           _stream_handle(stream)))

    def run_stages(self, mask: int, a=None, b=None, c=None, planes=None, batch: int = 1, stream=None):
        lib = load_library()
        ptr = (lambda t: None if t is None else _dptr(t))
        self._check(lib.bcrt_run_stages(self._h, mask, ptr(a), ptr(b), ptr(c), ptr(planes), batch,
                                        _stream_handle(stream)))

    def set_pipeline(self, chunk: int, streams: int):
        """Split batches above `chunk` products over `streams` worker streams."""
        self._check(load_library().bcrt_set_pipeline(self._h, chunk, streams))
        self.pipeline = (chunk, streams)

    def set_encode_tile(self, ts_log: int):
        """Encode tile of 2^ts_log consecutive s by 128 >> ts_log consecutive r."""
        self._check(load_library().bcrt_set_encode_tile(self._h, ts_log))
        self.encode_tile = ts_log

    FUSE_NEVER = 0xFFFFFFFF

    def set_fusion(self, min_tiles: int):
        """Fuse encode with the forward columns from `min_tiles` workgroups
        ((N/256) * 2 * batch; 0 = whenever N >= 16,384, FUSE_NEVER = never)."""
        self._check(load_library().bcrt_set_fusion(self._h, min_tiles))
        self.fuse_min_tiles = min_tiles

    FLAG_NO_XCD, FLAG_NO_PRIO, FLAG_WEIGHTS_L2, FLAG_CRT_WIDE, FLAG_ENC_ONE_CHAIN = 1, 2, 4, 8, 16
    FLAG_GENERIC_Q, FLAG_MULMOD_TW6 = 32, 64

    def set_kernel_flags(self, flags: int):
        """Kernel-level choices (bytecrt.h bcrt_set_kernel_flags): FLAG_NO_XCD
        plain workgroup order, FLAG_NO_PRIO no raised MFMA issue priority in the
        fused encode, FLAG_WEIGHTS_L2 fused-encode weights from L2,
        FLAG_CRT_WIDE the CRT built for five workgroups per CU (weights one
        k-step ahead), FLAG_ENC_ONE_CHAIN the separate encode with one MFMA
        chain per wave at four workgroups per CU, FLAG_GENERIC_Q the CRT's
        generic-q reduction also for a near-power q = 2^W - gap,
        FLAG_MULMOD_TW6 the NTT kernels' six-operation twiddle products."""
        self._check(load_library().bcrt_set_kernel_flags(self._h, flags))
        self.kernel_flags = flags

    def apply_tuning(self, batch: int, table: dict | None = None) -> bool:
        """Use the pipelining the autotuner recorded for this geometry and batch
        (tune.py's table), if any; returns whether one was found."""
        from . import tune
        return tune.apply(self, batch, table)

    def synchronize(self):
        self._check(load_library().bcrt_synchronize(self._h))


class Mix:
    """Products of several engines (different N, W, q) run as one launch per
    stage and geometry class under the default schedule (set_schedule changes
    it).  ``groups`` is a list of (engine, d_a, d_b, d_c, batch)."""

    def __init__(self, groups):
        lib = load_library()
        arr = (_Group * len(groups))()
        for i, (eng, a, b, c, batch) in enumerate(groups):
            arr[i] = _Group(eng._h.value, _dptr(a), _dptr(b), _dptr(c), batch)
        # The library defers an engine's destruction while a mix holds it; the
        # references keep the Python objects (and their handles) alive too.
        self._engines = [g[0] for g in groups]
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.bcrt_mix_create(ctypes.byref(handle), arr, len(groups), err, 512)
        if rc != BCRT_OK:
            _raise(rc, err.value.decode())
        self._h = handle
        self.schedule = (1, 1, 0)

    MERGE_CLASSES = 1
    PERSISTENT = 2

    def run(self, stream=None, graph: bool = False):
        lib = load_library()
        rc = lib.bcrt_mix_run(self._h, _stream_handle(stream), 1 if graph else 0)
        if rc != BCRT_OK:
            _raise(rc, lib.bcrt_mix_last_error(self._h).decode())

    def set_schedule(self, chunks: int = 1, streams: int = 1, flags: int = 0):
        """Chunks per geometry class, worker streams (1: all chunks on the
        caller's stream) and flags (MERGE_CLASSES, PERSISTENT), each count in
        1..16; results are unchanged."""
        lib = load_library()
        rc = lib.bcrt_mix_set_schedule(self._h, chunks, streams, flags)
        if rc != BCRT_OK:
            _raise(rc, lib.bcrt_mix_last_error(self._h).decode())
        self.schedule = (chunks, streams, flags)

    @property
    def chunks(self) -> int:
        return int(load_library().bcrt_mix_chunks(self._h))

    def close(self):
        if getattr(self, "_h", None):
            load_library().bcrt_mix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


class Server:
    """Continuous-batching product server over host buffers.

    ``submit`` returns a ticket; ``wait`` blocks until that product is in the
    output array and returns its submit-to-completion latency in seconds.
    """

    def __init__(self, hbm_budget_bytes: int = 8 << 30, max_wave: int = 128):
        lib = load_library()
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.bcrt_server_create(ctypes.byref(handle), hbm_budget_bytes, max_wave, err, 512)
        if rc != BCRT_OK:
            _raise(rc, err.value.decode())
        self._h = handle
        self._live = {}  # ticket -> arrays kept alive until waited on

    def _check(self, rc: int):
        if rc != BCRT_OK:
            _raise(rc, load_library().bcrt_server_last_error(self._h).decode())

    def submit(self, N: int, W: int, q: int, a: np.ndarray, b: np.ndarray, c: np.ndarray | None = None):
        L = (W + 31) // 32
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if a.size != N * L or b.size != N * L:
            raise RuntimeError("ByteCRT input coefficient count")
        if c is None:
            c = np.empty((N, L), dtype=np.uint32)
        elif c.dtype != np.uint32 or not c.flags.c_contiguous or c.size != N * L:
            raise ValueError("ByteCRT server output must be a contiguous uint32 array of N*L words")
        qw = limbs_of(q, L)
        t = ctypes.c_uint64()
        self._check(load_library().bcrt_server_submit(
            self._h, N, W, L, qw.ctypes.data_as(_u32p), a.ctypes.data_as(_u32p), b.ctypes.data_as(_u32p),
            c.ctypes.data_as(_u32p), ctypes.byref(t)))
        self._live[t.value] = (a, b, c)
        return t.value

    def wait(self, ticket: int):
        """(result array, latency seconds) of a submitted product."""
        lat = ctypes.c_double()
        rc = load_library().bcrt_server_wait(self._h, ticket, ctypes.byref(lat))
        _, _, c = self._live.pop(ticket)
        self._check(rc)
        return c, lat.value

    def stats(self) -> dict:
        out = np.zeros(8, dtype=np.uint64)
        self._check(load_library().bcrt_server_stats(self._h, out.ctypes.data_as(_u64p)))
        keys = ("completed", "waves", "max_wave", "plans", "evictions", "table_bytes", "table_cap", "arena_bytes")
        return {k: int(v) for k, v in zip(keys, out)}

    def close(self):
        if getattr(self, "_h", None):
            load_library().bcrt_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


def plan_host(N: int, W: int, q: int) -> dict:
    """Host-only plan of the product library (no GPU needed)."""
    lib = load_library()
    L = (W + 31) // 32
    qw = limbs_of(q, L)
    geo = np.zeros(10, dtype=np.uint32)
    err = ctypes.create_string_buffer(512)
    rc = lib.bcrt_plan_host(N, W, L, qw.ctypes.data_as(_u32p), geo.ctypes.data_as(_u32p),
                            None, None, None, None, err, 512)
    if rc != BCRT_OK:
        _raise(rc, err.value.decode())
    P = int(geo[3])
    primes, psi, crt = (np.zeros(P, dtype=np.uint32) for _ in range(3))
    weights = np.zeros((P + 1, L), dtype=np.uint32)
    rc = lib.bcrt_plan_host(N, W, L, qw.ctypes.data_as(_u32p), geo.ctypes.data_as(_u32p),
                            primes.ctypes.data_as(_u32p), psi.ctypes.data_as(_u32p),
                            crt.ctypes.data_as(_u32p), weights.ctypes.data_as(_u32p), err, 512)
    if rc != BCRT_OK:
        _raise(rc, err.value.decode())
    keys = ("N", "W", "L", "P", "N1", "N2", "Kenc", "Kcrt", "Mcrt", "product_bits")
    out = {k: int(v) for k, v in zip(keys, geo)}
    out.update(primes=primes, psi=psi, crt_inverse=crt, weights=weights)
    return out


def near_power_gap(W: int, q: int) -> int:
    """2^W - q when below 2^32 (the CRT then takes its near-power
    reduction), else 0."""
    L = (W + 31) // 32
    gap = ctypes.c_uint32(0)
    qw = limbs_of(q, L)
    rc = load_library().bcrt_near_power_gap(W, L, qw.ctypes.data_as(_u32p), ctypes.byref(gap))
    if rc != BCRT_OK:
        _raise(rc, "ByteCRT near_power_gap: bad arguments")
    return int(gap.value)


def mfma_probe(a, b, c):
    rc = load_library().bcrt_mfma_probe(_dptr(a), _dptr(b), _dptr(c))
    if rc != BCRT_OK:
        _raise(rc, "ByteCRT MFMA probe failed")


# ---------------------------------------------------------------------------
# fherma.h mirror
# ---------------------------------------------------------------------------
@dataclass
class Tensor:
    shape: list = field(default_factory=list)
    data: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.uint32))

    def count(self) -> int:
        n = 1
        for d in self.shape:
            n *= int(d)
        return n


@dataclass
class Point:
    B: int = 0
    N: int = 0
    W: int = 0
    L: int = 0
    q: Tensor = field(default_factory=Tensor)


@dataclass
class Inputs:
    a: Tensor = field(default_factory=Tensor)
    b: Tensor = field(default_factory=Tensor)


@dataclass
class Outputs:
    c: Tensor = field(default_factory=Tensor)


class _State:
    def __init__(self, engine: Engine, B: int):
        self.engine, self.B = engine, B


def fherma_init(p: Point) -> _State:
    """native-batched/solve.cu:160-198 semantics."""
    if p.B == 0 or p.B > 128:
        raise RuntimeError("ByteCRT batch coverage B=1..128")
    qd = np.asarray(p.q.data, dtype=np.uint32)
    if qd.size != p.L:
        raise ValueError("ByteCRT RNS modulus limb count")
    q = int.from_bytes(qd.astype("<u4").tobytes(), "little")
    if p.L != (p.W + 31) // 32:
        raise ValueError("ByteCRT point L mismatch")
    return _State(Engine(p.N, p.W, q, max_batch=p.B), p.B)


def fherma_run(state: _State, inputs: Inputs) -> Outputs:
    """native-batched/solve.cu:200-244 semantics (whole call)."""
    e = state.engine
    total = state.B * e.N * e.L
    a = np.asarray(inputs.a.data, dtype=np.uint32).reshape(-1)
    b = np.asarray(inputs.b.data, dtype=np.uint32).reshape(-1)
    if a.size != total or b.size != total:
        raise RuntimeError("ByteCRT input coefficient count")
    c = e.run(a.reshape(state.B, e.N, e.L), b.reshape(state.B, e.N, e.L))
    return Outputs(c=Tensor(shape=[state.B, e.N, e.L], data=c.reshape(-1)))


def fherma_free(state: _State) -> None:
    state.engine.close()
