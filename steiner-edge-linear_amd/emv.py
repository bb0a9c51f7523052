"""Python binding of the MI355X EMV Steiner library (ctypes over the C ABI).

Mirrors the reference's entry points for the EMV path:
  Context.load_graph(n, m, pos, adj)  <- steinerq_t hand-over (reader-el.c:2269)
  Context.solve(kk, ...)              <- erickson_monma_veinott (reader-el.c:2141)
  Context.sssp(s)                     <- dijkstra (reader-el.c:1762)
Errors raise EmvError carrying the library's status and message; a missing
libemvhip.so raises at import of the library (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemvhip.so")
GEN_PATH = os.path.join(HERE, "libemvgen.so")

LIST = 0x1
FORCE_I64 = 0x2
KEEP_TABLE = 0x4
FORCE_P64 = 0x8
WIDE_LABELS = 0x10

ERRORS = {0: "OK", -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_NOMEM", -4: "ERR_STATE",
          -5: "ERR_DISCONNECTED", -6: "ERR_RANGE", -7: "ERR_NODEV", -8: "ERR_CAPACITY",
          -9: "ERR_ADMISSION"}

_i64p = ctypes.POINTER(ctypes.c_int64)


class EmvResult(ctypes.Structure):
    _fields_ = [
        ("cost", ctypes.c_int64), ("k", ctypes.c_int32), ("table_bytes", ctypes.c_int32),
        ("merge_bits", ctypes.c_int32), ("label_bits", ctypes.c_int32), ("n", ctypes.c_int64),
        ("dp_cells", ctypes.c_int64), ("zero_ms", ctypes.c_double),
        ("singleton_ms", ctypes.c_double), ("merge_ms", ctypes.c_double),
        ("relax_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
        ("traceback_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("merge_alg_bytes", ctypes.c_double), ("relax_alg_bytes", ctypes.c_double),
        ("merge_launches", ctypes.c_int64), ("relax_rounds", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64), ("npairs", ctypes.c_int64), ("pairs", _i64p),
    ]


class EmvError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{ERRORS.get(status, status)}: {msg}")
        self.status = status


_LIB = None
_GEN = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run make -C steiner-edge-linear_amd")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.emv_open.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.emv_close.argtypes = [vp]
        L.emv_close.restype = None
        L.emv_last_error.argtypes = [vp]
        L.emv_last_error.restype = ctypes.c_char_p
        L.emv_load_graph.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p]
        L.emv_load_edges.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _i64p]
        L.emv_solve.argtypes = [vp, ctypes.c_int64, _i64p, ctypes.c_uint32,
                                ctypes.POINTER(EmvResult)]
        L.emv_read_rows.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p]
        L.emv_release_table.argtypes = [vp]
        L.emv_sssp.argtypes = [vp, ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(ctypes.c_double)]
        L.emv_result_free.argtypes = [ctypes.POINTER(EmvResult)]
        L.emv_result_free.restype = None
        L.emv_solve_footprint.argtypes = [vp, ctypes.c_int64, ctypes.c_uint32]
        L.emv_solve_footprint.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_i64p)


class Solution:
    def __init__(self, r, pairs):
        self.cost = r.cost
        self.k = r.k
        self.table_bytes = r.table_bytes
        self.merge_bits = r.merge_bits
        self.label_bits = r.label_bits
        self.dp_cells = r.dp_cells
        self.singleton_ms = r.singleton_ms
        self.merge_ms = r.merge_ms
        self.relax_ms = r.relax_ms
        self.kernel_ms = r.kernel_ms
        self.traceback_ms = r.traceback_ms
        self.total_ms = r.total_ms
        self.merge_alg_bytes = r.merge_alg_bytes
        self.relax_alg_bytes = r.relax_alg_bytes
        self.merge_launches = r.merge_launches
        self.relax_rounds = r.relax_rounds
        self.device_bytes = r.device_bytes
        self.pairs = pairs


class Context:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        st = lib().emv_open(ctypes.byref(self._h), device)
        if st != 0:
            raise EmvError(st, "emv_open failed")
        self.n = 0

    def _check(self, st):
        if st != 0:
            raise EmvError(st, lib().emv_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().emv_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_edges(self, n, edges):
        e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
        self._check(lib().emv_load_edges(self._h, n, len(e) // 3, _p(e)))
        self.n = n

    def load_graph(self, n, m, pos, adj):
        pos = np.ascontiguousarray(pos, dtype=np.int64)
        adj = np.ascontiguousarray(adj, dtype=np.int64)
        self._check(lib().emv_load_graph(self._h, n, m, _p(pos), _p(adj)))
        self.n = n

    def solve(self, kk, flags=0):
        kk = np.ascontiguousarray(kk, dtype=np.int64)
        r = EmvResult()
        self._check(lib().emv_solve(self._h, len(kk), _p(kk), flags, ctypes.byref(r)))
        pairs = None
        if flags & LIST:
            pairs = np.ctypeslib.as_array(r.pairs, shape=(2 * max(r.npairs, 1),))[: 2 * r.npairs]
            pairs = pairs.reshape(-1, 2).copy()
        sol = Solution(r, pairs)
        lib().emv_result_free(ctypes.byref(r))
        return sol

    def read_rows(self, X0, nrows, codes=True):
        f = np.empty(nrows * self.n, dtype=np.int64)
        c = np.empty(nrows * self.n, dtype=np.int64) if codes else None
        self._check(lib().emv_read_rows(self._h, X0, nrows, _p(f), _p(c) if codes else None))
        return f.reshape(nrows, self.n), (c.reshape(nrows, self.n) if codes else None)

    def release_table(self):
        self._check(lib().emv_release_table(self._h))

    def sssp(self, s):
        d = np.empty(self.n, dtype=np.int64)
        p = np.empty(self.n, dtype=np.int64)
        ms = ctypes.c_double(0)
        self._check(lib().emv_sssp(self._h, s, _p(d), _p(p), ctypes.byref(ms)))
        return d, p, ms.value

    def footprint(self, k, flags=0):
        return lib().emv_solve_footprint(self._h, k, flags)


# ------------------------------------------------------------- generator

def gen():
    global _GEN
    if _GEN is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is not built")
        G = ctypes.CDLL(GEN_PATH)
        G.emvgen_regular.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                     ctypes.c_int, _i64p]
        G.emvgen_regular.restype = ctypes.c_int64
        G.emvgen_grid.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                  _i64p]
        G.emvgen_grid.restype = ctypes.c_int64
        G.emvgen_terminals.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, _i64p]
        G.emvgen_distinct.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int,
                                      _i64p]
        G.emvgen_groups.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64,
                                    _i64p]
        G.emvgen_u64.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int64]
        G.emvgen_u64.restype = ctypes.c_uint64
        _GEN = G
    return _GEN


def regular_graph(n, d, wmax, seed, method="repair"):
    e = np.empty(3 * (n * d // 2), dtype=np.int64)
    tries = gen().emvgen_regular(n, d, wmax, seed, 0 if method == "reject" else 1, _p(e))
    if tries < 0:
        raise ValueError("bad regular graph arguments")
    return e.reshape(-1, 3)


def grid_graph(rows, cols, wmax, seed):
    e = np.empty(3 * (2 * rows * cols), dtype=np.int64)
    m = gen().emvgen_grid(rows, cols, wmax, seed, _p(e))
    if m < 0:
        raise ValueError("bad grid arguments")
    return e[: 3 * m].reshape(-1, 3)


def terminals(n, k, seed):
    kk = np.empty(k, dtype=np.int64)
    if gen().emvgen_terminals(n, k, seed, _p(kk)):
        raise ValueError("bad terminal arguments")
    return kk


def distinct(n, k, seed, stream):
    kk = np.empty(k, dtype=np.int64)
    if gen().emvgen_distinct(n, k, seed, stream, _p(kk)):
        raise ValueError("bad arguments")
    return kk


def groups(n, g, size, seed):
    mem = np.empty(g * size, dtype=np.int64)
    if gen().emvgen_groups(n, g, size, seed, _p(mem)):
        raise ValueError("bad group arguments")
    return mem.reshape(g, size)


def group_reduce(n, edges, grp):
    """Reduced instance: virtual terminals n..n+g-1 joined to their group's
    members by weight M = 1 + sum of weights (the same reduction as the
    emv-reader -groups option)."""
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 3)
    M = 1 + int(edges[:, 2].sum())
    g = len(grp)
    extra = []
    for i, mem in enumerate(grp):
        for v in mem:
            extra.append((n + i, int(v), M))
    red = np.concatenate([edges, np.asarray(extra, dtype=np.int64).reshape(-1, 3)])
    kk = np.arange(n, n + g, dtype=np.int64)
    return n + g, red, kk, M


def uniform(seed, stream, index):
    return gen().emvgen_u64(seed, stream, index)
