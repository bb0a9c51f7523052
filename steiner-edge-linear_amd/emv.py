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
TABLE32 = 0x20
MERGE_1V = 0x40
MERGE_ROWS = 0x80
MERGE_CUBE = 0x100
COUNT_VISITS = 0x200
DUMP_RECORDS = 0x1  # emv_write_table flag

ERRORS = {0: "OK", -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_NOMEM", -4: "ERR_STATE",
          -5: "ERR_DISCONNECTED", -6: "ERR_RANGE", -7: "ERR_NODEV", -8: "ERR_CAPACITY",
          -9: "ERR_ADMISSION", -10: "ERR_CANCELLED"}

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
        ("relax_blocks", ctypes.c_int64), ("unpack_ms", ctypes.c_double),
        ("singleton_rounds", ctypes.c_int64), ("block_dist_bits", ctypes.c_int64),
        ("merge_split_ms", ctypes.c_double), ("relative_label_layers", ctypes.c_int64),
        ("compact_merge_layers", ctypes.c_int64), ("relax_item_visits", ctypes.c_int64),
        ("relax_chunk_rows", ctypes.c_int64),
    ]


class Certificate(ctypes.Structure):
    """emv_certificate: cells of a table that break the EMV fixed point or
    the record rule (all zero: the table is the reference's, cell for cell)."""
    _fields_ = [("cells", ctypes.c_int64), ("bad_value", ctypes.c_int64),
                ("not_finite", ctypes.c_int64), ("bad_record", ctypes.c_int64),
                ("first_X", ctypes.c_int64), ("first_v", ctypes.c_int64), ("ms", ctypes.c_double)]

    @property
    def ok(self):
        return self.bad_value == 0 and self.not_finite == 0 and self.bad_record == 0

    def __repr__(self):
        return (f"Certificate(cells={self.cells}, bad_value={self.bad_value}, "
                f"not_finite={self.not_finite}, bad_record={self.bad_record}, "
                f"first=({self.first_X}, {self.first_v}), ms={self.ms:.1f})")


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
        L.emv_write_table.argtypes = [vp, ctypes.c_char_p, ctypes.c_uint32]
        L.emv_sssp.argtypes = [vp, ctypes.c_int64, _i64p, _i64p, ctypes.POINTER(ctypes.c_double)]
        L.emv_result_free.argtypes = [ctypes.POINTER(EmvResult)]
        L.emv_result_free.restype = None
        L.emv_solve_footprint.argtypes = [vp, ctypes.c_int64, ctypes.c_uint32]
        L.emv_solve_footprint.restype = ctypes.c_int64
        L.emv_certify.argtypes = [vp, ctypes.POINTER(Certificate)]
        L.emv_sssp_noalloc.argtypes = [vp, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.emv_certify_table.argtypes = [vp, ctypes.c_int64, _i64p, _i64p,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(Certificate)]
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
        self.zero_ms = r.zero_ms
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
        self.relax_blocks = r.relax_blocks
        self.unpack_ms = r.unpack_ms
        self.singleton_rounds = r.singleton_rounds
        self.block_dist_bits = r.block_dist_bits
        self.merge_split_ms = r.merge_split_ms
        self.relative_label_layers = r.relative_label_layers
        self.compact_merge_layers = r.compact_merge_layers
        self.relax_item_visits = r.relax_item_visits
        self.relax_chunk_rows = r.relax_chunk_rows
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

    def write_table(self, path, records=False):
        """The kept table as a canonical table file (read_table_file)."""
        self._check(lib().emv_write_table(self._h, os.fsencode(path),
                                          DUMP_RECORDS if records else 0))

    def sssp(self, s):
        d = np.empty(self.n, dtype=np.int64)
        p = np.empty(self.n, dtype=np.int64)
        ms = ctypes.c_double(0)
        self._check(lib().emv_sssp(self._h, s, _p(d), _p(p), ctypes.byref(ms)))
        return d, p, ms.value

    def sssp_noalloc(self, s, d_dist_ptr, d_pred_ptr):
        """emv_sssp into caller-owned device memory (raw device addresses,
        e.g. torch tensors' data_ptr()), with no allocation in the library."""
        self._check(lib().emv_sssp_noalloc(self._h, s, d_dist_ptr, d_pred_ptr))

    def footprint(self, k, flags=0):
        return lib().emv_solve_footprint(self._h, k, flags)

    def certify(self):
        """Certificate of the kept table (solve with KEEP_TABLE)."""
        c = Certificate()
        self._check(lib().emv_certify(self._h, ctypes.byref(c)))
        return c

    def certify_table(self, kk, f, code):
        """Certificate of a caller's table (f int64 with INF = INT64_MAX,
        code int32, both 2^k x n) for terminals kk on the loaded graph."""
        kk = np.ascontiguousarray(kk, dtype=np.int64)
        f = np.ascontiguousarray(f, dtype=np.int64)
        code = np.ascontiguousarray(code, dtype=np.int32)
        if f.size != (1 << len(kk)) * self.n or code.size != f.size:
            raise ValueError("table shape does not match 2^k x n")
        c = Certificate()
        self._check(lib().emv_certify_table(
            self._h, len(kk), _p(kk), _p(f),
            code.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(c)))
        return c


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


# ----------------------------------------------------- multi-instance engine

class EmvInstance(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64), ("edges", _i64p),
                ("k", ctypes.c_int64), ("kk", _i64p)]


class EngineParams(ctypes.Structure):
    _fields_ = [("relax_blocks_per_cu", ctypes.c_int32), ("merge_rows_per_wave", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("graph_round_margin", ctypes.c_int32),
                ("dense_rounds", ctypes.c_int32), ("split_items", ctypes.c_int32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("steps", ctypes.c_int64),
                ("relax_rounds", ctypes.c_int64), ("dp_cells", ctypes.c_int64),
                ("graph_used", ctypes.c_int32), ("graph_fallback", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("cost", ctypes.c_int64),
                ("k", ctypes.c_int32), ("stage", ctypes.c_int32), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("npairs", ctypes.c_int64),
                ("reduced_cost", ctypes.c_int64), ("groups", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("root", ctypes.c_int64)]


_ENG_BOUND = False


def _bind_engine():
    global _ENG_BOUND
    if _ENG_BOUND:
        return lib()
    L = lib()
    vp = ctypes.c_void_p
    L.emv_engine_default_params.argtypes = [ctypes.POINTER(EngineParams)]
    L.emv_engine_set_params.argtypes = [vp, ctypes.POINTER(EngineParams)]
    L.emv_engine_last_error.argtypes = [vp]
    L.emv_engine_last_error.restype = ctypes.c_char_p
    L.emv_engine_destroy.argtypes = [vp]
    L.emv_engine_destroy.restype = None
    L.emv_batch_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.POINTER(EmvInstance),
                                   ctypes.POINTER(vp)]
    L.emv_batch_solve.argtypes = [vp, ctypes.c_uint32, _i64p, ctypes.POINTER(BatchStats)]
    L.emv_server_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i64p,
                                    ctypes.c_int64, ctypes.c_int, ctypes.POINTER(vp)]
    L.emv_server_submit.argtypes = [vp, ctypes.c_int64, _i64p, ctypes.c_uint32,
                                    ctypes.POINTER(ctypes.c_int64)]
    L.emv_server_step.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int64)]
    L.emv_server_poll.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(RequestInfo)]
    L.emv_server_free_bytes.argtypes = [vp]
    L.emv_server_free_bytes.restype = ctypes.c_int64
    L.emv_request_footprint.argtypes = [vp, ctypes.c_int64]
    L.emv_request_footprint.restype = ctypes.c_int64
    L.emv_engine_read_rows.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p,
                                       _i64p]
    L.emv_engine_pairs.argtypes = [vp, ctypes.c_int64, _i64p, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_int64)]
    L.emv_engine_release.argtypes = [vp, ctypes.c_int64]
    L.emv_engine_certify.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(Certificate)]
    L.emv_server_submit_groups.argtypes = [vp, ctypes.c_int64, _i64p, _i64p, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_int64)]
    L.emv_server_group_weight.argtypes = [vp]
    L.emv_server_group_weight.restype = ctypes.c_int64
    L.emv_server_cancel.argtypes = [vp, ctypes.c_int64]
    L.emv_engine_fail_alloc.argtypes = [vp, ctypes.c_int64]
    _ENG_BOUND = True
    return L


class _Engine:
    def __init__(self):
        self._h = ctypes.c_void_p()
        self._keep = []

    def _check(self, st):
        if st != 0:
            msg = _bind_engine().emv_engine_last_error(self._h) if self._h else b"engine"
            raise EmvError(st, msg.decode())

    def close(self):
        if self._h:
            _bind_engine().emv_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, relax_blocks_per_cu=64, merge_rows_per_wave=0, use_graph=1,
                   graph_round_margin=2, dense_rounds=0, split_items=0):
        p = EngineParams(relax_blocks_per_cu, merge_rows_per_wave, use_graph, graph_round_margin,
                         dense_rounds, split_items)
        self._check(_bind_engine().emv_engine_set_params(self._h, ctypes.byref(p)))

    def read_rows(self, ticket, X0, nrows, n, codes=True):
        f = np.empty(nrows * n, dtype=np.int64)
        c = np.empty(nrows * n, dtype=np.int64) if codes else None
        self._check(_bind_engine().emv_engine_read_rows(self._h, ticket, X0, nrows, _p(f),
                                                        _p(c) if codes else None))
        return f.reshape(nrows, n), (c.reshape(nrows, n) if codes else None)

    def pairs(self, ticket):
        L = _bind_engine()
        cnt = ctypes.c_int64(0)
        self._check(L.emv_engine_pairs(self._h, ticket, None, 0, ctypes.byref(cnt)))
        out = np.empty(2 * max(cnt.value, 1), dtype=np.int64)
        self._check(L.emv_engine_pairs(self._h, ticket, _p(out), cnt.value, ctypes.byref(cnt)))
        return out[: 2 * cnt.value].reshape(-1, 2)

    def release(self, ticket):
        self._check(_bind_engine().emv_engine_release(self._h, ticket))

    def certify(self, ticket):
        """Certificate of a finished instance's kept table (KEEP_TABLE)."""
        c = Certificate()
        self._check(_bind_engine().emv_engine_certify(self._h, ticket, ctypes.byref(c)))
        return c

    def fail_alloc(self, after=0):
        """Test hook: the step allocation `after` allocations from now fails
        as if the device were out of memory (-1 disarms it)."""
        self._check(_bind_engine().emv_engine_fail_alloc(self._h, after))


class Batch(_Engine):
    """Many independent instances solved together (one merge launch, shared
    relaxation rounds and one unpack launch per DP layer)."""

    def __init__(self, instances, device=0):
        super().__init__()
        L = _bind_engine()
        arr = (EmvInstance * len(instances))()
        for i, (n, edges, kk) in enumerate(instances):
            e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
            t = np.ascontiguousarray(kk, dtype=np.int64)
            self._keep += [e, t]
            arr[i] = EmvInstance(n, len(e) // 3, _p(e), len(t), _p(t))
        st = L.emv_batch_create(device, len(instances), arr, ctypes.byref(self._h))
        if st != 0:
            msg = L.emv_engine_last_error(self._h).decode() if self._h else "emv_batch_create"
            self.close()
            raise EmvError(st, msg)
        self.count = len(instances)
        self._keep = []

    def solve(self, flags=0):
        costs = np.empty(self.count, dtype=np.int64)
        st = BatchStats()
        self._check(_bind_engine().emv_batch_solve(self._h, flags, _p(costs), ctypes.byref(st)))
        return costs, st


class Server(_Engine):
    """Continuous-batching query server on one resident graph."""

    def __init__(self, n, edges, arena_bytes=0, max_inflight=64, device=0):
        super().__init__()
        L = _bind_engine()
        e = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
        st = L.emv_server_create(device, n, len(e) // 3, _p(e), arena_bytes, max_inflight,
                                 ctypes.byref(self._h))
        if st != 0:
            msg = L.emv_engine_last_error(self._h).decode() if self._h else "emv_server_create"
            self.close()
            raise EmvError(st, msg)
        self.n = n

    def submit(self, kk, flags=0):
        t = np.ascontiguousarray(kk, dtype=np.int64)
        ticket = ctypes.c_int64(-1)
        self._check(_bind_engine().emv_server_submit(self._h, len(t), _p(t), flags,
                                                     ctypes.byref(ticket)))
        return ticket.value

    def submit_groups(self, groups, flags=0):
        """Group Steiner query: groups = list of member lists (2..30 groups)."""
        off = np.zeros(len(groups) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(g) for g in groups])
        mem = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.int64).reshape(-1)
                                                   for g in groups]), dtype=np.int64)
        ticket = ctypes.c_int64(-1)
        self._check(_bind_engine().emv_server_submit_groups(self._h, len(groups), _p(off),
                                                            _p(mem), flags,
                                                            ctypes.byref(ticket)))
        return ticket.value

    def group_weight(self):
        """M = 1 + sum of the resident graph's weights (the reduction's edge weight)."""
        return _bind_engine().emv_server_group_weight(self._h)

    def cancel(self, ticket):
        """Cancels a queued or running request (it ends with ERR_CANCELLED)."""
        self._check(_bind_engine().emv_server_cancel(self._h, ticket))

    def step(self):
        r, q = ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(_bind_engine().emv_server_step(self._h, ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def poll(self, ticket):
        info = RequestInfo()
        self._check(_bind_engine().emv_server_poll(self._h, ticket, ctypes.byref(info)))
        return info

    def free_bytes(self):
        return _bind_engine().emv_server_free_bytes(self._h)

    def footprint(self, k):
        return _bind_engine().emv_request_footprint(self._h, k)


TABLE_MAGIC = b"EMVTABLE"
TABLE_HEADER = 64


def read_table_file(path):
    """A canonical table file (emv_write_table): dict with k, n, inf, cost,
    f (2^k x n int64, INF = INT64_MAX) and code (2^k x n int32, or None)."""
    with open(path, "rb") as fh:
        hdr = fh.read(TABLE_HEADER)
    if len(hdr) != TABLE_HEADER or hdr[:8] != TABLE_MAGIC:
        raise ValueError(f"{path}: not a table file")
    version, cell_bytes = np.frombuffer(hdr[8:16], dtype="<u4")
    k, n, inf = np.frombuffer(hdr[16:40], dtype="<i8")
    records = int(np.frombuffer(hdr[40:44], dtype="<u4")[0])
    cost = int(np.frombuffer(hdr[48:56], dtype="<i8")[0])
    if version != 1 or cell_bytes != 8:
        raise ValueError(f"{path}: version {version}, cell bytes {cell_bytes} not supported")
    rows, n = 1 << int(k), int(n)
    f = np.fromfile(path, dtype="<i8", count=rows * n, offset=TABLE_HEADER).reshape(rows, n)
    code = None
    if records:
        code = np.fromfile(path, dtype="<i4", count=rows * n,
                           offset=TABLE_HEADER + 8 * rows * n).reshape(rows, n)
    if f.size != rows * n or (records and code.size != rows * n):
        raise ValueError(f"{path}: truncated")
    return dict(k=int(k), n=n, inf=int(inf), cost=cost, f=f, code=code)
