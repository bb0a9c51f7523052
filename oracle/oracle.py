"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker.  The product path never does.

emv()        full EMV table, records and cost (restates reader-el.c:1964-2139)
sssp()       single-source distances + smallest-id predecessors (1762-1880)
traceback()  depth-first record walk (1907-1944)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = ctypes.POINTER(ctypes.c_int64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.oracle_emv.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,
                                 _i64p, _i64p, _i64p, ctypes.c_int]
        L.oracle_emv.restype = ctypes.c_int
        L.oracle_sssp.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64, _i64p,
                                  _i64p]
        L.oracle_sssp.restype = ctypes.c_int
        L.oracle_traceback.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, ctypes.c_int64,
                                       ctypes.c_int64, _i64p, ctypes.c_int64]
        L.oracle_traceback.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_i64p)


def emv(n, edges, kk, threads=0):
    """Returns (cost, f, code) with f, code of shape (2^k, n) int64."""
    edges = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    kk = np.ascontiguousarray(kk, dtype=np.int64)
    k = len(kk)
    m = len(edges) // 3
    f = np.empty((1 << k) * n, dtype=np.int64)
    b = np.empty((1 << k) * n, dtype=np.int64)
    cost = ctypes.c_int64(0)
    rc = lib().oracle_emv(n, m, _p(edges), k, _p(kk), _p(f), _p(b), ctypes.byref(cost), threads)
    if rc != 0:
        raise RuntimeError(f"oracle_emv failed ({rc})")
    return cost.value, f.reshape(1 << k, n), b.reshape(1 << k, n)


def sssp(n, edges, s):
    edges = np.ascontiguousarray(edges, dtype=np.int64).reshape(-1)
    m = len(edges) // 3
    d = np.empty(n, dtype=np.int64)
    p = np.empty(n, dtype=np.int64)
    rc = lib().oracle_sssp(n, m, _p(edges), s, _p(d), _p(p))
    if rc != 0:
        raise RuntimeError(f"oracle_sssp failed ({rc})")
    return d, p


def traceback(n, k, code, q, C):
    code = np.ascontiguousarray(code, dtype=np.int64).reshape(-1)
    cap = 2 * (n + k) + 64
    pairs = np.empty(2 * cap, dtype=np.int64)
    cnt = lib().oracle_traceback(n, k, _p(code), q, C, _p(pairs), cap)
    if cnt < 0:
        raise RuntimeError("oracle_traceback overflow")
    return pairs[: 2 * cnt].reshape(cnt, 2)
