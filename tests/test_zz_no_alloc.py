"""emv_sssp_noalloc (no allocation, scratch in a fixed per-GPU array) against
emv_sssp, which it must equal exactly (integer distances and smallest-id
predecessors), on two calls in a row -- the second on a smaller graph, so
that what the first left in the shared array must not leak into it -- plus
the capacity refusal and an unchanged free device memory across a call.

The caller's device buffers come from the HIP runtime the library itself
uses (libamdhip64, through ctypes), not from another framework's runtime."""
import ctypes

import numpy as np
import pytest

import emv

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        emv.lib()  # the library's runtime is loaded first; the same one is reused
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]
        self.L.hipSetDevice.argtypes = [ctypes.c_int]
        assert self.L.hipSetDevice(0) == 0

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def free(self, p):
        assert self.L.hipFree(p) == 0

    def fill(self, p, arr):
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def read(self, p, n):
        out = np.empty(n, dtype=np.int64)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free_bytes(self):
        fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
        assert self.L.hipMemGetInfo(ctypes.byref(fr), ctypes.byref(tot)) == 0
        return fr.value


def _check_equal(hip, ctx, n, s, dd, pp):
    d, p, _ = ctx.sssp(s)
    junk = np.full(n, -7, dtype=np.int64)
    hip.fill(dd, junk)
    hip.fill(pp, junk)
    ctx.sssp_noalloc(s, dd, pp)
    np.testing.assert_array_equal(hip.read(dd, n), d)
    np.testing.assert_array_equal(hip.read(pp, n), p)


def test_sssp_noalloc_matches_sssp():
    hip = Hip()
    n1, n2 = 1024 * 1024, 6000
    with emv.Context(0) as ctx:
        dd, pp = hip.alloc(8 * n1), hip.alloc(8 * n1)
        try:
            # first call: a 1024 x 1024 grid (a long-hop graph)
            ctx.load_edges(n1, emv.grid_graph(1024, 1024, 1000, 77))
            print(f"case noalloc grid n={n1}", flush=True)
            _check_equal(hip, ctx, n1, 12345, dd, pp)
            # second call, smaller: a random 5-regular graph
            ctx.load_edges(n2, emv.regular_graph(n2, 5, 1000, 78))
            print(f"case noalloc regular n={n2}", flush=True)
            _check_equal(hip, ctx, n2, 17, dd, pp)
            # no allocation: free device memory is unchanged across a call
            free0 = hip.free_bytes()
            ctx.sssp_noalloc(5, dd, pp)
            assert hip.free_bytes() == free0
            # a NULL output is allowed
            ctx.sssp_noalloc(3, None, pp)
            _, p, _ = ctx.sssp(3)
            np.testing.assert_array_equal(hip.read(pp, n2), p)
        finally:
            hip.free(dd)
            hip.free(pp)


def test_sssp_noalloc_refuses_graphs_beyond_its_scratch():
    hip = Hip()
    with emv.Context(0) as ctx:
        n = 1500 * 1500  # more vertices than the 32 MiB array holds
        ctx.load_edges(n, emv.grid_graph(1500, 1500, 10, 79))
        out = hip.alloc(64)
        try:
            with pytest.raises(emv.EmvError) as ei:
                ctx.sssp_noalloc(0, out, out)
            assert ei.value.status == -8
            with pytest.raises(emv.EmvError) as ei:
                ctx.sssp_noalloc(n, out, out)
            assert ei.value.status == -1
        finally:
            hip.free(out)
