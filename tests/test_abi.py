"""CPU-side checks of the product: the C ABI library loads and exports what
include/emv_hip.h declares, its argument checks return status codes without
a device, and the C host (emv-reader) keeps the reference's input and
root-build lines (verify.py / report.py regexes) and fails loudly when no
GPU is present."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import emv
from golden_data import load_instance
from stputil import write_stp

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "steiner-edge-linear_amd")
HEADERS = [os.path.join(REPO, "include", h) for h in os.listdir(os.path.join(REPO, "include"))
           if h.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(emv\w*)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["emv_open", "emv_close", "emv_load_graph", "emv_load_edges", "emv_solve",
                 "emv_read_rows", "emv_sssp", "emv_result_free", "emv_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(emv.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_result_struct_layout_matches_header():
    # the ctypes mirror and the C struct agree on size (offsets follow)
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "emv_hip.h"
int main(void){ printf("%zu %zu %zu\n", sizeof(emv_result), offsetof(emv_result, pairs),
                       offsetof(emv_result, merge_alg_bytes)); return 0; }
"""
    tmp = os.path.join("/tmp", f"emv_layout_{os.getpid()}")
    with open(tmp + ".c", "w") as fh:
        fh.write(src)
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", tmp, tmp + ".c"], check=True)
    out = subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()
    os.remove(tmp)
    os.remove(tmp + ".c")
    assert int(out[0]) == ctypes.sizeof(emv.EmvResult)
    assert int(out[1]) == emv.EmvResult.pairs.offset
    assert int(out[2]) == emv.EmvResult.merge_alg_bytes.offset


@pytest.mark.parametrize("struct,header,fields", [
    ("emv_certificate", "emv_hip.h", ["cells", "bad_record", "first_v", "ms"]),
    ("emv_request_info", "emv_engine.h", ["cost", "npairs", "reduced_cost", "groups", "root"]),
    ("emv_batch_stats", "emv_engine.h", ["relax_rounds", "graph_fallback"]),
    ("emv_engine_params", "emv_engine.h", ["merge_rows_per_wave", "graph_round_margin",
                                           "dense_rounds", "split_items"]),
])
def test_struct_layouts_match_headers(struct, header, fields):
    """Every ctypes mirror agrees with its C struct in size and field offsets."""
    mirror = {"emv_certificate": emv.Certificate, "emv_request_info": emv.RequestInfo,
              "emv_batch_stats": emv.BatchStats, "emv_engine_params": emv.EngineParams}[struct]
    prints = " ".join(f'printf("%zu ", offsetof({struct}, {f}));' for f in fields)
    src = (f'#include <stdio.h>\n#include <stddef.h>\n#include "{header}"\n'
           f'int main(void){{ printf("%zu ", sizeof({struct})); {prints} return 0; }}\n')
    tmp = os.path.join("/tmp", f"emv_layout_{struct}_{os.getpid()}")
    with open(tmp + ".c", "w") as fh:
        fh.write(src)
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", tmp, tmp + ".c"], check=True)
    out = [int(x) for x in subprocess.run([tmp], capture_output=True, text=True,
                                          check=True).stdout.split()]
    os.remove(tmp)
    os.remove(tmp + ".c")
    assert out[0] == ctypes.sizeof(mirror)
    assert out[1:] == [getattr(mirror, f).offset for f in fields]


def test_status_codes_without_device():
    lib = emv.lib()
    h = ctypes.c_void_p()
    st = lib.emv_open(ctypes.byref(h), 0)
    if st == 0:
        lib.emv_close(h)
        pytest.skip("a HIP device is present")
    assert st in (-7, -2)
    assert lib.emv_load_edges(None, 4, 0, None) == -1
    assert lib.emv_solve(None, 2, None, 0, None) == -1
    assert lib.emv_solve_footprint(None, 3, 0) == -1
    assert lib.emv_certify(None, None) == -1
    assert lib.emv_certify_table(None, 2, None, None, None, None) == -1
    eng = emv._bind_engine()
    assert eng.emv_server_cancel(None, 0) == -1
    assert eng.emv_engine_fail_alloc(None, 0) == -1


INPUT_RE = re.compile(r"(.*)n = (.*), m = (.*), k = (.*), cost = (.*) \[(.*)ms\] {peak:(.*)GiB} "
                      r"{curr:(.*)GiB}", re.M | re.I)
ROOT_RE = re.compile(r"root build:(.*)\[zero:(.*)ms\] \[pos:(.*)ms\] \[adj:(.*)ms\] \[term:(.*)ms\] "
                     r"done. \[(.*)ms\] {peak:(.*)GiB} {curr:(.*)GiB}")
BUILD_RE = re.compile(r"build:(.*), (.*)")


def run_reader(args, **kw):
    return subprocess.run([os.path.join(PKG, "emv-reader")] + args, capture_output=True, text=True,
                          timeout=60, **kw)


def test_reader_lines_without_command(tmp_path):
    inst = load_instance("small-instances/B/b01")
    p = str(tmp_path / "b01.stp")
    write_stp(p, inst)
    out = run_reader(["-in", p])
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    inp = [x for x in lines if x.startswith("input")][0]
    g = INPUT_RE.search(inp)
    assert [g.group(i).strip() for i in (2, 3, 4)] == ["50", "63", "9"]
    assert ROOT_RE.search([x for x in lines if x.startswith("root build")][0])
    assert [x for x in lines if x.startswith("terminals")][0] == \
        "terminals: 48 49 22 35 27 12 37 34 24"
    assert BUILD_RE.search([x for x in lines if x.startswith("build")][0])
    for prefix in ["invoked as:", "random seed =", "command:", "command done", "grand total",
                   "host:", "list solution:", "num threads:", "compiler:"]:
        assert any(x.startswith(prefix) for x in lines), prefix


def test_reader_stdin_and_seed(tmp_path):
    inst = load_instance("small-instances/SP/se03")
    p = str(tmp_path / "s.stp")
    write_stp(p, inst)
    with open(p) as fh:
        out = run_reader(["-seed", "5"], stdin=fh)
    assert out.returncode == 0
    assert "no input file specified, defaulting to stdin" in out.stdout
    assert "random seed = 5" in out.stdout


@pytest.mark.parametrize("text,msg", [
    ("section graph\nsection graph\n", "nested sections"),
    ("section bogus\nend\n", "invalid section"),
    ("end\n", "no section to end"),
    ("section graph\nnodes 3\nedges 1\ne 1 9 2\nend\n", "outside"),
    ("section graph\nnodes 3\nedges 2\ne 1 2 2\nend\nsection terminals\nterminals 1\nt 1\nend\n",
     "edge count"),
])
def test_reader_malformed_input_errors(tmp_path, text, msg):
    p = str(tmp_path / "bad.stp")
    with open(p, "w") as fh:
        fh.write(text)
    out = run_reader(["-in", p, "-el"])
    assert out.returncode != 0
    assert "ERROR" in out.stderr and msg in out.stderr


def test_reader_fails_loudly_without_gpu(tmp_path):
    lib = emv.lib()
    h = ctypes.c_void_p()
    if lib.emv_open(ctypes.byref(h), 0) == 0:
        lib.emv_close(h)
        pytest.skip("a HIP device is present")
    inst = load_instance("small-instances/B/b01")
    p = str(tmp_path / "b01.stp")
    write_stp(p, inst)
    out = run_reader(["-in", p, "-el"])
    assert out.returncode != 0 and "ERROR" in out.stderr and "emv_open" in out.stderr


def test_missing_file_and_args():
    out = run_reader(["-in"])
    assert out.returncode != 0 and "file name missing" in out.stderr
    out = run_reader(["-in", "/nonexistent.stp"])
    assert out.returncode != 0 and "unable to open" in out.stderr
    out = run_reader(["-h"])
    assert out.returncode == 0 and "usage" in out.stdout


def test_python_reduction_matches_host_reduction_rule():
    e = np.array([[0, 1, 5], [1, 2, 7]])
    n2, red, kk, M = emv.group_reduce(3, e, [[0], [2]])
    assert M == 13 and n2 == 5 and list(kk) == [3, 4]
    assert red.tolist()[-2:] == [[3, 0, 13], [4, 2, 13]]


def _table_bytes(k, n, cost, f, code):
    hdr = bytearray(64)
    hdr[0:8] = b"EMVTABLE"
    hdr[8:16] = np.array([1, 8], dtype="<u4").tobytes()
    hdr[16:40] = np.array([k, n, 2 ** 63 - 1], dtype="<i8").tobytes()
    hdr[40:44] = np.array([0 if code is None else 1], dtype="<u4").tobytes()
    hdr[48:56] = np.array([cost], dtype="<i8").tobytes()
    body = np.asarray(f, dtype="<i8").tobytes()
    if code is not None:
        body += np.asarray(code, dtype="<i4").tobytes()
    return bytes(hdr) + body


def test_table_file_reader_follows_the_header(tmp_path):
    """read_table_file parses the documented layout (64-byte little-endian
    header, int64 rows, optional int32 records) of the oracle's table, and
    rejects a wrong magic and a truncated file."""
    import oracle

    n, k = 30, 4
    e = emv.regular_graph(n, 3, 50, 5)
    kk = emv.terminals(n, k, 5)
    cost, f, code = oracle.emv(n, e, kk)
    for rec in (None, code):
        p = tmp_path / "t.emvt"
        p.write_bytes(_table_bytes(k, n, cost, f, rec))
        t = emv.read_table_file(p)
        assert (t["k"], t["n"], t["inf"], t["cost"]) == (k, n, 2 ** 63 - 1, cost)
        np.testing.assert_array_equal(t["f"], np.asarray(f).reshape(1 << k, n))
        if rec is None:
            assert t["code"] is None
        else:
            np.testing.assert_array_equal(t["code"], np.asarray(code).reshape(1 << k, n))
    bad = tmp_path / "bad.emvt"
    bad.write_bytes(b"EMVTABLX" + _table_bytes(k, n, cost, f, None)[8:])
    with pytest.raises(ValueError):
        emv.read_table_file(bad)
    short = tmp_path / "short.emvt"
    short.write_bytes(_table_bytes(k, n, cost, f, None)[:-8])
    with pytest.raises(ValueError):
        emv.read_table_file(short)
