"""GPU parity: the HIP path (through the C ABI) against the reference's golden
vectors and the pinned CPU oracle.  Integer work, so everything is bit-exact:
every DP cell, the cost, and the record codes (the GPU and the oracle share
the deterministic tie rule; the reference's own record kinds are matched via
digests).  Listed trees are validated, not compared, since the reference's
edge-record choice depends on its heap order."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import emv
import oracle
from golden_data import CFG1, META, NAMES, digest, golden_table, kinds, load_instance
from stputil import INF, Instance, adjacency, check_listed_tree, dijkstra, write_stp

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
READER = os.path.join(REPO, "steiner-edge-linear_amd", "emv-reader")


def solve_full(ctx, inst, flags=0):
    ctx.load_edges(inst.n, inst.edges)
    sol = ctx.solve(inst.terminals, emv.LIST | emv.KEEP_TABLE | flags)
    f, code = ctx.read_rows(0, 1 << inst.k)
    ctx.release_table()
    return sol, f, code


@pytest.mark.parametrize("name", NAMES)
def test_testset_table_bit_exact(gpu_ctx, name):
    inst = load_instance(name)
    mt = META[name]
    print(f"case {name} n={inst.n} k={inst.k}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst)
    assert sol.cost == mt["ref_cost"] == inst.cost
    assert digest(f) == mt["f_sha256"]
    assert digest(kinds(code, inst.k)) == mt["kind_sha256"]
    ocost, of, ocode = oracle.emv(inst.n, inst.edges, inst.terminals)
    np.testing.assert_array_equal(code, ocode)
    C = (1 << (inst.k - 1)) - 1
    opairs = oracle.traceback(inst.n, inst.k, ocode, int(inst.terminals[-1]), C)
    np.testing.assert_array_equal(sol.pairs, opairs)
    ok, why = check_listed_tree(inst, sol.pairs, sol.cost)
    assert ok, why


def test_regular_1000_full_table(gpu_ctx):
    inst = load_instance(CFG1)
    sol, f, code = solve_full(gpu_ctx, inst)
    gf, gcode = golden_table(CFG1)
    np.testing.assert_array_equal(f, gf)
    np.testing.assert_array_equal(kinds(code, inst.k), kinds(gcode, inst.k))
    assert sol.cost == META[CFG1]["ref_cost"]
    # the bound fits 16-bit cells
    assert sol.table_bytes == 2 and sol.merge_bits == 32 and sol.label_bits == 16


@pytest.mark.parametrize("flags", [emv.WIDE_LABELS, emv.FORCE_P64, emv.FORCE_I64,
                                   emv.FORCE_I64 | emv.FORCE_P64, emv.TABLE32,
                                   emv.TABLE32 | emv.WIDE_LABELS, emv.MERGE_1V,
                                   emv.MERGE_1V | emv.WIDE_LABELS, emv.MERGE_ROWS, emv.MERGE_CUBE])
@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/D/d17", CFG1])
def test_wide_paths_bit_exact(gpu_ctx, name, flags):
    """Every width combination (16/32/64-bit labels, 32/64-bit merge
    accumulators, 16-bit/int32/int64 table) and every merge of 16-bit cells
    (stars of rows sharing a base set, row by row: MERGE_ROWS, one vertex
    per lane: MERGE_1V) give the reference's table."""
    inst = load_instance(name)
    print(f"case {name} flags={flags}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst, flags)
    assert digest(f) == META[name]["f_sha256"]
    assert digest(kinds(code, inst.k)) == META[name]["kind_sha256"]
    if flags & emv.FORCE_P64:
        assert sol.merge_bits == 64 and sol.label_bits == 64
    if flags == emv.WIDE_LABELS:
        assert sol.label_bits == 32
    if flags & emv.FORCE_I64:
        assert sol.table_bytes == 8 and sol.merge_bits == 64
    elif flags & emv.TABLE32:
        assert sol.table_bytes == 4


@pytest.mark.parametrize("seed", range(4))
def test_random_regular_vs_oracle(gpu_ctx, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2000, 6000)) // 2 * 2
    d = int(rng.integers(3, 9))
    k = int(rng.integers(5, 10))
    e = emv.regular_graph(n, d, 1000, 500 + seed)
    kk = emv.terminals(n, k, 500 + seed)
    inst = Instance(n, e, kk)
    print(f"case regular n={n} d={d} k={k} seed={seed}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst)
    ocost, of, ocode = oracle.emv(n, e, kk)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
    assert sol.cost == ocost


@pytest.mark.parametrize("seed,k", [(0, 10), (1, 11), (2, 12)])
def test_cube_merge_vs_oracle(gpu_ctx, seed, k):
    """The LDS-staged cube merge (EMV_MERGE_CUBE): with k >= 10 the layers
    3..8 are covered by several 9-element cubes, rows split into parts over
    the waves; tables and records equal the oracle's."""
    n = 1500 + 500 * seed
    e = emv.regular_graph(n, 6, 1000, 700 + seed)
    kk = emv.terminals(n, k, 700 + seed)
    inst = Instance(n, e, kk)
    print(f"case cube n={n} k={k}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst, emv.MERGE_CUBE)
    assert sol.table_bytes == 2
    ocost, of, ocode = oracle.emv(n, e, kk)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
    assert sol.cost == ocost


def test_grid_vs_oracle(gpu_ctx):
    e = emv.grid_graph(60, 70, 100, 9)
    n = 60 * 70
    kk = emv.terminals(n, 8, 9)
    inst = Instance(n, e, kk)
    sol, f, code = solve_full(gpu_ctx, inst)
    ocost, of, ocode = oracle.emv(n, e, kk)
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


def test_edge_cases_k1_k2_duplicates(gpu_ctx):
    inst = load_instance("small-instances/B/b01")
    gpu_ctx.load_edges(inst.n, inst.edges)
    # k = 1: the answer row C = 0 is never written (reader-el.c:2137)
    sol = gpu_ctx.solve(inst.terminals[:1], emv.LIST)
    assert sol.cost == INF and len(sol.pairs) == 0
    # k = 2 through the DP equals the shortest path
    sol = gpu_ctx.solve(inst.terminals[:2], emv.LIST)
    d = dijkstra(adjacency(inst.n, inst.edges), int(inst.terminals[0]))
    assert sol.cost == d[int(inst.terminals[1])]
    # a repeated terminal changes nothing but the table shape
    kk = np.array([inst.terminals[0], inst.terminals[1], inst.terminals[0], inst.terminals[2]])
    sol = gpu_ctx.solve(kk, emv.KEEP_TABLE)
    f, _ = gpu_ctx.read_rows(0, 16)
    gpu_ctx.release_table()
    ocost, of, _ = oracle.emv(inst.n, inst.edges, kk)
    np.testing.assert_array_equal(f, of)
    assert sol.cost == ocost


def test_errors_are_status_codes(gpu_ctx):
    inst = load_instance("small-instances/B/b01")
    gpu_ctx.load_edges(inst.n, inst.edges)
    with pytest.raises(emv.EmvError) as ei:
        gpu_ctx.solve(np.array([0, inst.n + 5]))
    assert ei.value.status == -1
    with pytest.raises(emv.EmvError) as ei:
        gpu_ctx.read_rows(0, 1)
    assert ei.value.status == -4
    # disconnected graph: two components
    e = np.array([[0, 1, 3], [2, 3, 4]])
    gpu_ctx.load_edges(4, e)
    with pytest.raises(emv.EmvError) as ei:
        gpu_ctx.solve(np.array([0, 2, 1]))
    assert ei.value.status == -5


def test_sssp_vs_dijkstra(gpu_ctx):
    inst = load_instance(CFG1)
    gpu_ctx.load_edges(inst.n, inst.edges)
    d, p, _ = gpu_ctx.sssp(17)
    od, op = oracle.sssp(inst.n, inst.edges, 17)
    np.testing.assert_array_equal(d, od)
    np.testing.assert_array_equal(p, op)


def test_load_graph_reference_layout(gpu_ctx):
    # the reference's steinerq pos/adj layout (reader-el.c:696-859) gives the
    # same solve as the edge list
    inst = load_instance("small-instances/C/c01")
    n = inst.n
    deg = np.zeros(n, dtype=np.int64)
    for u, v, _ in inst.edges:
        deg[u] += 1
        deg[v] += 1
    pos = np.zeros(n + 1, dtype=np.int64)
    pos[1:] = np.cumsum(1 + 2 * deg)
    adj = np.zeros(pos[-1], dtype=np.int64)
    fill = pos[:-1].copy()
    for u, v, w in inst.edges:
        for a, b in ((u, v), (v, u)):
            p = pos[a]
            c = adj[p]
            adj[p + 1 + 2 * c] = b
            adj[p + 2 + 2 * c] = w
            adj[p] += 1
    gpu_ctx.load_graph(n, inst.m, pos, adj)
    sol = gpu_ctx.solve(inst.terminals)
    assert sol.cost == inst.cost


VERIFY_INPUT = re.compile(r"(.*)n = (.*), m = (.*), k = (.*), cost = (.*) \[(.*)ms\] {peak:(.*)GiB} "
                          r"{curr:(.*)GiB}")
VERIFY_EL_LIST = re.compile(r"erickson: \[zero:(.*)ms\] \[kernel:(.*)ms (.*)GiB/s\] \[traceback: "
                            r"(.*)ms\] done. \[(.*)ms\] \[cost:(.*)] {peak:(.*)GiB} {curr:(.*)GiB}")
VERIFY_EL = re.compile(r"erickson: \[zero:(.*)ms\] \[kernel:(.*)ms (.*)GiB/s\] done. \[(.*)ms\] "
                       r"\[cost:(.*)] {peak:(.*)GiB} {curr:(.*)GiB}")


def run_reader(args):
    out = subprocess.run([READER] + args, capture_output=True, text=True, timeout=120)
    return out.returncode, out.stdout, out.stderr


@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/I080/i080-001",
                                  "small-instances/SP/se03", CFG1])
def test_cli_drop_in_lines(name):
    """The reference harness's regexes (verify.py:120-169) parse our lines."""
    inst = load_instance(name)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "x.stp")
        write_stp(path, inst)
        if inst.cost >= 0:
            with open(path, "a") as fh:
                fh.write(f"cost {inst.cost}\n")
        rc, out, err = run_reader(["-in", path, "-el", "-list"])
        assert rc == 0, err
        lines = out.splitlines()
        el = [x for x in lines if x.startswith("erickson")][0]
        toks = VERIFY_EL_LIST.search(el)
        assert toks is not None, el
        cost = int(toks.group(6).strip())
        assert cost == META[name]["ref_cost"]
        sol = [x for x in lines if x.startswith("solution")][0]
        pairs = [tuple(int(a) - 1 for a in p.split()) for p in re.findall(r'"([^"]*)"', sol)]
        ok, why = check_listed_tree(inst, pairs, cost)
        assert ok, why
        assert VERIFY_INPUT.search([x for x in lines if x.startswith("input")][0])
        assert any(x.startswith("gpu:") for x in lines)
        rc, out, err = run_reader(["-in", path, "-el"])
        assert rc == 0
        el = [x for x in out.splitlines() if x.startswith("erickson")][0]
        assert VERIFY_EL.search(el) is not None


def test_cli_cost_mismatch_aborts():
    inst = load_instance("small-instances/B/b01")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "x.stp")
        write_stp(path, inst)
        with open(path, "a") as fh:
            fh.write("cost 81\n")
        rc, out, err = run_reader(["-in", path, "-el"])
        assert rc != 0 and "ERROR" in err and "min_cost != cost" in err


def test_cli_k2_path_and_dijkstra():
    inst = load_instance("path/p02")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "x.stp")
        write_stp(path, inst)
        rc, out, err = run_reader(["-in", path, "-el", "-list"])
        assert rc == 0, err
        el = [x for x in out.splitlines() if x.startswith("erickson")][0]
        assert "[zero:" not in el and "[cost: 20]" in el
        sol = [x for x in out.splitlines() if x.startswith("solution")][0]
        pairs = [tuple(int(a) - 1 for a in p.split()) for p in re.findall(r'"([^"]*)"', sol)]
        adj = adjacency(inst.n, inst.edges)
        # a real path from kk[1] to kk[0] of total weight 20
        assert pairs[0][0] == inst.terminals[1] and pairs[-1][1] == inst.terminals[0]
        tot = 0
        for a, b in pairs:
            tot += min(w for x, w in adj[a] if x == b)
        assert tot == 20
        rc, out, err = run_reader(["-in", path, "-dijkstra", "-seed", "7"])
        assert rc == 0, err


@pytest.mark.parametrize("name", ["small-instances/B/b01", CFG1])
def test_table_file_round_trip(gpu_ctx, tmp_path, name):
    """emv_write_table writes the canonical table file: header fields, every
    cell (INF = INT64_MAX) and record code equal the oracle's, and the CLI's
    -dump writes the same bytes."""
    inst = load_instance(name)
    print(f"case table file {name}", flush=True)
    gpu_ctx.load_edges(inst.n, inst.edges)
    sol = gpu_ctx.solve(inst.terminals, emv.KEEP_TABLE)
    path = tmp_path / "t.emvt"
    gpu_ctx.write_table(path, records=True)
    gpu_ctx.release_table()
    t = emv.read_table_file(path)
    assert (t["k"], t["n"], t["inf"], t["cost"]) == (inst.k, inst.n, 2 ** 63 - 1, sol.cost)
    assert os.path.getsize(path) == 64 + (1 << inst.k) * inst.n * 12
    ocost, of, ocode = oracle.emv(inst.n, inst.edges, inst.terminals)
    np.testing.assert_array_equal(t["f"], of)
    np.testing.assert_array_equal(t["code"], ocode)
    stp = tmp_path / "x.stp"
    write_stp(str(stp), inst)
    cli = tmp_path / "cli.emvt"
    rc, out, err = run_reader(["-in", str(stp), "-el", "-dump", str(cli)])
    assert rc == 0, err
    assert any(x.startswith("dump: [file: ") for x in out.splitlines())
    assert cli.read_bytes() == path.read_bytes()


def test_write_table_needs_a_kept_table(gpu_ctx, tmp_path):
    inst = load_instance("small-instances/B/b01")
    gpu_ctx.load_edges(inst.n, inst.edges)
    gpu_ctx.solve(inst.terminals)
    with pytest.raises(emv.EmvError, match="ERR_STATE"):
        gpu_ctx.write_table(tmp_path / "t.emvt")


def test_cli_group_reduction(tmp_path):
    """Group Steiner by reduction: cost = OPT(reduced) - g*M and the printed
    group tree is a real tree of that cost touching every group."""
    R, C = 9, 11
    e = emv.grid_graph(R, C, 50, 21)
    n = R * C
    grp = emv.groups(n, 4, 5, 21)
    inst = Instance(n, e, emv.terminals(n, 2, 21))
    path = str(tmp_path / "g.stp")
    write_stp(path, inst)
    gpath = str(tmp_path / "g.grp")
    with open(gpath, "w") as fh:
        for row in grp:
            fh.write("g " + " ".join(str(int(v) + 1) for v in row) + "\n")
    rc, out, err = run_reader(["-in", path, "-el", "-list", "-groups", gpath])
    assert rc == 0, err
    line = [x for x in out.splitlines() if x.startswith("group: [groups")][0]
    gcost = int(re.search(r"\[cost: (-?\d+)\]", line).group(1))
    nn, red, kk, M = emv.group_reduce(n, e, grp)
    ocost, _, _ = oracle.emv(nn, red, kk)
    assert gcost == ocost - len(grp) * M
    sol = [x for x in out.splitlines() if x.startswith("group solution")][0]
    pairs = [tuple(int(a) - 1 for a in p.split()) for p in re.findall(r'"([^"]*)"', sol)]
    adj = adjacency(n, e)
    tot = sum(min(w for x, w in adj[a] if x == b) for a, b in pairs)
    assert tot == gcost
    verts = {v for p in pairs for v in p}
    for row in grp:
        assert verts & set(int(v) for v in row)


def test_csr_and_ell_paths_agree(gpu_ctx, monkeypatch):
    """The padded fixed-degree adjacency and the CSR walk give the same table."""
    n, d, k = 3000, 6, 7
    e = emv.regular_graph(n, d, 1000, 77)
    kk = emv.terminals(n, k, 77)
    inst = Instance(n, e, kk)
    _, f1, c1 = solve_full(gpu_ctx, inst)
    monkeypatch.setenv("EMV_NO_ELL", "1")
    _, f2, c2 = solve_full(gpu_ctx, inst)
    monkeypatch.delenv("EMV_NO_ELL")
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(c1, c2)


@pytest.mark.parametrize("chunk128", ["1", "0"])
@pytest.mark.parametrize("ell,flags", [(True, 0), (False, 0), (True, emv.TABLE32)])
@pytest.mark.parametrize("n,d,k", [(2999, 8, 10), (1234, 5, 8), (778, 3, 3), (4100, 7, 9)])
def test_chunk_widths_vs_oracle(gpu_ctx, monkeypatch, n, d, k, ell, flags, chunk128):
    """16-bit labels relax vertex-wise in chunks of 128 rows (two labels per
    32-bit word, packed saturating add and min; records two rows per lane)
    or, with EMV_CHUNK128=0, in chunks of 64; both adjacency paths; 16-bit
    and int32 tables; layers whose last chunk is partly filled (k = 8: 70
    rows, k = 10: 252) and k = 3.  Cells, records and cost equal the
    oracle's."""
    monkeypatch.setenv("EMV_CHUNK128", chunk128)
    if not ell:
        monkeypatch.setenv("EMV_NO_ELL", "1")
    e = emv.regular_graph(n, d, 1000, 900 + k)
    kk = emv.terminals(n, k, 900 + k)
    inst = Instance(n, e, kk)
    print(f"case chunks n={n} d={d} k={k} ell={ell} flags={flags} chunk128={chunk128}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst, flags)
    assert sol.label_bits == 16
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("qcap", ["64", "100000"])
def test_worklist_rounds_agree(gpu_ctx, monkeypatch, qcap):
    """Rounds taken from the worklist (EMV_QCAP > 0) give the flag-scan
    result cell for cell; 64 forces list overflow back to the flag scan."""
    inst = Instance(3600, emv.grid_graph(60, 60, 100, 12), emv.terminals(3600, 7, 12))
    monkeypatch.setenv("EMV_BLOCKS", "0")  # vertex-wise rounds, where the worklist lives
    _, f1, c1 = solve_full(gpu_ctx, inst)
    monkeypatch.setenv("EMV_QCAP", qcap)
    ctx = emv.Context(0)
    try:
        _, f2, c2 = solve_full(ctx, inst)
    finally:
        ctx.close()
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(c1, c2)


@pytest.mark.parametrize("closure", ["1", "0"])
@pytest.mark.parametrize("flags", [0, emv.WIDE_LABELS, emv.FORCE_P64])
def test_block_relax_grid_vs_oracle(gpu_ctx, monkeypatch, flags, closure):
    """Grids take the block-local relaxation (LDS sweeps inside vertex
    blocks, then in-block distances from the lowered border slots on later
    visits, or sweeps on every visit with EMV_BLOCK_CLOSURE=0); every cell
    and record equal the oracle's at every label width."""
    R, C = 50, 60
    n = R * C
    e = emv.grid_graph(R, C, 100, 21)
    kk = emv.terminals(n, 8, 21)
    monkeypatch.setenv("EMV_BLOCK_CLOSURE", closure)
    print(f"case grid {R}x{C} flags={flags} closure={closure}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk), flags)
    if flags != emv.FORCE_P64:
        assert sol.relax_blocks > 0
        # in-block distances of a w <= 100 grid fit 16 bits
        assert sol.block_dist_bits == (16 if closure == "1" else 0)
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("full_sweeps", ["1", "12"])
@pytest.mark.parametrize("shape,flags", [("grid", 0), ("grid", emv.WIDE_LABELS), ("group", 0)])
def test_block_relax_many_chunks_vs_oracle(gpu_ctx, monkeypatch, shape, flags, full_sweeps):
    """Block-local relaxation over layers of many 64-row chunks (k = 12:
    layers 4 .. 8 hold 8 to 15 chunks, so one round's items span chunks and
    blocks): a grid, and a group-reduced grid whose virtual terminals join
    blocks far apart. A block's first visit hands over from full sweeps to
    worklists after one sweep or after twelve (EMV_BLOCK_FULL_SWEEPS). Every
    cell and record equal the oracle's."""
    import workloads

    if shape == "grid":
        R = C = 48
        n = R * C
        e = emv.grid_graph(R, C, 100, 43)
        kk = emv.terminals(n, 12, 43)
    else:
        _, _, _, n, e, kk, _ = workloads.group_instance(36, 36, 12, 5, 100, 44)
    monkeypatch.setenv("EMV_BLOCK_FULL_SWEEPS", full_sweeps)
    print(f"case block relax {shape} n={n} k={len(kk)} flags={flags} full_sweeps={full_sweeps}",
          flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk), flags)
    assert sol.relax_blocks > 0
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("n,d,seed", [(1000, 4, 1), (300, 20, 2), (1200, 16, 3)])
def test_block_relax_forced_vs_oracle(gpu_ctx, monkeypatch, n, d, seed):
    """EMV_BLOCKS=1 forces blocks on graphs without locality: most neighbours
    outside the block (d=16: more outside neighbours than the kernel holds in
    registers) or more than eight inside it (n=300, d=20) take the overflow
    walks of the kernel."""
    e = emv.regular_graph(n, d, 1000, 700 + seed)
    kk = emv.terminals(n, 7, 700 + seed)
    monkeypatch.setenv("EMV_BLOCKS", "1")
    print(f"case forced blocks n={n} d={d}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk))
    assert sol.relax_blocks > 0 and sol.block_dist_bits in (16, 32)
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


def test_block_and_vertex_rounds_agree_on_group_instance(gpu_ctx, monkeypatch):
    """A group-reduced grid (virtual terminals of degree 12, joined by weight-M
    edges) solved with blocks and with vertex-wise rounds: same table, same
    records, the oracle's cost."""
    import workloads

    n, e, grp, nn, red, kk, M = workloads.group_instance(36, 36, 5, 12, 100, 31)
    inst = Instance(nn, red, kk)
    print(f"case group grid nn={nn} k={len(kk)}", flush=True)
    s1, f1, c1 = solve_full(gpu_ctx, inst)
    assert s1.relax_blocks > 0
    monkeypatch.setenv("EMV_BLOCKS", "0")
    s2, f2, c2 = solve_full(gpu_ctx, inst)
    assert s2.relax_blocks == 0
    np.testing.assert_array_equal(f1, f2)
    np.testing.assert_array_equal(c1, c2)
    assert s1.cost == oracle.emv(nn, red, kk)[0]


@pytest.mark.parametrize("shape", ["regular", "grid"])
def test_large_weights_64bit_labels(gpu_ctx, shape):
    """Weights near 2^30: the shortest-path chunk runs on 64-bit labels
    (n * max w >= 2^32), the table is int64, and cells, records and cost
    still equal the oracle's (vertex-wise rounds and blocks)."""
    if shape == "regular":
        n = 1500
        e = emv.regular_graph(n, 5, 1 << 30, 41)
    else:
        n = 40 * 45
        e = emv.grid_graph(40, 45, 1 << 30, 41)
    kk = emv.terminals(n, 6, 41)
    print(f"case large weights {shape} n={n}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk))
    assert sol.table_bytes == 8
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
    d, p, _ = gpu_ctx.sssp(int(kk[0]))
    np.testing.assert_array_equal(d, of[1])




@pytest.mark.parametrize("grid", ["1", "0"])
@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/D/d17", CFG1])
def test_star_merges_bit_exact(gpu_ctx, monkeypatch, name, grid):
    """Every layer merged in grids of rows S | a_i | b_j (layers of >= 5
    terminals) or in stars of rows sharing a base set (EMV_MERGE_GRID=0, and
    the smaller layers), forced for small layers, which otherwise merge row
    by row: the reference's table and record kinds, as row by row."""
    monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    monkeypatch.setenv("EMV_MERGE_GRID", grid)
    inst = load_instance(name)
    print(f"case stars {name}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst)
    assert sol.table_bytes == 2
    assert digest(f) == META[name]["f_sha256"]
    assert digest(kinds(code, inst.k)) == META[name]["kind_sha256"]
    s2, f2, c2 = solve_full(gpu_ctx, inst, emv.MERGE_ROWS)
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(code, c2)


@pytest.mark.parametrize("flags", [emv.TABLE32, emv.TABLE32 | emv.FORCE_P64,
                                   emv.TABLE32 | emv.WIDE_LABELS])
@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/D/d17", CFG1])
def test_star_merges_32bit_cells(gpu_ctx, monkeypatch, name, flags):
    """Stars of rows sharing a base set on an int32 table (forced for small
    layers), with 32- and 64-bit packed words: the reference's table and
    record kinds, as row by row.  (The compact int32 copy of an int64 table
    takes the same stars with row bases: the group-instance test.)"""
    monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    inst = load_instance(name)
    print(f"case stars32 {name} flags={flags}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst, flags)
    assert sol.table_bytes == 4
    assert digest(f) == META[name]["f_sha256"]
    assert digest(kinds(code, inst.k)) == META[name]["kind_sha256"]
    s2, f2, c2 = solve_full(gpu_ctx, inst, flags | emv.MERGE_ROWS)
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(code, c2)


@pytest.mark.parametrize("flags", [emv.TABLE32, emv.TABLE32 | emv.FORCE_P64])
@pytest.mark.parametrize("n,d,k", [(3001, 6, 11), (2000, 8, 9), (230, 3, 13)])
def test_star_merges_32bit_cells_vs_oracle(gpu_ctx, monkeypatch, n, d, k, flags):
    """32-bit-cell stars on random regular graphs (k = 9 .. 13: stars of up
    to four rows, parts over the outer steps): cells, records and cost equal
    the oracle's."""
    if n * d % 2:
        n += 1
    monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    e = emv.regular_graph(n, d, 1000, 31 + k)
    kk = emv.terminals(n, k, 31 + k)
    print(f"case stars32 n={n} d={d} k={k} flags={flags}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk), flags)
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("tiles", ["1", "0"])
@pytest.mark.parametrize("n,d,k", [(3001, 6, 11), (2000, 8, 9), (777, 4, 12), (130, 3, 10)])
def test_grid_merges_vs_oracle(gpu_ctx, monkeypatch, n, d, k, tiles):
    """16-bit cells merged in grids of rows S | a_i | b_j on every layer of
    >= 5 terminals (forced; 1 x 1 up to 3 x 3 grids, parts over the outer
    steps), on the tiled table and (EMV_TILE_TABLE=0) on rows; n = 3001, 777
    and 130 end in partly filled tiles.  Cells, records and cost equal the
    oracle's."""
    if n * d % 2:
        n += 1
    monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    monkeypatch.setenv("EMV_MERGE_GRID", "1")
    monkeypatch.setenv("EMV_TILE_TABLE", tiles)
    e = emv.regular_graph(n, d, 1000, 53 + k)
    kk = emv.terminals(n, k, 53 + k)
    print(f"case grids n={n} d={d} k={k} tiles={tiles}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk))
    assert sol.table_bytes == 2
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("n,d,k", [(100, 4, 15), (48, 3, 16), (230, 3, 13)])
@pytest.mark.parametrize("flags,grid", [(0, "1"), (0, "0"), (emv.MERGE_ROWS, "1"),
                                        (emv.MERGE_1V, "1")])
def test_many_terminals_16bit_cells(gpu_ctx, monkeypatch, n, d, k, flags, grid):
    """k = 13 .. 16 on 16-bit cells: the two-vertex merges' keys hold Y in
    16 bits and their low rows in 32-bit lane offsets at the largest k they
    take; cells, records and cost equal the oracle's, in stars (forced for
    these small layers; grids of rows S | a_i | b_j, or stars with
    EMV_MERGE_GRID=0), row by row and one vertex per lane.  n = 100, 48 and
    230 end in partly filled 128-vertex tiles."""
    monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    monkeypatch.setenv("EMV_MERGE_GRID", grid)
    e = emv.regular_graph(n, d, 1000, 7)
    kk = emv.terminals(n, k, 7)
    print(f"case n={n} d={d} k={k} flags={flags}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk), flags)
    assert sol.table_bytes == 2
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


@pytest.mark.parametrize("k", [17, 19])
def test_more_than_16_terminals_vs_oracle(gpu_ctx, k):
    """k = 17 and 19: 16-bit cells, but the two-vertex merges' keys hold Y in
    16 bits, so these layers take the generic merges with 64-bit packed words
    ((sum << k) | Y no longer fits 32 bits).  Cells, records and cost equal
    the oracle's."""
    n = 24
    e = emv.regular_graph(n, 3, 1000, 11)
    kk = emv.terminals(n, k, 11)
    print(f"case n={n} k={k}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk))
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
    check_listed_tree(Instance(n, e, kk), sol.pairs, sol.cost)


def test_k22_certified(gpu_ctx):
    """k = 22 (2^22 rows; the oracle would take minutes): the fixed-point
    certificate over all 2^22 x n cells and a validated listed tree."""
    from bigcheck import validate_tree

    n, k = 24, 22
    e = emv.regular_graph(n, 3, 1000, 12)
    kk = emv.terminals(n, k, 12)
    print(f"case n={n} k={k}", flush=True)
    gpu_ctx.load_edges(n, e)
    sol = gpu_ctx.solve(kk, emv.LIST | emv.KEEP_TABLE)
    cert = gpu_ctx.certify()
    print(f"certificate {cert}", flush=True)
    assert cert.ok and cert.cells == (1 << k) * n, cert
    ok, why = validate_tree(n, e, kk, sol.pairs, sol.cost)
    assert ok, why
    gpu_ctx.release_table()


@pytest.mark.parametrize("stars", ["0", "1"])
@pytest.mark.parametrize("blocks", ["1", "0"])
def test_relative_32bit_labels_on_group_instance(gpu_ctx, monkeypatch, blocks, stars):
    """Group-reduced grid with weights up to 2^20: the bound needs 64-bit
    labels (M is about 9e8), but every row's merged values span less than
    2^32, so the layers relax on 32-bit labels relative to each row's
    minimum, and merge from the int32 copy of the rows (cell - row base),
    row by row or in stars (forced for these small layers).  Cells, records
    and cost equal the oracle's, with blocks and with vertex-wise rounds,
    and with 64-bit labels forced (FORCE_P64)."""
    import workloads

    n, e, grp, nn, red, kk, M = workloads.group_instance(30, 30, 5, 6, 1 << 20, 35)
    monkeypatch.setenv("EMV_BLOCKS", blocks)
    if stars == "1":
        monkeypatch.setenv("EMV_STAR_MIN_CELLS", "0")
    inst = Instance(nn, red, kk)
    print(f"case relative labels group grid nn={nn} k={len(kk)} M={M} blocks={blocks}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst)
    assert sol.label_bits == 64 and sol.table_bytes == 8
    assert sol.relative_label_layers == len(kk) - 1
    # every layer merges from the int32 copy of the int64 rows
    assert sol.compact_merge_layers == len(kk) - 1
    assert (sol.relax_blocks > 0) == (blocks == "1")
    ocost, of, ocode = oracle.emv(nn, red, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
    s64, f64, c64 = solve_full(gpu_ctx, inst, emv.FORCE_P64)
    assert s64.relative_label_layers == 0 and s64.compact_merge_layers == 0
    np.testing.assert_array_equal(f64, of)
    np.testing.assert_array_equal(c64, ocode)


@pytest.mark.parametrize("flags", [0, emv.FORCE_I64, emv.FORCE_P64])
def test_merge_k11_vs_oracle(gpu_ctx, flags):
    """The merge at k=11 -- whole rows on the lower layers, rows split into
    parts folded with atomicMin on the upper ones: every cell and record
    equal the oracle's, for int32/int64 tables and 32/64-bit accumulators."""
    n, d, k = 3000, 5, 11
    e = emv.regular_graph(n, d, 1000, 89)
    kk = emv.terminals(n, k, 89)
    print(f"case merge k=11 flags={flags}", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk), flags)
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


def test_block_distances_32bit(gpu_ctx):
    """A grid whose in-block distances need 32 bits (w up to 20000): the
    closure reads 32-bit rows; cells and records equal the oracle's."""
    R, C = 40, 45
    n = R * C
    e = emv.grid_graph(R, C, 20000, 23)
    kk = emv.terminals(n, 7, 23)
    print(f"case grid {R}x{C} w<=20000", flush=True)
    sol, f, code = solve_full(gpu_ctx, Instance(n, e, kk))
    assert sol.relax_blocks > 0 and sol.block_dist_bits == 32
    ocost, of, ocode = oracle.emv(n, e, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)


def test_block_closure_with_64bit_singleton_labels(gpu_ctx):
    """Group-reduced grid whose singleton rows need 64-bit labels (n * M >
    2^32) while its blocks take in-block distances: a border slot that holds
    no path yet in some row (INF) must lower nothing there -- INF + d wraps
    in 64 bits.  Cells, records and cost equal the oracle's."""
    import workloads

    n, e, grp, nn, red, kk, M = workloads.group_instance(40, 40, 6, 8, 3000, 33)
    assert nn * M > 2 ** 32
    inst = Instance(nn, red, kk)
    print(f"case group grid nn={nn} k={len(kk)} M={M}", flush=True)
    sol, f, code = solve_full(gpu_ctx, inst)
    assert sol.relax_blocks > 0 and sol.block_dist_bits == 32
    ocost, of, ocode = oracle.emv(nn, red, kk)
    assert sol.cost == ocost
    np.testing.assert_array_equal(f, of)
    np.testing.assert_array_equal(code, ocode)
