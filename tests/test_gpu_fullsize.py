"""Parity at the measured workloads' full sizes (the bench's workloads, not
reduced copies): the 8-regular n = 2^18, k = 14 solve; the 256-instance
batch; the 512 x 512 served query stream; the 1024 x 1024 group instance.

The oracle cannot run these tables on the CPU in test time (the k = 14 table
alone is 2^32 cells), so each is proven with the fixed-point certificate
(emv_certify: every cell equal to the least of its edge, split and source
candidates, every record code by its rule; a connected graph with positive
weights has exactly one table that passes, so a clean certificate means
every cell equals the reference's).  The certificate kernel is itself pinned
first: it passes the oracle's table and the reference's own golden table and
catches slips planted in them on the CPU.  Beside it, independent CPU
checks: singleton rows against the oracle's Dijkstra, sampled rows against
a numpy fixed point, listed trees against scipy distances, and whole tables
against the oracle where it finishes (k <= 8 at n = 2^18, the batch's
smallest and largest instances)."""
import numpy as np
import pytest

import emv
import oracle
import workloads
from bigcheck import (INF, row_fixed_point, subsets, validate_direct_group_tree,
                      validate_group_tree, validate_tree)
from golden_data import CFG1, golden_table, load_instance

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def assert_clean(cert, what):
    assert cert.ok, f"{what}: {cert}"


# ------------------------------------------------------------ the checker

def test_certificate_passes_oracle_and_reference_tables(gpu_ctx):
    inst = load_instance(CFG1)
    gpu_ctx.load_edges(inst.n, inst.edges)
    _, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    cert = gpu_ctx.certify_table(inst.terminals, f, code)
    assert_clean(cert, "oracle table")
    assert cert.cells == (1 << inst.k) * inst.n
    # the reference's own table (built from reader-el.c): every value passes;
    # its edge records follow its heap order, so only the values are compared
    gf, gcode = golden_table(CFG1)
    cert = gpu_ctx.certify_table(inst.terminals, gf, gcode)
    assert cert.bad_value == 0 and cert.not_finite == 0, cert


def test_certificate_catches_planted_slips(gpu_ctx):
    inst = load_instance(CFG1)
    gpu_ctx.load_edges(inst.n, inst.edges)
    _, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    k = inst.k
    S2 = 1 << k
    v = 123
    merge_cell = next((X, u) for X in range(3, S2) for u in range(inst.n)
                      if X & (X - 1) and 0 < code[X, u] < S2 and code[X, u] != X)
    edge_cell = next((X, u) for X in range(3, S2) for u in range(inst.n)
                     if X & (X - 1) and code[X, u] >= S2)
    cases = [
        ("value +1", lambda g, c: g.__setitem__((37, v), g[37, v] + 1), "bad_value"),
        ("value -1", lambda g, c: g.__setitem__((37, v), g[37, v] - 1), "bad_value"),
        ("singleton value", lambda g, c: g.__setitem__((4, v), g[4, v] + 1), "bad_value"),
        ("terminal not zero", lambda g, c: g.__setitem__((1, int(inst.terminals[0])), 1),
         "bad_value"),
        ("INF cell", lambda g, c: g.__setitem__((200, v), INF), "not_finite"),
        ("row 0 record", lambda g, c: c.__setitem__((0, v), 0), "bad_record"),
        ("merge split", lambda g, c: c.__setitem__(merge_cell, c[merge_cell] ^ 1), "bad_record"),
        ("edge predecessor", lambda g, c: c.__setitem__(edge_cell, c[edge_cell] + 1),
         "bad_record"),
        ("shortcut record", lambda g, c: c.__setitem__((2, v), c[2, v] + 1), "bad_record"),
    ]
    for name, plant, counter in cases:
        g, c = f.copy(), code.copy()
        plant(g, c)
        cert = gpu_ctx.certify_table(inst.terminals, g, c)
        say(f"planted {name}: {cert}")
        assert getattr(cert, counter) >= 1, name
        diff = np.argwhere((g != f) | (c != code))[0]
        assert (cert.first_X, cert.first_v) <= (int(diff[0]), int(diff[1])), name


# ------------------------------------------------- 8-regular n=2^18, k=14

def test_regular_n262144_k14_certified():
    n, e, kk = workloads.regular_single(262144, 8, 1000, 14, 2)
    k = len(kk)
    say(f"case 8-regular n={n} m={len(e)} k={k}")
    with emv.Context(0) as ctx:
        ctx.load_edges(n, e)
        sol = ctx.solve(kk, emv.LIST | emv.KEEP_TABLE)
        assert sol.table_bytes == 2 and sol.cost == 14108
        cert = ctx.certify()
        say(f"certificate {cert}")
        assert_clean(cert, "k=14 table")
        assert cert.cells == (1 << k) * n
        # singleton rows: the oracle's Dijkstra
        for t in range(k):
            d, _ = oracle.sssp(n, e, int(kk[t]))
            row, _ = ctx.read_rows(1 << t, 1, codes=False)
            np.testing.assert_array_equal(row[0], d)
        # sampled rows: a numpy fixed point from the rows of their subsets
        for X in [0b11, 0b10110, 0b1000100010001, (1 << 13) - 1 - 0b101]:
            Ys = subsets(X)
            rows = {}
            for Y in Ys:
                rows[Y] = ctx.read_rows(Y, 1, codes=False)[0][0]
            assert row_fixed_point(e, X, rows), X
        C = (1 << (k - 1)) - 1
        assert ctx.read_rows(C, 1, codes=False)[0][0][kk[-1]] == sol.cost
        ok, why = validate_tree(n, e, kk, sol.pairs, sol.cost)
        assert ok, why
        ctx.release_table()
        # the int32 table and the 32-bit labels give the same table
        sol32 = ctx.solve(kk, emv.KEEP_TABLE | emv.TABLE32 | emv.WIDE_LABELS)
        assert sol32.table_bytes == 4 and sol32.label_bits == 32 and sol32.cost == sol.cost
        assert_clean(ctx.certify(), "k=14 int32 table")
        ctx.release_table()
        # the LDS-staged cube merge (opt-in) gives the same table
        solc = ctx.solve(kk, emv.KEEP_TABLE | emv.MERGE_CUBE)
        assert solc.cost == sol.cost
        assert_clean(ctx.certify(), "k=14 table, cube merge")
        ctx.release_table()


# --------------------------------------------------- the 256-instance batch

def test_batch_256_instances_certified():
    insts = workloads.batch_instances(256)
    say(f"case batch of {len(insts)} instances, {sum((1 << len(kk)) * n for n, _, kk in insts)} cells")
    plain = []
    with emv.Context(0) as ctx:
        for n, e, kk in insts:
            ctx.load_edges(n, e)
            plain.append(ctx.solve(kk).cost)
    with emv.Batch(insts) as b:
        for rep in range(2):  # eager, then the graph replay
            costs, st = b.solve(emv.KEEP_TABLE)
            say(f"batch solve {rep}: graph={st.graph_used} fallback={st.graph_fallback}")
            np.testing.assert_array_equal(costs, plain)
            bad = [(i, c) for i in range(len(insts)) for c in [b.certify(i)] if not c.ok]
            assert not bad, bad[:3]
        assert st.graph_used == 1
        # whole tables against the oracle: smallest n, largest n, largest k,
        # most cells (read from the replayed solve)
        sizes = [(n, len(kk), (1 << len(kk)) * n) for n, _, kk in insts]
        picks = sorted(set([min(range(256), key=lambda i: sizes[i][0]),
                            max(range(256), key=lambda i: sizes[i][0]),
                            max(range(256), key=lambda i: (sizes[i][1], -sizes[i][0])),
                            max(range(256), key=lambda i: sizes[i][2])]))
        for i in picks:
            n, e, kk = insts[i]
            say(f"oracle instance {i}: n={n} k={len(kk)}")
            ocost, of, ocode = oracle.emv(n, e, kk)
            f, code = b.read_rows(i, 0, 1 << len(kk), n)
            assert costs[i] == ocost
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)


# --------------------------------------------------- the served query stream

def test_served_stream_grid512_certified():
    R = C = 512
    n = R * C
    e = emv.grid_graph(R, C, 100, 4)
    qs = workloads.grid_queries(n, 24, 40)
    rng = np.random.default_rng(41)
    gqs = [[rng.choice(n, size=8, replace=False) for _ in range(g)] for g in (4, 6, 9)]
    say(f"case grid {R}x{C}: {len(qs)} terminal queries k={[len(q) for q in qs]}, "
        f"{len(gqs)} group queries")
    with emv.Context(0) as ctx:
        ctx.load_edges(n, e)
        plain = [ctx.solve(q).cost for q in qs]
    small = [i for i, q in enumerate(qs) if len(q) <= 8][:3]
    assert len(small) == 3
    with emv.Server(n, e, max_inflight=64) as srv:
        tk = [srv.submit(q, emv.KEEP_TABLE | emv.LIST) for q in qs]
        gk = [srv.submit_groups(g, emv.KEEP_TABLE | emv.LIST) for g in gqs]
        while any(srv.poll(t).state != 2 for t in tk + gk):
            srv.step()
        for i, t in enumerate(tk):
            info = srv.poll(t)
            assert info.status == 0 and info.cost == plain[i], i
            assert_clean(srv.certify(t), f"query {i} k={len(qs[i])}")
        for i in small:
            q = qs[i]
            say(f"oracle query {i}: k={len(q)}")
            ocost, of, ocode = oracle.emv(n, e, q)
            f, code = srv.read_rows(tk[i], 0, 1 << len(q), n)
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)
            ok, why = validate_tree(n, e, q, srv.pairs(tk[i]), ocost)
            assert ok, why
        M = srv.group_weight()
        assert M == 1 + int(np.asarray(e)[:, 2].sum())
        for j, t in enumerate(gk):
            info = srv.poll(t)
            assert info.status == 0 and info.groups == len(gqs[j])
            assert info.reduced_cost == info.cost + len(gqs[j]) * M
            assert_clean(srv.certify(t), f"group query {j}")
            ok, why = validate_direct_group_tree(n, e, gqs[j], srv.pairs(t), info.cost, info.root)
            assert ok, why
            say(f"group query {j}: groups={len(gqs[j])} cost={info.cost}")
        for t in tk + gk:
            srv.release(t)


# --------------------------------------------- the 1024 x 1024 group instance

def test_group_grid1024_certified_and_validated():
    R = C = 1024
    g, size = 13, 32
    n, e, grp, nn, red, kk, M = workloads.group_instance(R, C, g, size, 1000, 5)
    say(f"case group grid {R}x{C}: {g} groups of {size}, reduced n={nn} M={M}")
    with emv.Context(0) as ctx:
        ctx.load_edges(nn, red)
        sol = ctx.solve(kk, emv.LIST | emv.KEEP_TABLE)
        assert sol.table_bytes == 8
        cert = ctx.certify()
        say(f"certificate {cert}")
        assert_clean(cert, "reduced k=13 int64 table")
        ok, why, gcost, _ = validate_group_tree(n, e, grp, nn, red, M, sol.pairs, sol.cost)
        assert ok, why
        assert gcost == sol.cost - g * M
        say(f"reduced cost {sol.cost}, group cost {gcost}")
        sample = [1 << 3, 0b1010011, (1 << 12) - 1, (1 << 13) - 1]
        red_rows = {X: ctx.read_rows(X, 1, codes=True) for X in sample}
        ctx.release_table()
    # the same query on a resident graph: the DP runs on the grid itself
    with emv.Server(n, e, max_inflight=4) as srv:
        t = srv.submit_groups([list(x) for x in grp], emv.KEEP_TABLE | emv.LIST)
        while srv.poll(t).state != 2:
            srv.step()
        info = srv.poll(t)
        assert info.status == 0
        assert srv.group_weight() == M
        assert info.reduced_cost == sol.cost and info.cost == gcost
        assert_clean(srv.certify(t), "server group table")
        ok, why = validate_direct_group_tree(n, e, grp, srv.pairs(t), info.cost, info.root)
        assert ok, why
        # reduced cells at graph vertices = |X| M + the server's cells;
        # records agree away from the singleton rows
        for X in sample:
            f, code = srv.read_rows(t, X, 1, n)
            rf, rcode = red_rows[X]
            np.testing.assert_array_equal(rf[0][:n], f[0] + bin(X).count("1") * M)
            if X & (X - 1):
                np.testing.assert_array_equal(rcode[0][:n], code[0])
        srv.release(t)
