"""Batched solves and the query server (multi-instance engine) on the GPU:
every instance's table, records and cost equal the oracle's, alone and in
the mix, eagerly and from a hipGraph replay.  Integer work: tolerance 0."""
import numpy as np
import pytest

import emv
import oracle
from golden_data import META, load_instance
from bigcheck import validate_direct_group_tree
from stputil import check_listed_tree, Instance

pytestmark = pytest.mark.gpu


def mixed_instances(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(2 ** rng.uniform(8, 11.5))
        d = int(rng.integers(3, 9))
        if (n * d) % 2:
            n += 1
        k = int(rng.integers(3, 9))
        e = emv.regular_graph(n, d, 1000, 900 + 17 * seed + i)
        kk = emv.terminals(n, k, 900 + 17 * seed + i)
        out.append((n, e, kk))
    return out


def test_batch_matches_oracle_eager_and_graph():
    insts = mixed_instances(12, 1)
    with emv.Batch(insts) as b:
        c1, s1 = b.solve(emv.KEEP_TABLE | emv.LIST)
        assert s1.graph_used == 0
        for i, (n, e, kk) in enumerate(insts):
            ocost, of, ocode = oracle.emv(n, e, kk)
            assert c1[i] == ocost
            f, code = b.read_rows(i, 0, 1 << len(kk), n)
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)
            ok, why = check_listed_tree(Instance(n, e, kk), b.pairs(i), c1[i])
            assert ok, why
        # second solve: replayed from a graph, identical tables
        c2, s2 = b.solve(emv.KEEP_TABLE)
        assert s2.graph_used == 1 or s2.graph_fallback == 1
        np.testing.assert_array_equal(c1, c2)
        for i, (n, e, kk) in enumerate(insts):
            _, of, ocode = oracle.emv(n, e, kk)
            f, code = b.read_rows(i, 0, 1 << len(kk), n)
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)


def test_batch_more_than_16_terminals():
    """A batch mixing k = 17 and 18 instances (16-bit cells, generic merges:
    the two-vertex merges hold Y in 16 bits) with small ones: every table,
    record and cost equals the oracle's, eagerly and from the replay."""
    insts = []
    for i, (n, k) in enumerate([(24, 17), (26, 18), (200, 6), (64, 9)]):
        e = emv.regular_graph(n, 3, 1000, 70 + i)
        insts.append((n, e, emv.terminals(n, k, 70 + i)))
    want = [oracle.emv(n, e, kk) for n, e, kk in insts]
    with emv.Batch(insts) as b:
        for rep in range(2):
            costs, st = b.solve(emv.KEEP_TABLE)
            print(f"batch solve {rep}: graph={st.graph_used}", flush=True)
            for i, (n, e, kk) in enumerate(insts):
                ocost, of, ocode = want[i]
                assert costs[i] == ocost, i
                f, code = b.read_rows(i, 0, 1 << len(kk), n)
                np.testing.assert_array_equal(f, of)
                np.testing.assert_array_equal(code, ocode)


def test_batch_of_testset_instances():
    names = ["small-instances/B/b01", "small-instances/C/c01", "small-instances/P6E/p619",
             "small-instances/I080/i080-011", "small-instances/D/d06", "small-instances/SP/se03"]
    insts = [load_instance(nm) for nm in names]
    with emv.Batch([(x.n, x.edges, x.terminals) for x in insts]) as b:
        costs, _ = b.solve()
        for x, c in zip(insts, costs):
            assert c == x.cost
        b.set_params(use_graph=1, graph_round_margin=0)
        costs2, st = b.solve()
        np.testing.assert_array_equal(costs, costs2)


def test_batch_graph_replay_short_of_rounds_falls_back():
    """A replay given too few relaxation rounds (forced from the host with a
    negative margin) is caught by the per-step convergence words and the
    batch is solved again eagerly, with the same results."""
    insts = mixed_instances(4, 2)
    with emv.Batch(insts) as b:
        c1, _ = b.solve()
        b.set_params(use_graph=1, graph_round_margin=-6)
        c2, st = b.solve()
        assert st.graph_fallback == 1 and st.graph_used == 0
        np.testing.assert_array_equal(c1, c2)
        b.set_params(use_graph=1, graph_round_margin=2)
        c3, st = b.solve()
        assert st.graph_used == 1
        np.testing.assert_array_equal(c1, c3)


def test_server_results_alone_equal_in_the_mix():
    R, C = 40, 50
    n = R * C
    e = emv.grid_graph(R, C, 100, 41)
    rng = np.random.default_rng(5)
    queries = [emv.distinct(n, int(rng.integers(3, 8)), 4000 + i, 2) for i in range(10)]
    # alone: one request at a time
    alone = []
    with emv.Server(n, e, arena_bytes=256 << 20, max_inflight=16) as s:
        for q in queries:
            t = s.submit(q)
            while s.poll(t).state != 2:
                s.step()
            info = s.poll(t)
            assert info.status == 0
            alone.append(info.cost)
    # in the mix: all submitted at once, stepped together at different layers
    with emv.Server(n, e, arena_bytes=256 << 20, max_inflight=16) as s:
        tickets = []
        for i, q in enumerate(queries):
            tickets.append(s.submit(q, emv.KEEP_TABLE))
            if i % 3 == 0:
                s.step()  # later requests join while earlier ones are mid-way
        while any(s.poll(t).state != 2 for t in tickets):
            s.step()
        for t, q, c in zip(tickets, queries, alone):
            info = s.poll(t)
            assert info.status == 0 and info.cost == c
            ocost, of, ocode = oracle.emv(n, e, q)
            assert c == ocost
            f, code = s.read_rows(t, 0, 1 << len(q), n)
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)
            s.release(t)


def _group_queries(n, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        g = int(rng.integers(2, 7))
        out.append([rng.choice(n, size=int(rng.integers(1, 6)), replace=False) for _ in range(g)])
    return out


@pytest.mark.parametrize("wmax", [100, 3000])
def test_server_group_queries_match_reduced_oracle(wmax):
    """Group queries on the resident graph against the oracle on the reduced
    instance (virtual terminals joined to their members by weight M, the
    reduction of README.txt:4-5): reduced cells at graph vertices are
    |X| M + the served cells, records agree outside the singleton rows, the
    cost is the reduced cost less g M; alone equals in the mix with terminal
    queries (tolerance 0).  wmax 3000 takes 32-bit cells, 100 16-bit ones."""
    R, C = 30, 40
    n = R * C
    e = emv.grid_graph(R, C, wmax, 47)
    gqs = _group_queries(n, 6, 3)
    tqs = [emv.distinct(n, int(k), 4100 + i, 2) for i, k in enumerate([3, 5, 7, 4])]
    alone = []
    with emv.Server(n, e, arena_bytes=512 << 20, max_inflight=16) as s:
        M = s.group_weight()
        assert M == 1 + int(np.asarray(e)[:, 2].sum())
        for g in gqs:
            t = s.submit_groups(g, emv.KEEP_TABLE | emv.LIST)
            while s.poll(t).state != 2:
                s.step()
            info = s.poll(t)
            assert info.status == 0 and info.groups == len(g)
            nn, red, kk, M2 = emv.group_reduce(n, e, g)
            assert M2 == M
            ocost, of, ocode = oracle.emv(nn, red, kk)
            assert info.reduced_cost == ocost and info.cost == ocost - len(g) * M
            f, code = s.read_rows(t, 0, 1 << len(g), n)
            for X in range(1, 1 << len(g)):
                np.testing.assert_array_equal(of[X, :n], f[X] + bin(X).count("1") * M)
                if X & (X - 1):
                    np.testing.assert_array_equal(ocode[X, :n], code[X])
            cert = s.certify(t)
            assert cert.ok, cert
            ok, why = validate_direct_group_tree(n, e, g, s.pairs(t), info.cost, info.root)
            assert ok, why
            alone.append((info.cost, f, code))
            s.release(t)
    with emv.Server(n, e, arena_bytes=512 << 20, max_inflight=16) as s:
        tickets = []
        for i in range(max(len(gqs), len(tqs))):
            if i < len(gqs):
                tickets.append(("g", i, s.submit_groups(gqs[i], emv.KEEP_TABLE)))
            if i < len(tqs):
                tickets.append(("t", i, s.submit(tqs[i], emv.KEEP_TABLE)))
            s.step()  # later requests join while earlier ones are mid-way
        while any(s.poll(t).state != 2 for _, _, t in tickets):
            s.step()
        for kind, i, t in tickets:
            info = s.poll(t)
            assert info.status == 0
            if kind == "g":
                cost, f0, c0 = alone[i]
                f, code = s.read_rows(t, 0, 1 << len(gqs[i]), n)
                assert info.cost == cost
                np.testing.assert_array_equal(f, f0)
                np.testing.assert_array_equal(code, c0)
            else:
                assert info.cost == oracle.emv(n, e, tqs[i])[0]
            s.release(t)


def test_server_rejects_bad_group_queries():
    R, C = 10, 10
    n = R * C
    e = emv.grid_graph(R, C, 100, 1)
    with emv.Server(n, e, arena_bytes=64 << 20, max_inflight=4) as s:
        for bad in ([[1, 2]], [[1], []], [[1], [n]], [[0]] * 31):
            with pytest.raises(emv.EmvError) as ei:
                s.submit_groups(bad)
            assert ei.value.status == -1


def test_server_admission_by_footprint():
    R, C = 30, 30
    n = R * C
    e = emv.grid_graph(R, C, 100, 43)
    with emv.Server(n, e, arena_bytes=0, max_inflight=4) as probe:
        fp8 = probe.footprint(8)
    # arena for one k=8 request at a time
    with emv.Server(n, e, arena_bytes=fp8 + 4096, max_inflight=4) as s:
        with pytest.raises(emv.EmvError) as ei:
            s.submit(emv.distinct(n, 10, 1, 2))  # 4x the arena: never fits
        assert ei.value.status == -9
        t1 = s.submit(emv.distinct(n, 8, 2, 2))
        t2 = s.submit(emv.distinct(n, 8, 3, 2))
        s.step()
        assert s.poll(t1).state == 1 and s.poll(t2).state == 0  # t2 waits for memory
        while s.poll(t2).state != 2:
            s.step()
        for t, seed in ((t1, 2), (t2, 3)):
            q = emv.distinct(n, 8, seed, 2)
            assert s.poll(t).cost == oracle.emv(n, e, q)[0]
        assert s.poll(t2).queue_ms > 0.0


def test_server_cancel_queued_and_running():
    """A cancelled request ends with ERR_CANCELLED and frees its table and
    slot; the others' results are unchanged, and a request admitted into
    the freed region afterwards is solved exactly."""
    R, C = 30, 30
    n = R * C
    e = emv.grid_graph(R, C, 100, 44)
    qs = [emv.distinct(n, k, 50 + k, 2) for k in (6, 7, 8, 5)]
    with emv.Server(n, e, arena_bytes=0, max_inflight=2) as s:
        full = s.free_bytes()
        t = [s.submit(q) for q in qs[:3]]
        s.step()
        s.step()
        assert s.poll(t[0]).state == 1 and s.poll(t[1]).state == 1 and s.poll(t[2]).state == 0
        s.cancel(t[2])  # queued
        s.cancel(t[1])  # running, mid-way through its layers
        for x in (t[1], t[2]):
            info = s.poll(x)
            assert info.state == 2 and info.status == -10
        with pytest.raises(emv.EmvError) as ei:
            s.cancel(t[1])  # already finished
        assert ei.value.status == -4
        t3 = s.submit(qs[3], emv.KEEP_TABLE)  # takes the freed slot and region
        while s.poll(t[0]).state != 2 or s.poll(t3).state != 2:
            s.step()
        for x, q in ((t[0], qs[0]), (t3, qs[3])):
            info = s.poll(x)
            assert info.status == 0 and info.cost == oracle.emv(n, e, q)[0]
        f, code = s.read_rows(t3, 0, 1 << len(qs[3]), n)
        _, of, ocode = oracle.emv(n, e, qs[3])
        np.testing.assert_array_equal(f, of)
        np.testing.assert_array_equal(code, ocode)
        s.release(t3)
        assert s.free_bytes() == full
        s.step()  # nothing left: an empty step
    with emv.Batch([(n, e, qs[0])]) as b:
        with pytest.raises(emv.EmvError) as ei:
            b._check(emv._bind_engine().emv_server_cancel(b._h, 0))
        assert ei.value.status == -4


def test_server_step_survives_injected_allocation_failure():
    """The step's scratch allocation fails (forced from the host, hipMalloc
    is not called): the step returns ERR_NOMEM with every request at its
    stage, and the next step carries on to the exact results."""
    R, C = 30, 30
    n = R * C
    e = emv.grid_graph(R, C, 100, 45)
    q1, q2 = emv.distinct(n, 5, 61, 2), emv.distinct(n, 9, 62, 2)
    with emv.Server(n, e, arena_bytes=0, max_inflight=4) as s:
        t1 = s.submit(q1)
        s.fail_alloc(0)  # the first step allocates its staging
        with pytest.raises(emv.EmvError) as ei:
            s.step()
        assert ei.value.status == -3
        info = s.poll(t1)
        assert info.state == 1 and info.stage == 1
        s.step()
        assert s.poll(t1).stage == 2
        # a larger request grows the staging: fail each of its allocations in turn
        t2 = s.submit(q2)
        for after in range(3):
            s.fail_alloc(after)
            try:
                s.step()
            except emv.EmvError as err:
                assert err.status == -3
        s.fail_alloc(-1)
        while s.poll(t1).state != 2 or s.poll(t2).state != 2:
            s.step()
        for t, q in ((t1, q1), (t2, q2)):
            info = s.poll(t)
            assert info.status == 0 and info.cost == oracle.emv(n, e, q)[0]


@pytest.mark.parametrize("wide", [emv.WIDE_LABELS, emv.TABLE32, emv.WIDE_LABELS | emv.TABLE32,
                                  emv.MERGE_1V])
def test_batch_label_widths_agree(wide):
    """Instances whose bound allows it relax on 16-bit labels (and keep
    16-bit table cells), the others on 32-bit ones, in the same step (two
    launches per round).  Forcing 32-bit labels (WIDE_LABELS), int32 cells
    (TABLE32) or one vertex per lane in the merges of 16-bit cells (MERGE_1V)
    everywhere gives the same tables and records."""
    insts = mixed_instances(6, 3)
    # one instance with large weights, whose bound needs 32-bit labels
    n = 700
    e = emv.regular_graph(n, 4, 20000, 77)
    insts.append((n, e, emv.terminals(n, 6, 77)))
    with emv.Batch(insts) as b:
        c1, _ = b.solve(emv.KEEP_TABLE)
        t1 = [b.read_rows(i, 0, 1 << len(kk), n_) for i, (n_, _, kk) in enumerate(insts)]
        c2, _ = b.solve(emv.KEEP_TABLE | wide)
        np.testing.assert_array_equal(c1, c2)
        for i, (n_, e_, kk) in enumerate(insts):
            f2, code2 = b.read_rows(i, 0, 1 << len(kk), n_)
            np.testing.assert_array_equal(t1[i][0], f2)
            np.testing.assert_array_equal(t1[i][1], code2)
            ocost, of, ocode = oracle.emv(n_, e_, kk)
            assert c1[i] == ocost
            np.testing.assert_array_equal(f2, of)
            np.testing.assert_array_equal(code2, ocode)


def _check_batch_vs_oracle(insts, costs, b):
    for i, (n, e, kk) in enumerate(insts):
        ocost, of, ocode = oracle.emv(n, e, kk)
        assert costs[i] == ocost
        f, code = b.read_rows(i, 0, 1 << len(kk), n)
        np.testing.assert_array_equal(f, of)
        np.testing.assert_array_equal(code, ocode)


def test_batch_block_relaxation_forced(monkeypatch):
    """EMV_BLOCKS=1 puts every graph on block-local relaxation (random regular
    graphs: blocks with few inside edges, many border flags): the tables and
    records equal the oracle's, eagerly and from a graph replay."""
    monkeypatch.setenv("EMV_BLOCKS", "1")
    insts = mixed_instances(6, 4)
    n = 700
    insts.append((n, emv.regular_graph(n, 4, 20000, 78), emv.terminals(n, 6, 78)))
    with emv.Batch(insts) as b:
        c1, s1 = b.solve(emv.KEEP_TABLE)
        _check_batch_vs_oracle(insts, c1, b)
        c2, s2 = b.solve(emv.KEEP_TABLE)
        assert s2.graph_used == 1 or s2.graph_fallback == 1
        _check_batch_vs_oracle(insts, c2, b)


@pytest.mark.parametrize("closure", ["1", "0"])
def test_batch_grids_and_regular_graphs_in_one_step(monkeypatch, closure):
    """Grids (block-local relaxation by the locality rule) and random regular
    graphs (vertex-wise) step together: four launch groups per round (two
    label widths x two relaxation kinds) sharing the round's flags.  Block
    visits after the first apply in-block distances (32-bit rows here: one
    grid has w up to 30000) or, with EMV_BLOCK_CLOSURE=0, sweep."""
    monkeypatch.setenv("EMV_BLOCK_CLOSURE", closure)
    insts = []
    for i, (R, C, w) in enumerate([(40, 30, 100), (33, 47, 100), (36, 36, 30000)]):
        n = R * C
        insts.append((n, emv.grid_graph(R, C, w, 60 + i), emv.terminals(n, 5 + i, 60 + i)))
    insts += mixed_instances(4, 5)
    with emv.Batch(insts) as b:
        c1, _ = b.solve(emv.KEEP_TABLE)
        _check_batch_vs_oracle(insts, c1, b)
        c2, _ = b.solve(emv.KEEP_TABLE | emv.WIDE_LABELS)
        np.testing.assert_array_equal(c1, c2)
        _check_batch_vs_oracle(insts, c2, b)


@pytest.mark.parametrize("pair,no_ell", [("1", False), ("1", True), ("0", False)])
def test_batch_paired_chunks(monkeypatch, pair, no_ell):
    """Batched solves relax 16-bit vertex-wise chunks in pairs (two chunks'
    labels side by side, 128 rows per vertex, packed saturating add and
    min), a last odd chunk with an INF partner, or one by one (EMV_PAIR=0);
    both adjacency paths.  Layers of 1 .. 4 chunks (k = 9 .. 11), a 32-bit
    label instance in the same step; tables and records equal the oracle's,
    eagerly and from the replay."""
    monkeypatch.setenv("EMV_PAIR", pair)
    if no_ell:
        monkeypatch.setenv("EMV_NO_ELL", "1")
    insts = []
    for i, (n, d, k, w) in enumerate([(640, 4, 9, 1000), (512, 6, 10, 1000), (300, 3, 11, 100),
                                      (410, 5, 7, 1000), (350, 4, 8, 200000)]):
        if (n * d) % 2:
            n += 1
        e = emv.regular_graph(n, d, w, 970 + i)
        insts.append((n, e, emv.terminals(n, k, 970 + i)))
    with emv.Batch(insts) as b:
        c1, _ = b.solve(emv.KEEP_TABLE)
        _check_batch_vs_oracle(insts, c1, b)
        c2, s2 = b.solve(emv.KEEP_TABLE)
        assert s2.graph_used == 1 or s2.graph_fallback == 1
        _check_batch_vs_oracle(insts, c2, b)


@pytest.mark.parametrize("no_split,stars,star_pairs", [(False, "1", "63"), (False, "1", "3"),
                                                        (False, "0", "63"), (True, "1", "63")])
def test_batch_split_merge_upper_layers(monkeypatch, no_split, stars, star_pairs):
    """Instances with k >= 9 reach layers whose rows have >= 128 pairs: the
    engine merges the 16-bit tables' rows in stars (rows sharing a base set,
    128-vertex tiles, a last tile half past the vertices; also the 16-bit
    rows of >= EMV_STAR_PAIRS pairs, 63 by default, 3 here for every layer
    from 3 terminals up), cuts the other rows into parts (EMV_MERGE_STAR=0
    for all), both folded with atomicMin (16/32-bit labels, 32- and 64-bit
    packed words: the last instance's bound needs 64), or, with
    EMV_NO_SPLIT, merges them whole.  Tables and records equal the oracle's,
    eagerly and from the replay."""
    monkeypatch.setenv("EMV_MERGE_STAR", stars)
    monkeypatch.setenv("EMV_STAR_PAIRS", star_pairs)
    if no_split:
        monkeypatch.setenv("EMV_NO_SPLIT", "1")
    insts = []
    for i, (n, d, k, w) in enumerate([(300, 4, 9, 1000), (260, 3, 10, 1000), (200, 5, 11, 100),
                                      (240, 4, 10, 800000), (384, 3, 12, 1000)]):
        e = emv.regular_graph(n, d, w, 950 + i)
        insts.append((n, e, emv.terminals(n, k, 950 + i)))
    with emv.Batch(insts) as b:
        c1, _ = b.solve(emv.KEEP_TABLE)
        _check_batch_vs_oracle(insts, c1, b)
        c2, s2 = b.solve(emv.KEEP_TABLE)
        assert s2.graph_used == 1 or s2.graph_fallback == 1
        _check_batch_vs_oracle(insts, c2, b)


@pytest.mark.parametrize("no_pack", [False, True])
@pytest.mark.parametrize("shape", ["regular", "grid"])
def test_server_packed_relaxation(monkeypatch, shape, no_pack):
    """Requests on one resident graph whose layers leave chunks partly
    filled share packed relaxation chunks (rows copied into a bin before the
    rounds and back after), vertex-wise (random graph) or block-local (grid);
    EMV_NO_PACK relaxes every chunk in place.  Tables and records equal the
    oracle's either way."""
    if no_pack:
        monkeypatch.setenv("EMV_NO_PACK", "1")
    if shape == "grid":
        n = 45 * 40
        e = emv.grid_graph(45, 40, 100, 71)
    else:
        n = 600
        e = emv.regular_graph(n, 5, 1000, 71)
    qs = [emv.distinct(n, k, 7100 + i, 2) for i, k in enumerate([3, 4, 5, 6, 7, 4, 8, 5])]
    with emv.Server(n, e, arena_bytes=256 << 20, max_inflight=16) as s:
        tickets = [s.submit(q, emv.KEEP_TABLE) for q in qs]
        while any(s.poll(t).state != 2 for t in tickets):
            s.step()
        for t, q in zip(tickets, qs):
            info = s.poll(t)
            ocost, of, ocode = oracle.emv(n, e, q)
            assert info.status == 0 and info.cost == ocost
            f, code = s.read_rows(t, 0, 1 << len(q), n)
            np.testing.assert_array_equal(f, of)
            np.testing.assert_array_equal(code, ocode)
