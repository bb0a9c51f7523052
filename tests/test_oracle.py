"""Pins the CPU oracle (oracle/emv_oracle.c) before anything is checked
against it.

* golden vectors produced by the reference itself (tests/golden/, written by
  make_golden.py from oracle/_ref): full-table SHA-256, record-kind SHA-256
  and cost for all 207 SteinLib instances of the reference testset plus the
  n=1000 4-regular instance;
* SteinLib optimal cost lines (independent of the reference);
* scipy's Dijkstra for the singleton rows;
* brute-force minimum Steiner trees on tiny graphs for every cell;
* the fixed-point certificate of the table;
* validity of the oracle's and the reference's listed trees.
Each check is chosen so that a plausible oracle slip (dropped term, wrong
index, transposed operand, wrong tie rule) fails at least one of them.
"""
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

import oracle
from stputil import INF, Instance, adjacency, check_listed_tree, dijkstra

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _meta():
    with open(os.path.join(GOLD, "ref_golden.json")) as fh:
        return json.load(fh)


META = _meta()
TESTSET = np.load(os.path.join(GOLD, "testset.npz"))
TABLES = np.load(os.path.join(GOLD, "tables.npz"))
NAMES = sorted(n for n in META if not n.startswith("regular"))


def load_instance(name):
    if name.startswith("regular"):
        from stputil import parse_stp
        return parse_stp(os.path.join(GOLD, name + ".stp"))
    key = name.replace("/", "__")
    return Instance(META[name]["n"], TESTSET[key + "__edges"].astype(np.int64),
                    TESTSET[key + "__terms"].astype(np.int64), META[name]["stp_cost"], name)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i8").tobytes()).hexdigest()


def kinds(code, k):
    out = code.copy()
    out[code >= (1 << k)] = 1 << k
    return out


def test_golden_costs_match_steinlib():
    # the reference itself agrees with the SteinLib optimum on every instance
    for name in NAMES:
        mt = META[name]
        assert mt["ref_rc"] == 0, name
        assert mt["ref_cost"] == mt["stp_cost"] == mt["kernel_cost"], name


@pytest.mark.parametrize("name", NAMES)
def test_oracle_table_matches_reference(name):
    inst = load_instance(name)
    mt = META[name]
    cost, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    assert cost == mt["ref_cost"] == inst.cost
    assert digest(f) == mt["f_sha256"]
    # merge records (first strict minimum over increasing Y) and the set of
    # edge-record cells are the reference's exactly
    assert digest(kinds(code, inst.k)) == mt["kind_sha256"]


def test_oracle_regular_1000_full_table():
    name = "regular_n1000_d4_k8_s1"
    inst = load_instance(name)
    cost, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    ref_f = TABLES[name + "__f"].astype(np.int64)
    ref_f[ref_f == -1] = INF
    np.testing.assert_array_equal(f, ref_f)
    assert cost == META[name]["ref_cost"]
    ref_code = TABLES[name + "__code"].astype(np.int64)
    np.testing.assert_array_equal(kinds(code, inst.k), kinds(ref_code, inst.k))


def test_oracle_small_tables_elementwise():
    # every stored small table, cell by cell (digests above cover the rest)
    for key in TABLES.files:
        if not key.endswith("__f") or key.startswith("regular"):
            continue
        name = key[:-3].replace("__", "/")
        inst = load_instance(name)
        _, f, _ = oracle.emv(inst.n, inst.edges, inst.terminals)
        ref = TABLES[key].astype(np.int64)
        ref[ref == -1] = INF
        np.testing.assert_array_equal(f, ref, err_msg=name)


def test_edge_records_are_shortest_path_predecessors():
    for name in ["small-instances/B/b01", "small-instances/C/c01", "regular_n1000_d4_k8_s1"]:
        inst = load_instance(name)
        _, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
        S = 1 << inst.k
        adj = adjacency(inst.n, inst.edges)
        for X in range(1, S):
            if bin(X).count("1") == 1:
                continue
            for v in range(inst.n):
                c = code[X, v]
                if c >= S:
                    u = c - S
                    ws = [w for (x, w) in adj[v] if x == u]
                    assert ws and f[X, u] + min(ws) == f[X, v]
                    # the smallest such neighbour
                    cands = [x for (x, w) in adj[v] if f[X, x] + w == f[X, v]]
                    assert u == min(cands)
                else:
                    assert 0 < c < X and (c & ~X) == 0
                    assert f[c, v] + f[X ^ c, v] == f[X, v]


def test_singletons_match_scipy():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import dijkstra as sp_dijkstra

    for name in ["small-instances/D/d01", "small-instances/I080/i080-001", "regular_n1000_d4_k8_s1"]:
        inst = load_instance(name)
        e = inst.edges
        # keep the lightest of parallel edges, as a shortest path does
        g = coo_matrix((e[:, 2].astype(float), (e[:, 0], e[:, 1])), shape=(inst.n, inst.n)).tocsr()
        g.sum_duplicates()
        w = {}
        for u, v, ww in e:
            key = (min(u, v), max(u, v))
            w[key] = min(w.get(key, INF), ww)
        rows = [a for a, b in w]
        cols = [b for a, b in w]
        g = coo_matrix(([w[x] for x in w], (rows, cols)), shape=(inst.n, inst.n)).tocsr()
        dist = sp_dijkstra(g, directed=False, indices=inst.terminals)
        _, f, _ = oracle.emv(inst.n, inst.edges, inst.terminals)
        for t in range(inst.k):
            np.testing.assert_array_equal(f[1 << t], dist[t].astype(np.int64))


def brute_steiner(n, edges, req):
    """Minimum Steiner tree over the vertex set req by enumerating Steiner
    vertex subsets and taking the MST of the induced subgraph."""
    req = set(req)
    best = INF
    others = [v for v in range(n) if v not in req]
    wmin = {}
    for u, v, w in edges:
        if u == v:
            continue
        key = (min(u, v), max(u, v))
        wmin[key] = min(wmin.get(key, INF), w)
    for r in range(len(others) + 1):
        for extra in itertools.combinations(others, r):
            vs = req | set(extra)
            es = sorted((w, a, b) for (a, b), w in wmin.items() if a in vs and b in vs)
            parent = {v: v for v in vs}

            def find(x):
                while parent[x] != x:
                    parent[x] = parent[parent[x]]
                    x = parent[x]
                return x

            tot, cnt = 0, 0
            for w, a, b in es:
                ra, rb = find(a), find(b)
                if ra != rb:
                    parent[ra] = rb
                    tot += w
                    cnt += 1
            if cnt == len(vs) - 1:
                best = min(best, tot)
    return best


@pytest.mark.parametrize("seed", range(6))
def test_oracle_bruteforce_tiny(seed):
    rng = np.random.default_rng(seed)
    n = 7
    # random connected multigraph: a spanning tree plus extra edges (some parallel)
    edges = [(int(rng.integers(0, v)), v, int(rng.integers(1, 9))) for v in range(1, n)]
    for _ in range(5):
        u, v = rng.integers(0, n, 2)
        edges.append((int(u), int(v), int(rng.integers(1, 9))))
    edges = np.array(edges, dtype=np.int64)
    k = 4
    kk = rng.choice(n, k, replace=False)
    cost, f, _ = oracle.emv(n, edges, kk)
    for X in range(1, 1 << k):
        for v in range(n):
            req = [int(kk[t]) for t in range(k) if X >> t & 1] + [v]
            assert f[X, v] == brute_steiner(n, edges, req), (X, v)
    assert cost == f[(1 << (k - 1)) - 1, kk[k - 1]]


def fixed_point_ok(n, edges, kk, f):
    """f[X][v] <= f[X][u] + w(u,v), f[X][v] <= f[Y][v] + f[X\\Y][v], with
    equality attained (or |X| = 1 and v = kk[t], value 0)."""
    k = len(kk)
    e = np.asarray(edges).reshape(-1, 3)
    for X in range(1, 1 << k):
        row = f[X]
        best = np.full(n, INF)
        np.minimum.at(best, e[:, 1], row[e[:, 0]] + e[:, 2])
        np.minimum.at(best, e[:, 0], row[e[:, 1]] + e[:, 2])
        Y = (X - 1) & X
        while Y:
            best = np.minimum(best, f[Y] + f[X ^ Y])
            Y = (Y - 1) & X
        if bin(X).count("1") == 1:
            t = X.bit_length() - 1
            best[kk[t]] = min(best[kk[t]], 0)
        if not np.array_equal(best, row):
            return False
    return True


def test_oracle_fixed_point_certificate():
    inst = load_instance("regular_n1000_d4_k8_s1")
    _, f, _ = oracle.emv(inst.n, inst.edges, inst.terminals)
    assert fixed_point_ok(inst.n, inst.edges, inst.terminals, f)
    # a planted slip fails the certificate
    g = f.copy()
    g[37, 5] += 1
    assert not fixed_point_ok(inst.n, inst.edges, inst.terminals, g)


@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/P6E/p619",
                                  "small-instances/C/c06", "regular_n1000_d4_k8_s1"])
def test_oracle_traceback_valid_tree(name):
    inst = load_instance(name)
    cost, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    C = (1 << (inst.k - 1)) - 1
    pairs = oracle.traceback(inst.n, inst.k, code, int(inst.terminals[-1]), C)
    ok, why = check_listed_tree(inst, pairs, cost)
    assert ok, why


def test_reference_listed_trees_pass_validator():
    # the validator accepts the reference's own -list output
    for name in ["small-instances/B/b01", "small-instances/D/d01", "regular_n1000_d4_k8_s1"]:
        inst = load_instance(name)
        sol = np.array(META[name]["ref_solution"], dtype=np.int64) - 1
        ok, why = check_listed_tree(inst, sol, META[name]["ref_cost"])
        assert ok, why


def test_readme_b01_solution_line():
    # README.txt:87 prints this solution for b01
    sol = ["24 28", "28 18", "18 43", "43 22", "22 12", "22 41", "41 49", "41 37", "22 20",
           "20 48", "20 35", "20 27", "27 34"]
    got = [f"{a} {b}" for a, b in META["small-instances/B/b01"]["ref_solution"]]
    assert got == sol


def test_oracle_sssp_pred():
    inst = load_instance("small-instances/B/b01")
    d, p = oracle.sssp(inst.n, inst.edges, 3)
    adj = adjacency(inst.n, inst.edges)
    ref = dijkstra(adj, 3)
    assert list(d) == ref
    assert p[3] == -1
    for v in range(inst.n):
        if v != 3:
            assert any(u == p[v] and d[u] + w == d[v] for u, w in adj[v])
