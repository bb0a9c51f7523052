"""The full-size checkers (tests/bigcheck.py) against the pinned oracle on
small instances: each accepts the oracle's output and rejects planted
mistakes, so a pass at full size means what it says."""
import numpy as np
import pytest

import emv
import oracle
from bigcheck import (INF, row_fixed_point, subsets, validate_direct_group_tree,
                      validate_group_tree, validate_tree)
from golden_data import load_instance


@pytest.mark.parametrize("name", ["small-instances/B/b01", "small-instances/C/c06",
                                  "regular_n1000_d4_k8_s1"])
def test_validate_tree_accepts_oracle_trees(name):
    inst = load_instance(name)
    cost, f, code = oracle.emv(inst.n, inst.edges, inst.terminals)
    C = (1 << (inst.k - 1)) - 1
    pairs = oracle.traceback(inst.n, inst.k, code, int(inst.terminals[-1]), C)
    ok, why = validate_tree(inst.n, inst.edges, inst.terminals, pairs, cost)
    assert ok, why
    # planted: a wrong cost, a dropped pair, a pair that is no edge
    assert not validate_tree(inst.n, inst.edges, inst.terminals, pairs, cost + 1)[0]
    assert not validate_tree(inst.n, inst.edges, inst.terminals, pairs[1:], cost)[0]
    bad = pairs.copy()
    nonterm = [v for v in range(inst.n) if v not in set(inst.terminals.tolist())]
    j = [i for i, (a, b) in enumerate(bad) if a not in inst.terminals and b not in inst.terminals]
    if j:
        bad[j[0], 1] = [v for v in nonterm if v not in bad.reshape(-1)][0]
        assert not validate_tree(inst.n, inst.edges, inst.terminals, bad, cost)[0]


def _tiny_group_instance(seed):
    rows, cols = 9, 11
    e = emv.grid_graph(rows, cols, 30, seed)
    n = rows * cols
    rng = np.random.default_rng(seed)
    groups = [rng.choice(n, size=int(rng.integers(2, 4)), replace=False) for _ in range(4)]
    nn, red, kk, M = emv.group_reduce(n, e, groups)
    return n, e, groups, nn, red, kk, M


def _brute_group_cost(n, e, groups):
    best = INF
    import itertools
    for pick in itertools.product(*[list(g) for g in groups]):
        kk = np.unique(np.asarray(pick, dtype=np.int64))
        if len(kk) == 1:
            return 0
        c, _, _ = oracle.emv(n, e, kk)
        best = min(best, c)
    return best


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_validate_group_tree_on_oracle_reduced_instance(seed):
    n, e, groups, nn, red, kk, M = _tiny_group_instance(seed)
    cost, f, code = oracle.emv(nn, red, kk)
    C = (1 << (len(kk) - 1)) - 1
    pairs = oracle.traceback(nn, len(kk), code, int(kk[-1]), C)
    ok, why, gcost, tree = validate_group_tree(n, e, groups, nn, red, M, pairs, cost)
    assert ok, why
    assert gcost == cost - len(groups) * M == _brute_group_cost(n, e, groups)
    # the expanded tree is a direct group tree of that cost
    root = int(tree[0][0]) if tree else int(groups[-1][0])
    ok, why = validate_direct_group_tree(n, e, groups, tree, gcost, root)
    assert ok, why
    assert not validate_direct_group_tree(n, e, groups, tree, gcost + 1, root)[0]
    assert not validate_group_tree(n, e, groups, nn, red, M, pairs, cost + 1)[0]
    # a group the tree does not touch
    far = [[v for v in range(n) if v not in set(x for p in tree for x in p)][:1]]
    if tree and far[0]:
        assert not validate_direct_group_tree(n, e, list(groups) + far, tree, gcost, root)[0]


def test_row_fixed_point_accepts_table_and_rejects_slips():
    inst = load_instance("regular_n1000_d4_k8_s1")
    _, f, _ = oracle.emv(inst.n, inst.edges, inst.terminals)
    kk = inst.terminals
    for X in [1, 4, 3, 37, 200, 127, 255]:
        rows = {Y: f[Y] for Y in subsets(X)}
        src = [kk[X.bit_length() - 1]] if X & (X - 1) == 0 else None
        assert row_fixed_point(inst.edges, X, rows, src), X
        g = dict(rows)
        g[X] = rows[X].copy()
        g[X][17] -= 1
        assert not row_fixed_point(inst.edges, X, g, src), X
        g[X] = rows[X].copy()
        g[X][17] += 1
        assert not row_fixed_point(inst.edges, X, g, src), X
