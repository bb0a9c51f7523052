"""The seeded instance generator behind the workloads: structure, ranges,
determinism, and the group-reduction identity on tiny inputs (CPU)."""
import itertools

import numpy as np
import pytest

import emv
import oracle
from test_oracle import brute_steiner


def degrees(n, e):
    deg = np.zeros(n, dtype=np.int64)
    np.add.at(deg, e[:, 0], 1)
    np.add.at(deg, e[:, 1], 1)
    return deg


def is_connected(n, e):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for u, v, _ in e:
        parent[find(u)] = find(v)
    return len({find(v) for v in range(n)}) == 1


@pytest.mark.parametrize("n,d,method", [(1000, 4, "reject"), (4096, 8, "repair"),
                                        (1025 + 1, 3, "repair"), (20000, 6, "repair")])
def test_regular_simple_connected(n, d, method):
    e = emv.regular_graph(n, d, 1000, 7, method)
    assert len(e) == n * d // 2
    assert np.all(degrees(n, e) == d)
    assert np.all(e[:, 0] != e[:, 1])
    key = np.minimum(e[:, 0], e[:, 1]) * n + np.maximum(e[:, 0], e[:, 1])
    assert len(np.unique(key)) == len(key)
    assert is_connected(n, e)
    assert e[:, 2].min() >= 1 and e[:, 2].max() <= 1000
    e2 = emv.regular_graph(n, d, 1000, 7, method)
    np.testing.assert_array_equal(e, e2)


def test_regular_1000_fixture_is_reproduced():
    import os
    from golden_data import GOLD
    from stputil import parse_stp
    inst = parse_stp(os.path.join(GOLD, "regular_n1000_d4_k8_s1.stp"))
    e = emv.regular_graph(1000, 4, 100, 1, "reject")
    np.testing.assert_array_equal(inst.edges, e)
    np.testing.assert_array_equal(inst.terminals, emv.terminals(1000, 8, 1))


def test_grid_structure():
    R, C = 7, 9
    e = emv.grid_graph(R, C, 100, 3)
    assert len(e) == R * (C - 1) + (R - 1) * C
    diff = np.abs(e[:, 0] - e[:, 1])
    assert set(diff.tolist()) == {1, C}
    assert is_connected(R * C, e)
    assert e[:, 2].min() >= 1 and e[:, 2].max() <= 100


def test_terminals_and_groups_distinct():
    kk = emv.terminals(100, 30, 4)
    assert len(set(kk.tolist())) == 30 and kk.min() >= 0 and kk.max() < 100
    g = emv.groups(1000, 13, 32, 5)
    for row in g:
        assert len(set(row.tolist())) == 32


def brute_group(n, edges, groups):
    best = None
    for pick in itertools.product(*groups):
        c = brute_steiner(n, edges, sorted(set(int(x) for x in pick)))
        best = c if best is None else min(best, c)
    return best


@pytest.mark.parametrize("seed", range(3))
def test_group_reduction_identity(seed):
    """OPT_group = OPT(reduced) - g*M, M = 1 + sum of weights."""
    rng = np.random.default_rng(seed)
    n = 6
    edges = [(int(rng.integers(0, v)), v, int(rng.integers(1, 9))) for v in range(1, n)]
    for _ in range(3):
        u, v = rng.integers(0, n, 2)
        if u != v:
            edges.append((int(u), int(v), int(rng.integers(1, 9))))
    edges = np.array(edges, dtype=np.int64)
    groups = [rng.choice(n, 2, replace=False) for _ in range(3)]
    nn, red, kk, M = emv.group_reduce(n, edges, groups)
    cost, _, _ = oracle.emv(nn, red, kk)
    assert cost - len(groups) * M == brute_group(n, edges, groups)
