"""Checks of solver output at full workload sizes, on the CPU, independent of
the solver: scipy's Dijkstra for distances, numpy for the fixed point.

validate_tree()           a -list tree (reader-el.c:1891-1904): the pairs form a
                          tree over a vertex set holding every terminal; each
                          pair is a graph edge or a metric-closure shortcut
                          (one endpoint a terminal, reader-el.c:2025); the
                          pairs' lengths sum to the cost
validate_group_tree()     the tree of a group-reduced instance (virtual
                          terminals n..n+g-1, README.txt:4-5): a valid tree
                          of the reduced instance whose virtual terminals are
                          leaves; expanded into graph paths it is a group
                          Steiner tree of cost (cost - g M)
validate_direct_group_tree()
                          a tree of graph edges touching every group, of the
                          given cost (the server's group queries)
row_fixed_point()         one table row X against its edge and split
                          candidates, from the rows of X's subsets
"""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

INF = np.iinfo(np.int64).max


def min_weight_csr(n, edges):
    """Symmetric CSR holding, for each vertex pair, the least weight among
    its parallel edges (scipy would sum duplicates)."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 3)
    u = np.concatenate([e[:, 0], e[:, 1]])
    v = np.concatenate([e[:, 1], e[:, 0]])
    w = np.concatenate([e[:, 2], e[:, 2]])
    key = u * n + v
    order = np.lexsort((w, key))
    key, w = key[order], w[order]
    first = np.ones(len(key), dtype=bool)
    first[1:] = key[1:] != key[:-1]
    key, w = key[first], w[first]
    return sp.csr_matrix((w.astype(np.float64), (key // n, key % n)), shape=(n, n))


def edge_weights(csr, pairs):
    """Least weight of each pair's edge, -1 where the pair is no edge."""
    out = np.full(len(pairs), -1, dtype=np.int64)
    for i, (a, b) in enumerate(pairs):
        lo, hi = csr.indptr[a], csr.indptr[a + 1]
        j = np.searchsorted(csr.indices[lo:hi], b)
        if j < hi - lo and csr.indices[lo + j] == b:
            out[i] = int(csr.data[lo + j])
    return out


def _is_tree(pairs, verts):
    if len(pairs) != len(verts) - 1:
        return False
    parent = {v: v for v in verts}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in pairs:
        ra, rb = find(a), find(b)
        if ra == rb:
            return False
        parent[ra] = rb
    return True


def validate_tree(n, edges, terminals, pairs, cost, csr=None):
    """(ok, why) for a listed tree of `cost` over terminals."""
    csr = min_weight_csr(n, edges) if csr is None else csr
    pairs = [(int(a), int(b)) for a, b in np.asarray(pairs).reshape(-1, 2)]
    terms = sorted(set(int(t) for t in terminals))
    if not pairs:
        return (len(terms) == 1 and cost == 0), "empty tree"
    verts = set(x for p in pairs for x in p)
    if not set(terms) <= verts:
        return False, f"terminals missing: {sorted(set(terms) - verts)[:5]}"
    if not _is_tree(pairs, verts):
        return False, "pairs do not form a tree"
    tset = set(terms)
    need = sorted(set(x for p in pairs for x in p if x in tset))
    dist = dijkstra(csr, indices=need) if need else None
    row = {t: i for i, t in enumerate(need)}
    w = edge_weights(csr, pairs)
    total = 0
    for (a, b), we in zip(pairs, w):
        if a in row or b in row:
            t, o = (a, b) if a in row else (b, a)
            total += int(dist[row[t], o])
        elif we >= 0:
            total += int(we)
        else:
            return False, f"pair ({a}, {b}) is neither an edge nor a shortcut from a terminal"
    if total != cost:
        return False, f"pairs sum to {total}, cost is {cost}"
    return True, "ok"


def validate_group_tree(n, edges, groups, nn, red_edges, M, pairs, cost):
    """(ok, why, group_cost, edges) for the listed tree of the reduced
    instance; edges = the expanded group Steiner tree as graph vertex pairs."""
    g = len(groups)
    csr = min_weight_csr(nn, red_edges)
    virt = list(range(n, n + g))
    ok, why = validate_tree(nn, red_edges, virt, pairs, cost, csr)
    if not ok:
        return False, "reduced tree: " + why, None, None
    pairs = [(int(a), int(b)) for a, b in np.asarray(pairs).reshape(-1, 2)]
    real, attach = [], {}
    for a, b in pairs:
        va, vb = a >= n, b >= n
        if va and vb:
            return False, "pair between two virtual terminals", None, None
        if va or vb:
            t, x = (a, b) if va else (b, a)
            if t in attach:
                return False, f"virtual terminal {t} is not a leaf", None, None
            attach[t] = x
        else:
            real.append((a, b))
    # expand: each virtual leaf (t_j, x) is the edge t_j - s plus a shortest
    # graph path s .. x from the nearest member s of group j
    gcsr = min_weight_csr(n, edges)
    union = set((min(a, b), max(a, b)) for a, b in real)
    touched = set(x for p in real for x in p) | set(attach.values())
    for j, mem in enumerate(groups):
        x = attach[n + j]
        d, pred, _ = dijkstra(gcsr, indices=[int(s) for s in mem], min_only=True,
                              return_predecessors=True)
        y = x
        while pred[y] >= 0 and y not in set(int(s) for s in mem):
            union.add((min(y, int(pred[y])), max(y, int(pred[y]))))
            y = int(pred[y])
        touched.add(y)
        if y not in set(int(s) for s in mem):
            return False, f"group {j}: path from {x} ends at {y}, not a member", None, None
    w = edge_weights(gcsr, list(union))
    if (w < 0).any():
        return False, "expanded tree uses a non-edge", None, None
    gcost = int(w.sum())
    want = cost - g * M
    verts = set(x for p in union for x in p) | touched
    if gcost != want:
        return False, f"expanded tree costs {gcost}, cost - g M is {want}", None, None
    if union and not _connected(union, verts):
        return False, "expanded tree is not connected", None, None
    for j, mem in enumerate(groups):
        if not verts & set(int(s) for s in mem):
            return False, f"group {j} untouched", None, None
    return True, "ok", gcost, sorted(union)


def _connected(edge_set, verts):
    parent = {v: v for v in verts}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in edge_set:
        parent[find(a)] = find(b)
    return len(set(find(v) for v in verts)) == 1


def validate_direct_group_tree(n, edges, groups, pairs, cost, root):
    """(ok, why) for a group query's listed tree: graph edges only."""
    pairs = [(int(a), int(b)) for a, b in np.asarray(pairs).reshape(-1, 2)]
    verts = set(x for p in pairs for x in p) | {int(root)}
    for j, mem in enumerate(groups):
        if not verts & set(int(s) for s in mem):
            return False, f"group {j} untouched"
    if pairs and not _is_tree(pairs, verts):
        return False, "pairs do not form a tree"
    w = edge_weights(min_weight_csr(n, edges), pairs)
    if (w < 0).any():
        return False, "a pair is not a graph edge"
    if int(w.sum()) != cost:
        return False, f"edges sum to {int(w.sum())}, cost is {cost}"
    return True, "ok"


def row_fixed_point(edges, X, rows, sources=None):
    """True when row X equals, at every vertex, the least of its edge
    candidates, its split candidates (rows of X's subsets, from `rows`) and,
    for a singleton row, 0 at its sources."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 3)
    row = rows[X]
    n = len(row)
    best = np.full(n, INF)
    fin0, fin1 = row[e[:, 0]] != INF, row[e[:, 1]] != INF
    np.minimum.at(best, e[fin0, 1], row[e[fin0, 0]] + e[fin0, 2])
    np.minimum.at(best, e[fin1, 0], row[e[fin1, 1]] + e[fin1, 2])
    Y = (X - 1) & X
    while Y:
        a, b = rows[Y], rows[X ^ Y]
        s = np.where((a == INF) | (b == INF), INF, a + b)
        best = np.minimum(best, s)
        Y = (Y - 1) & X
    if X & (X - 1) == 0 and sources is not None:
        best[np.asarray(sources, dtype=np.int64)] = 0
    return np.array_equal(best, row)


def subsets(X):
    """X and all its nonempty proper subsets."""
    out = [X]
    Y = (X - 1) & X
    while Y:
        out.append(Y)
        Y = (Y - 1) & X
    return out
