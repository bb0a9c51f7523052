"""STP parsing and Steiner-tree checks shared by the tests (no oracle here).

parse_stp() accepts the same lower-case DIMACS STP subset as reader-el
(reader/reader-el.c:473-618) and returns 0-based edges and terminals.
"""
import heapq

import numpy as np

INF = np.iinfo(np.int64).max


class Instance:
    def __init__(self, n, edges, terminals, cost=-1, name=""):
        self.n = n
        self.edges = np.asarray(edges, dtype=np.int64).reshape(-1, 3)
        self.terminals = np.asarray(terminals, dtype=np.int64)
        self.cost = cost
        self.name = name

    @property
    def m(self):
        return len(self.edges)

    @property
    def k(self):
        return len(self.terminals)


def parse_stp(path):
    n = 0
    edges, terms = [], []
    cost = -1
    with open(path) as fh:
        for line in fh:
            tok = line.split(" ")[0]
            if tok == "nodes":
                n = int(line.split()[1])
            elif tok == "e":
                _, u, v, w = line.split()[:4]
                edges.append((int(u) - 1, int(v) - 1, int(w)))
            elif tok == "t":
                terms.append(int(line.split()[1]) - 1)
            elif tok == "cost":
                cost = int(line.split()[1])
    return Instance(n, edges, terms, cost, path)


def write_stp(path, inst):
    with open(path, "w") as fh:
        fh.write("33d32945 stp file, stp format version 1.0\n\n")
        fh.write("section graph\n")
        fh.write(f"nodes {inst.n}\nedges {inst.m}\n")
        for u, v, w in inst.edges:
            fh.write(f"e {u + 1} {v + 1} {w}\n")
        fh.write("end\n\nsection terminals\n")
        fh.write(f"terminals {inst.k}\n")
        for t in inst.terminals:
            fh.write(f"t {t + 1}\n")
        fh.write("end\n\neof\n")


def adjacency(n, edges):
    adj = [[] for _ in range(n)]
    for u, v, w in np.asarray(edges).reshape(-1, 3):
        adj[u].append((int(v), int(w)))
        adj[v].append((int(u), int(w)))
    return adj


def dijkstra(adj, s):
    d = [INF] * len(adj)
    d[s] = 0
    pq = [(0, s)]
    while pq:
        du, u = heapq.heappop(pq)
        if du > d[u]:
            continue
        for v, w in adj[u]:
            if du + w < d[v]:
                d[v] = du + w
                heapq.heappush(pq, (d[v], v))
    return d


def check_listed_tree(inst, pairs, cost):
    """The reference's -list contract plus the metric: the pairs form a tree
    (connected, |E| = |V|-1) over a vertex set containing every terminal,
    and the sum of shortest-path distances over the pairs equals `cost`
    (singleton records list metric-closure pairs, reader-el.c:2025)."""
    adj = adjacency(inst.n, inst.edges)
    pairs = [tuple(int(x) for x in p) for p in pairs]
    verts = set()
    for u, v in pairs:
        verts.add(u)
        verts.add(v)
    terms = set(int(t) for t in inst.terminals)
    if len(terms) == 1 and not pairs:
        return True, "single terminal"
    if not terms <= verts:
        return False, f"terminals missing: {sorted(terms - verts)[:5]}"
    if len(pairs) != len(verts) - 1:
        return False, f"{len(pairs)} pairs over {len(verts)} vertices is not a tree"
    parent = {v: v for v in verts}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for u, v in pairs:
        ru, rv = find(u), find(v)
        if ru == rv:
            return False, "cycle"
        parent[ru] = rv
    total = 0
    cache = {}
    for u, v in pairs:
        if v not in cache:
            cache[v] = dijkstra(adj, v)
        total += cache[v][u]
    if total != cost:
        return False, f"pair distances sum to {total}, cost is {cost}"
    return True, "ok"
