"""Known-answer graphs of the reference's own tests, rebuilt as int64/uint64 CSR arrays.

Each builder cites the reference test it mirrors. A graph is a dict with
indptr (int64, n+1), indices (int64), edge_weights (uint64), vertex_weights (uint64).
"""
from __future__ import annotations

import numpy as np


def csr_from_rows(rows, vertex_weights, default_weight=1):
    """rows: list of lists of (neighbour, weight) or of plain neighbours."""
    indptr = [0]
    indices, weights = [], []
    for row in rows:
        for item in row:
            if isinstance(item, tuple):
                u, w = item
            else:
                u, w = item, default_weight
            indices.append(int(u))
            weights.append(int(w))
        indptr.append(len(indices))
    return dict(indptr=np.array(indptr, np.int64), indices=np.array(indices, np.int64),
                edge_weights=np.array(weights, np.uint64), vertex_weights=np.array(vertex_weights, np.uint64))


def path4():
    """tests/test_uncoarsen.cu:20-47 (make_hierarchy): fine path 0-1-2-3 with weights
    1, 5, 10; its coarse level pairs {0,1} and {2,3}."""
    rows = [[(1, 1)], [(0, 1), (2, 5)], [(1, 5), (3, 10)], [(2, 10)]]
    fine = csr_from_rows(rows, [1, 1, 1, 1])
    coarse = csr_from_rows([[(1, 5)], [(0, 5)]], [2, 2])
    return fine, coarse, np.array([0, 0, 1, 1], np.int32)


def hub259(weight=7):
    """tests/test_uncoarsen.cu:121-160: hub 0 joined twice to 1..257, a chain over
    1..257, isolated tail vertex 258; every entry carries `weight`."""
    n = 259
    rows = [[] for _ in range(n)]

    def add(v, u):
        rows[v].append(u)
        rows[u].append(v)

    for v in range(1, n - 1):
        add(0, v)
        add(0, v)
        if v > 1:
            add(v - 1, v)
    rows = [sorted(r) for r in rows]
    return csr_from_rows(rows, [1] * n, default_weight=weight)


def contraction_case(groups, weight=3):
    """tests/test_uncoarsen.cu:162-231: 35 vertices, rows 0..33 list every u < 34 twice
    (self loops included), row 34 empty; map v -> v % groups. Returns the graph, the
    map, the expected coarse vertex weights and the expected coarse CSR built by
    grouping cross entries by (map[v], map[u]) and summing them."""
    n = 35
    rows = []
    expected = {}
    for v in range(n):
        row = []
        if v < n - 1:
            for u in range(n - 1):
                for _ in range(2):
                    row.append((u, weight))
                    if v % groups != u % groups:
                        key = (v % groups, u % groups)
                        expected[key] = expected.get(key, 0) + weight
        rows.append(row)
    graph = csr_from_rows(rows, [1] * n)
    mapping = np.array([v % groups for v in range(n)], np.int32)
    cw = np.bincount(mapping, minlength=groups).astype(np.uint64)
    e_ptr, e_idx, e_w = [0], [], []
    for c in range(groups):
        for (a, b), w in sorted(expected.items()):
            if a == c:
                e_idx.append(b)
                e_w.append(w)
        e_ptr.append(len(e_idx))
    return graph, mapping, cw, (np.array(e_ptr, np.int64), np.array(e_idx, np.int64), np.array(e_w, np.uint64))


def isolated_packing():
    """tests/test_uncoarsen.cu:53-76: vertices 0 and 1 joined by weight 3, ten
    isolated vertices, unequal vertex weights."""
    vw = [7, 7, 1, 4, 4, 1, 1, 8, 1, 3, 6, 2]
    rows = [[(1, 3)], [(0, 3)]] + [[] for _ in range(10)]
    return csr_from_rows(rows, vw)


def exchange():
    """tests/test_uncoarsen.cu:454-471: both parts full; two improving moves form a
    balanced cycle. Expected: cut 2, three vertices in part 0."""
    indptr = [0, 2, 5, 7, 10, 11, 12]
    indices = [1, 3, 0, 2, 4, 1, 3, 0, 2, 5, 1, 3]
    weights = [1, 4, 1, 4, 100, 4, 1, 4, 1, 100, 100, 100]
    return dict(indptr=np.array(indptr, np.int64), indices=np.array(indices, np.int64),
                edge_weights=np.array(weights, np.uint64), vertex_weights=np.ones(6, np.uint64)), \
        np.array([0, 0, 1, 1, 0, 1], np.int32)


def repair():
    """tests/test_uncoarsen.cu:472-484: an improving move overloads part 1 and the
    rebalance must move an isolated vertex back. Expected: cut 1, four in part 0,
    at least one admission iteration."""
    indptr = [0, 2, 4, 5, 5, 7, 8, 8, 8]
    indices = [1, 4, 0, 2, 1, 0, 5, 4]
    weights = [1, 4, 1, 100, 100, 4, 100, 100]
    return dict(indptr=np.array(indptr, np.int64), indices=np.array(indices, np.int64),
                edge_weights=np.array(weights, np.uint64), vertex_weights=np.ones(8, np.uint64)), \
        np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int32)


def stress(k, scale=1):
    """tests/test_uncoarsen.cu:485-518: 384-vertex circulant with a duplicated +17
    edge and self loops, rows reversed (unsorted), weights 1 + (v/k) % 3, labels
    v % k."""
    n = 384
    rows = [[] for _ in range(n)]

    def add(a, b, w):
        rows[a].append((b, w))
        rows[b].append((a, w))

    for v in range(n):
        add(v, (v + 1) % n, scale * 3)
        add(v, (v + 17) % n, scale * 11)
        add(v, (v + 17) % n, scale * 7)
        add(v, v, scale * 13)
    rows = [list(reversed(r)) for r in rows]
    vw = [1 + (v // k) % 3 for v in range(n)]
    return csr_from_rows(rows, vw), np.array([v % k for v in range(n)], np.int32)


def two_cliques(edge_weight=3, vertex_weight=2):
    """tests/test_initial_partition_gpu.cu:37-45: two 4-cliques; expected cut 0 at k=2."""
    rows = [[(u, edge_weight) for u in range(8) if u != v and u // 4 == v // 4] for v in range(8)]
    return csr_from_rows(rows, [vertex_weight] * 8)


def isolated(n, vertex_weight=1):
    return csr_from_rows([[] for _ in range(n)], [vertex_weight] * n)


def initial_hub259():
    """tests/test_initial_partition_gpu.cu:55-70: hub with doubled entries, chain,
    isolated tail; edge weight 7."""
    n = 259
    rows = []
    for v in range(n):
        row = []
        if v == 0:
            for u in range(1, 258):
                row += [u, u]
        if 0 < v < 258:
            row += [0, 0]
        if 1 < v < 258:
            row.append(v - 1)
        if 0 < v < 257:
            row.append(v + 1)
        rows.append(sorted(row))
    return csr_from_rows(rows, [1] * n, default_weight=7)


def pairs(n):
    """tests/test_uncoarsen.cu:405-415: perfect matching v <-> v ^ 1."""
    return csr_from_rows([[v ^ 1] for v in range(n)], [1] * n)


def sclp_small():
    """tests/test_sclp_small.py:24-51: a self loop, parallel edges, a disconnected
    component, an isolated vertex, a separate pair and 2043 short chains (n = 4096)."""
    n = 4096
    adj = [[] for _ in range(n)]

    def add(u, v, copies=1):
        for _ in range(copies):
            adj[u].append(v)
            adj[v].append(u)

    adj[0].append(0)
    add(0, 1, 2)
    add(1, 2)
    add(2, 3)
    add(4, 5)
    add(5, 6, 2)
    add(8, 9)
    for u in range(10, 4095, 2):
        add(u, u + 1)
    return csr_from_rows([sorted(r) for r in adj], [1] * n)


def independent_cut(g, labels):
    """Closed form of tools/verify_partition.py:42-65: directed cross weight / 2."""
    indptr, indices, w = g["indptr"], g["indices"], g["edge_weights"]
    src = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    labels = np.asarray(labels)
    cross = labels[src] != labels[indices]
    directed = int(w[cross].astype(object).sum()) if cross.any() else 0
    assert directed % 2 == 0
    return directed // 2


def loads(g, labels, k):
    out = [0] * k
    for v, p in enumerate(labels):
        out[int(p)] += int(g["vertex_weights"][v])
    return out
