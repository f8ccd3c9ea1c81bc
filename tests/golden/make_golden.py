"""Writes the golden fixtures under tests/golden/.

Expected outputs come from libraries independent of this project and of the
reference: scipy.sparse.csgraph (Dijkstra, unweighted shortest paths) and
networkx (single-source Dijkstra).  The reference ships no golden vectors for
this path (its bmk2 expected outputs are downloaded from outside the repo),
so these are the pins.  Inputs are built here with numpy's seeded generator,
except the 256x256 grid, which is produced by liblsg's generator and pinned
by a checksum of its CSR arrays.

Run:  python tests/golden/make_golden.py      (rewrites the .npz files)
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0x7FFFFFFF


def csr_from_pairs(n, src, dst, w):
    """Rows sorted by (src, dst); duplicates and self loops KEPT (the .gr format allows them)."""
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.uint32)
    return row, dst.astype(np.uint32), w.astype(np.int32)


def scipy_dist(row, col, w, src, unweighted=False):
    n = len(row) - 1
    # scipy sums duplicate entries of a csr_matrix; keep the minimum weight per (u, v) instead,
    # and drop self loops (they never shorten a path).
    u = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    keep = u != col
    u, v, ww = u[keep], col[keep].astype(np.int64), w[keep].astype(np.float64)
    if unweighted:
        ww = np.ones_like(ww)
    key = u * n + v
    order = np.lexsort((ww, key))
    key, ww = key[order], ww[order]
    first = np.ones(len(key), dtype=bool)
    first[1:] = key[1:] != key[:-1]
    key, ww = key[first], ww[first]
    # explicit zeros are dropped by scipy; encode weight 0 as a tiny epsilon and round back
    eps = 1e-9
    mat = sp.csr_matrix((np.where(ww == 0, eps, ww), (key // n, key % n)), shape=(n, n))
    if unweighted:
        d = shortest_path(mat, directed=True, unweighted=True, indices=src)
    else:
        d = sp_dijkstra(mat, directed=True, indices=src)
    out = np.full(n, INF, dtype=np.int64)
    fin = np.isfinite(d)
    out[fin] = np.rint(d[fin]).astype(np.int64)
    return out.astype(np.int32)


def networkx_dist(row, col, w, src):
    import networkx as nx
    n = len(row) - 1
    g = nx.DiGraph()
    g.add_nodes_from(range(n))
    for u in range(n):
        for e in range(row[u], row[u + 1]):
            v, ww = int(col[e]), int(w[e])
            if u == v:
                continue
            if g.has_edge(u, v):
                ww = min(ww, g[u][v]["weight"])
            g.add_edge(u, v, weight=ww)
    d = nx.single_source_dijkstra_path_length(g, src)
    out = np.full(n, INF, dtype=np.int32)
    for k, v in d.items():
        out[k] = v
    return out


def random_graph(rng, n, m, wlo, whi):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(wlo, whi + 1, m)
    return csr_from_pairs(n, src, dst, w)


def main():
    rng = np.random.default_rng(20241202)
    cases = {}
    # tiny graphs (networkx and scipy must agree), including duplicates and self loops
    for i, (n, m) in enumerate([(1, 0), (2, 1), (5, 8), (16, 40), (33, 120), (64, 300)]):
        row, col, w = random_graph(rng, n, m, 1, 20)
        cases[f"tiny{i}"] = (row, col, w, 0)
    # zero weights, a disconnected graph, a path, a star
    row, col, w = random_graph(rng, 50, 200, 0, 3)
    cases["zeroish"] = (row, col, w, 3)
    n = 40
    src = np.concatenate([rng.integers(0, 20, 80), rng.integers(20, 40, 80)])
    dst = np.concatenate([rng.integers(0, 20, 80), rng.integers(20, 40, 80)])
    cases["twoparts"] = csr_from_pairs(n, src, dst, rng.integers(1, 50, 160)) + (0,)
    # medium graphs: sparse random and a heavy-tailed one with a hub
    row, col, w = random_graph(rng, 2000, 16000, 1, 255)
    cases["medium"] = (row, col, w, 7)
    hub_src = np.concatenate([np.zeros(3000, dtype=np.int64), rng.integers(0, 3000, 9000)])
    hub_dst = np.concatenate([rng.integers(0, 3000, 3000), rng.integers(0, 3000, 9000)])
    cases["hub"] = csr_from_pairs(3000, hub_src, hub_dst, rng.integers(1, 255, 12000)) + (0,)

    out = {}
    for name, (row, col, w, s) in cases.items():
        d = scipy_dist(row, col, w, s)
        b = scipy_dist(row, col, w, s, unweighted=True)
        if len(row) - 1 <= 64:
            dn = networkx_dist(row, col, w, s)
            assert np.array_equal(d, dn), name
        out[f"{name}_row"] = row
        out[f"{name}_col"] = col
        out[f"{name}_w"] = w
        out[f"{name}_src"] = np.array([s], dtype=np.int64)
        out[f"{name}_dist"] = d
        out[f"{name}_bfs"] = b
    np.savez_compressed(os.path.join(HERE, "small_graphs.npz"), names=np.array(sorted(cases)), **out)

    # C1: 256x256 grid from liblsg's generator, w in [1, 100], seed 1, source 0
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from galoisgpu_amd import CSRGraph
    g = CSRGraph.grid(256, 256, 1, 100, 1)
    row, col, w = g.row_start.copy(), g.edge_dst.copy(), g.edge_data.copy()
    h = hashlib.sha256(row.tobytes() + col.tobytes() + w.tobytes()).hexdigest()
    d = scipy_dist(row, col, w, 0)
    b = scipy_dist(row, col, w, 0, unweighted=True)
    np.savez_compressed(os.path.join(HERE, "grid256.npz"), sha256=np.array(h), dist=d, bfs=b,
                        nnodes=np.array(g.nnodes), nedges=np.array(g.nedges))
    print("wrote", len(cases), "small cases and grid256", h[:16])


if __name__ == "__main__":
    main()
