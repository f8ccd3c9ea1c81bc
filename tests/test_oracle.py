"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against (a) golden vectors from scipy and
networkx (tests/golden/), (b) closed forms, (c) brute-force Bellman-Ford on
tiny graphs, (d) the O(E) certificate, and the tests show that planted
mistakes (dropped weight term, off-by-one row bound, wrong comparison, wrong
BFS start level) fail at least one pin.
"""
import itertools

import numpy as np
import pytest

import oracle
from oracle import INF


def bellman_ford(row, col, w, src, unit=False):
    n = len(row) - 1
    d = [INF] * n
    d[src] = 0
    for _ in range(n):
        changed = False
        for u in range(n):
            if d[u] == INF:
                continue
            for e in range(row[u], row[u + 1]):
                nd = d[u] + (1 if unit else int(w[e]))
                if nd < d[col[e]]:
                    d[col[e]] = nd
                    changed = True
        if not changed:
            break
    return np.array(d, dtype=np.int32)


def grid_csr(rows, cols, weight=None):
    row = [0]
    col, w = [], []
    for r in range(rows):
        for c in range(cols):
            for rr, cc in ((r - 1, c), (r, c - 1), (r, c + 1), (r + 1, c)):
                if 0 <= rr < rows and 0 <= cc < cols:
                    col.append(rr * cols + cc)
                    w.append(1 if weight is None else weight)
            row.append(len(col))
    return np.array(row, np.uint32), np.array(col, np.uint32), np.array(w, np.int32)


def check_all(g, src, delta=3):
    """Runs every oracle function on one graph; returns a list of failures."""
    bad = []
    d = oracle.dijkstra(g["row"], g["col"], g["w"], src)
    if not np.array_equal(d, g["dist"]):
        bad.append("dijkstra")
    nf, _ = oracle.nearfar(g["row"], g["col"], g["w"], src, delta)
    if not np.array_equal(nf, g["dist"]):
        bad.append("nearfar")
    b, _ = oracle.bfs(g["row"], g["col"], src)
    if not np.array_equal(b, g["bfs"]):
        bad.append("bfs")
    return bad


def test_golden_vectors(small_graphs):
    for name, g in small_graphs.items():
        for delta in (1, 3, 32, 10000):
            assert check_all(g, g["src"], delta) == [], (name, delta)


def test_grid256_golden(grid256_golden):
    """Config 1 (256x256 grid, w in [1,100], source 0) against scipy's distances."""
    from galoisgpu_amd import CSRGraph
    import hashlib
    g = CSRGraph.grid(256, 256, 1, 100, 1)
    row, col, w = g.row_start, g.edge_dst, g.edge_data
    assert hashlib.sha256(row.tobytes() + col.tobytes() + w.tobytes()).hexdigest() == str(grid256_golden["sha256"])
    assert np.array_equal(oracle.dijkstra(row, col, w, 0), grid256_golden["dist"])
    nf, st = oracle.nearfar(row, col, w, 0, 32)
    assert np.array_equal(nf, grid256_golden["dist"])
    assert st["phases"] >= grid256_golden["dist"].max() // 32
    b, _ = oracle.bfs(row, col, 0)
    assert np.array_equal(b, grid256_golden["bfs"])


def test_closed_forms():
    # unit-weight grid: Manhattan distance from the source
    row, col, w = grid_csr(9, 13)
    src = 4 * 13 + 5
    rr, cc = np.divmod(np.arange(9 * 13), 13)
    man = (np.abs(rr - 4) + np.abs(cc - 5)).astype(np.int32)
    assert np.array_equal(oracle.dijkstra(row, col, w, src), man)
    assert np.array_equal(oracle.nearfar(row, col, w, src, 2)[0], man)
    assert np.array_equal(oracle.bfs(row, col, src)[0], man)
    # weight-7 grid: 7 x Manhattan
    row, col, w = grid_csr(6, 6, 7)
    rr, cc = np.divmod(np.arange(36), 6)
    assert np.array_equal(oracle.nearfar(row, col, w, 0, 5)[0], (7 * (rr + cc)).astype(np.int32))
    # directed path 0->1->...->n-1 with weights 1..n-1: prefix sums; reverse direction unreachable
    n = 50
    row = np.arange(n + 1, dtype=np.uint32).clip(max=n - 1)
    col = np.arange(1, n, dtype=np.uint32)
    ws = np.arange(1, n, dtype=np.int32)
    pref = np.concatenate([[0], np.cumsum(ws)]).astype(np.int32)
    assert np.array_equal(oracle.dijkstra(row, col, ws, 0), pref)
    assert np.array_equal(oracle.nearfar(row, col, ws, 0, 10)[0], pref)
    d = oracle.dijkstra(row, col, ws, 10)
    assert (d[:10] == INF).all() and np.array_equal(d[10:], pref[10:] - pref[10])
    # star: hub 0 to every leaf with weight i, leaves have no out-edges
    n = 30
    row = np.concatenate([[0], np.full(n, n - 1)]).astype(np.uint32)
    col = np.arange(1, n, dtype=np.uint32)
    ws = np.arange(1, n, dtype=np.int32)
    assert np.array_equal(oracle.nearfar(row, col, ws, 0, 4)[0], np.arange(n, dtype=np.int32))
    assert (oracle.dijkstra(row, col, ws, 5) == np.where(np.arange(n) == 5, 0, INF)).all()
    # zero weights: every reachable vertex at 0
    row, col, w = grid_csr(5, 5, 0)
    assert (oracle.nearfar(row, col, w, 12, 1)[0] == 0).all()


def test_brute_force_tiny():
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(0, 40))
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m)
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        row = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]).astype(np.uint32)
        col = dst.astype(np.uint32)
        w = rng.integers(0, 9, m).astype(np.int32)
        s = int(rng.integers(0, n))
        bf = bellman_ford(row, col, w, s)
        assert np.array_equal(oracle.dijkstra(row, col, w, s), bf)
        for delta in (1, 4, 100):
            assert np.array_equal(oracle.nearfar(row, col, w, s, delta)[0], bf)
        assert np.array_equal(oracle.bfs(row, col, s)[0], bellman_ford(row, col, w, s, unit=True))


def test_certificate(small_graphs):
    g = small_graphs["medium"]
    d = g["dist"]
    assert oracle.certificate(g["row"], g["col"], g["w"], g["src"], d) == 0
    assert oracle.certificate(g["row"], g["col"], None, g["src"], g["bfs"]) == 0
    reached = np.flatnonzero((d != INF) & (d > 0))
    for delta in (+1, -1):
        bad = d.copy()
        bad[reached[len(reached) // 2]] += delta
        assert oracle.certificate(g["row"], g["col"], g["w"], g["src"], bad) > 0
    bad = d.copy()
    bad[reached[0]] = INF
    assert oracle.certificate(g["row"], g["col"], g["w"], g["src"], bad) > 0


def test_reach(small_graphs):
    g = small_graphs["twoparts"]
    vr, er = oracle.reach(g["row"], g["dist"])
    fin = g["dist"] != INF
    assert vr == fin.sum()
    assert er == np.diff(g["row"].astype(np.int64))[fin].sum()


@pytest.mark.parametrize("bug", [1, 2, 3, 4])
def test_pins_catch_planted_bugs(small_graphs, bug):
    oracle.plant_bug(bug)
    try:
        failures = set()
        for name, g in small_graphs.items():
            failures.update(check_all(g, g["src"]))
        row, col, w = grid_csr(7, 7)
        if not np.array_equal(oracle.bfs(row, col, 0)[0], oracle.dijkstra(row, col, w, 0)):
            failures.add("bfs-vs-grid")
        assert failures, f"planted bug {bug} was not caught by any pin"
    finally:
        oracle.plant_bug(0)


# ---- connected components (apps/cc) ------------------------------------------

def min_id_labels(n, row, col):
    """Weak components from scipy (an independent library), relabelled by the
    smallest vertex id of each component."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    m = csr_matrix((np.ones(len(col)), np.asarray(col, np.int64), np.asarray(row, np.int64)), shape=(n, n))
    k, lab = connected_components(m, directed=True, connection="weak")
    first = np.full(k, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    return first[lab].astype(np.int32), k


def random_csr(n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, dtype=np.uint32)
    row[1:] = np.cumsum(np.bincount(src, minlength=n))
    return row, dst[order].astype(np.uint32)


def test_cc_matches_scipy(small_graphs):
    for name, g in small_graphs.items():
        lab, k = oracle.cc(g["row"], g["col"])
        ref, rk = min_id_labels(len(g["row"]) - 1, g["row"], g["col"])
        assert np.array_equal(lab, ref) and k == rk, name
    for seed, (n, m) in enumerate([(50, 30), (300, 200), (2000, 1500), (5000, 9000)]):
        row, col = random_csr(n, m, seed)
        lab, k = oracle.cc(row, col)
        ref, rk = min_id_labels(n, row, col)
        assert np.array_equal(lab, ref) and k == rk, (n, m)


def test_cc_closed_forms():
    row, col, _ = grid_csr(9, 13)
    lab, k = oracle.cc(row, col)
    assert k == 1 and not lab.any()
    # edges only from larger to smaller ids still join everything (directions ignored)
    n = 40
    col = np.arange(0, n - 1, dtype=np.uint32)  # vertex i+1 -> i
    row = np.concatenate([[0, 0], np.arange(1, n, dtype=np.uint32)]).astype(np.uint32)
    lab, k = oracle.cc(row, col)
    assert k == 1 and not lab.any()
    # isolated vertices and two paths: labels are the first vertex of each part
    row = np.array([0, 1, 2, 2, 2, 3, 4, 4], np.uint32)  # 0->1->2, 3 alone, 4->5->6
    col = np.array([1, 2, 5, 6], np.uint32)
    lab, k = oracle.cc(row, col)
    assert k == 3 and lab.tolist() == [0, 0, 0, 3, 4, 4, 4]


def test_cc_pins_catch_planted_bug():
    oracle.plant_bug(4)
    try:
        caught = False
        for seed, (n, m) in enumerate([(300, 200), (2000, 1500)]):
            row, col = random_csr(n, m, seed)
            lab, k = oracle.cc(row, col)
            ref, rk = min_id_labels(n, row, col)
            caught = caught or not np.array_equal(lab, ref)
        assert caught, "skipping the last edge of each row was not caught"
    finally:
        oracle.plant_bug(0)


# ---- PageRank (apps/pr) --------------------------------------------------------

def pagerank_solve(n, row, col, alpha=0.85):
    """The same fixpoint by a dense linear solve (numpy), an independent method:
    (I - alpha M^T) p = (1 - alpha) 1 with M[u, v] = multiplicity(u -> v) / outdeg(u)."""
    row = np.asarray(row, np.int64)
    M = np.zeros((n, n))
    for u in range(n):
        od = row[u + 1] - row[u]
        for e in range(row[u], row[u + 1]):
            M[u, col[e]] += 1.0 / od
    return np.linalg.solve(np.eye(n) - alpha * M.T, np.full(n, 1.0 - alpha))


def test_pagerank_closed_forms():
    n = 17
    cyc_row = np.arange(n + 1, dtype=np.uint32)
    cyc_col = ((np.arange(n) + 1) % n).astype(np.uint32)
    p, _ = oracle.pagerank(cyc_row, cyc_col)
    assert np.allclose(p, 1.0, rtol=0, atol=1e-12)  # p = 0.15 + 0.85 p
    p, _ = oracle.pagerank(np.zeros(6, np.uint32), np.zeros(0, np.uint32))
    assert np.allclose(p, 0.15, rtol=0, atol=1e-15)
    star_row = np.array([0, n - 1] + [n - 1] * (n - 1), np.uint32)  # 0 -> 1..n-1
    star_col = np.arange(1, n, dtype=np.uint32)
    p, _ = oracle.pagerank(star_row, star_col)
    assert abs(p[0] - 0.15) < 1e-15 and np.allclose(p[1:], 0.15 + 0.85 * 0.15 / (n - 1), rtol=0, atol=1e-14)


def test_pagerank_matches_linear_solve():
    for seed, (n, m) in enumerate([(30, 60), (200, 700), (400, 300)]):  # duplicates, dangling vertices
        row, col = random_csr(n, m, seed)
        p, _ = oracle.pagerank(row, col)
        assert np.allclose(p, pagerank_solve(n, row, col), rtol=1e-10, atol=0), (n, m)


def test_pagerank_pins_catch_planted_bug():
    oracle.plant_bug(5)
    try:
        row, col = random_csr(200, 700, 1)
        p, _ = oracle.pagerank(row, col)
        assert not np.allclose(p, pagerank_solve(200, row, col), rtol=1e-6, atol=0)
    finally:
        oracle.plant_bug(0)
