"""GPU parity of the one-workgroup near/far traversal (lsg_opts.single_block).

Its piles live in LDS and carry stale entries that are skipped, so its item
counts differ from the persistent kernel's; distances must equal the oracle's
Dijkstra / BFS bit for bit: on the golden graphs, the 256x256 grid of
configuration 1 at delta 32 and at the tuned delta, road-like and RMAT graphs
(wavefront-expanded heavy vertices), a graph whose far pile passes half full
(stale far entries dropped at the split) and whose rounds exceed one pass of
the workgroup, and graphs whose piles outgrow LDS early or late (the
persistent kernel then continues from the distances reached, still exact).
"""
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, Options
from galoisgpu_amd.apps import bfs_into, sssp_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def run(g, src, bfs=False, **kw):
    dev = g.copy_to_gpu()
    buf = DeviceArray(g.nnodes * 4)
    try:
        o = Options(start_node=src, **{"single_block": True, **kw})
        st = (bfs_into if bfs else sssp_into)(dev, buf.ptr, o)
        return buf.to_numpy(np.int32, g.nnodes), st
    finally:
        buf.close()
        dev.close()


def test_single_block_golden(small_graphs):
    for name, gg in small_graphs.items():
        g = CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"])
        say("single block golden", name)
        d, st = run(g, gg["src"], delta=32)
        assert np.array_equal(d, gg["dist"]), name
        assert st["grid_blocks"] == 1
        b, _ = run(g, gg["src"], bfs=True)
        assert np.array_equal(b, gg["bfs"]), name


@pytest.mark.parametrize("delta", [32, 256, 512])
def test_single_block_grid256(grid256_golden, delta):
    g = CSRGraph.grid(256, 256, 1, 100, 1)
    say("single block grid256 delta", delta)
    d, st = run(g, 0, delta=delta)
    assert np.array_equal(d, grid256_golden["dist"])
    assert st["grid_blocks"] == 1
    assert st["items"] >= g.nnodes  # every vertex expanded at least once
    b, bst = run(g, 0, bfs=True)
    assert np.array_equal(b, grid256_golden["bfs"])
    assert bst["near_iters"] == int(grid256_golden["bfs"].max()) + 1  # one round per level


def test_single_block_far_prune_and_wide_rounds():
    """Source 0 reaches 1500 leaves directly (weights 1000-1099) and through
    vertex 1 (weights 1 + 1).  At delta 1 the far pile first holds the direct
    entries, then also the lower ones: at the next split it is past half full
    and the direct entries, now stale, are dropped.  The leaves' round has 1500
    entries, more than one pass of the workgroup; 0 and 1 are wavefront-expanded."""
    n_leaf = 1500
    leaves = np.arange(2, 2 + n_leaf)
    row = np.zeros(n_leaf + 3, dtype=np.int64)
    row[1] = 1 + n_leaf
    row[2:] = 1 + 2 * n_leaf
    dst = np.concatenate([[1], leaves, leaves])
    w = np.concatenate([[1], 1000 + np.arange(n_leaf) % 100, np.ones(n_leaf, dtype=np.int64)])
    g = CSRGraph.from_arrays(row, dst, w)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    assert (ref[2:] == 2).all()
    say("single block far prune")
    d, st = run(g, 0, delta=1)
    assert np.array_equal(d, ref)
    assert st["grid_blocks"] == 1


@pytest.mark.parametrize("rows,seed", [(256, 3), (512, 4)])
def test_single_block_roadlike(rows, seed):
    g = CSRGraph.roadlike(rows, rows, 0.10, 0.001, 1, 10000, seed=seed)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    say("single block road", rows)
    d, _ = run(g, 0, delta=10000)
    assert np.array_equal(d, ref)
    lref, _ = oracle.bfs(g.row_start, g.edge_dst, 0)
    b, _ = run(g, 0, bfs=True)
    assert np.array_equal(b, lref)


def test_single_block_rmat_and_zero_weights():
    g = CSRGraph.rmat(14, 8, seed=9)
    src = g.max_degree_vertex()
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
    say("single block rmat14")
    d, _ = run(g, src, delta=32)
    assert np.array_equal(d, ref)
    lref, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    b, _ = run(g, src, bfs=True)
    assert np.array_equal(b, lref)
    # zero weights: every reachable vertex at distance 0
    z = CSRGraph.from_arrays(g.row_start, g.edge_dst, np.zeros(g.nedges, dtype=np.int64))
    d0, _ = run(z, src, delta=32)
    zref = oracle.dijkstra(z.row_start, z.edge_dst, z.edge_data, src)
    assert np.array_equal(d0, zref)


def test_graphs_one_workgroup_each():
    """lsg_sssp_graphs with single_block: one workgroup per graph, piles in HBM.  RMAT graphs of
    uneven sizes (wavefront-expanded hubs), a 64x64 grid (many rounds), road-like graphs, a
    one-node graph; each graph against Dijkstra, at several deltas."""
    from galoisgpu_amd.batch import sssp_graphs
    rng = np.random.default_rng(23)
    graphs, srcs = [], []
    for i in range(40):
        g = CSRGraph.rmat(int(rng.integers(6, 15)), 8, seed=200 + i)
        graphs.append(g)
        srcs.append(g.max_degree_vertex() if i % 3 else int(rng.integers(0, g.nnodes)))
    graphs.append(CSRGraph.grid(64, 64, 1, 100, seed=5))
    srcs.append(0)
    for k in range(3):
        graphs.append(CSRGraph.roadlike(96, 96, 0.10, 0.001, 1, 10000, seed=30 + k))
        srcs.append(k)
    graphs.append(CSRGraph.from_arrays([0, 0], [], []))
    srcs.append(0)
    refs = [oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, s) for g, s in zip(graphs, srcs)]
    devs = [g.copy_to_gpu() for g in graphs]
    try:
        for delta in (1, 32, 1 << 20):
            say("graphs one workgroup each, delta", delta)
            outs, st = sssp_graphs(devs, srcs, Options(delta=delta, single_block=True))
            assert st["grid_blocks"] == len(graphs) and st["block_size"] == 1024  # no fallback
            bad = [k for k in range(len(graphs)) if not np.array_equal(outs[k], refs[k])]
            assert not bad, (delta, bad[:8])
    finally:
        for d in devs:
            d.close()


def test_graphs_one_workgroup_each_falls_back_on_overflow(monkeypatch):
    """Piles capped at 64 entries (LSG_GRAPHS_PILE_CAP): the RMAT graphs' rounds outgrow them, the
    call runs the union launch instead (one persistent grid) and every graph is still exact."""
    from galoisgpu_amd.batch import sssp_graphs
    graphs = [CSRGraph.rmat(10 + (i % 3), 8, seed=300 + i) for i in range(12)]
    srcs = [g.max_degree_vertex() for g in graphs]
    refs = [oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, s) for g, s in zip(graphs, srcs)]
    devs = [g.copy_to_gpu() for g in graphs]
    try:
        monkeypatch.setenv("LSG_GRAPHS_PILE_CAP", "64")
        say("graphs fallback")
        outs, st = sssp_graphs(devs, srcs, Options(delta=32, single_block=True))
        assert st["block_size"] != 1024 or st["grid_blocks"] != len(graphs)  # the union launch ran
        for k in range(len(graphs)):
            assert np.array_equal(outs[k], refs[k]), k
        monkeypatch.delenv("LSG_GRAPHS_PILE_CAP")
        outs, st = sssp_graphs(devs, srcs, Options(delta=32, single_block=True))
        assert st["grid_blocks"] == len(graphs) and st["block_size"] == 1024
        for k in range(len(graphs)):
            assert np.array_equal(outs[k], refs[k]), k
    finally:
        for d in devs:
            d.close()


def test_single_block_seeded_sweep():
    """24 seeded graphs of the three families at random sizes, sources and deltas (weights from 0,
    delta from 1): every single-source run, and the batch of all of them, equal Dijkstra."""
    from galoisgpu_amd.batch import sssp_graphs
    rng = np.random.default_rng(77)
    graphs, srcs, deltas = [], [], []
    for i in range(24):
        kind = i % 3
        if kind == 0:
            g = CSRGraph.grid(int(rng.integers(2, 90)), int(rng.integers(2, 90)), 0, int(rng.integers(1, 300)),
                              seed=400 + i)
        elif kind == 1:
            n = int(rng.integers(8, 120))
            g = CSRGraph.roadlike(n, n, 0.15, 0.002, 1, int(rng.integers(2, 20000)), seed=400 + i)
        else:
            g = CSRGraph.rmat(int(rng.integers(4, 13)), int(rng.integers(1, 12)), seed=400 + i)
        graphs.append(g)
        srcs.append(int(rng.integers(0, g.nnodes)))
        deltas.append(int(rng.choice([1, 7, 32, 300, 5000, 1 << 24])))
    refs = [oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, s) for g, s in zip(graphs, srcs)]
    say("single block seeded sweep")
    for k, (g, s, dl) in enumerate(zip(graphs, srcs, deltas)):
        d, _ = run(g, s, delta=dl)
        assert np.array_equal(d, refs[k]), (k, g.nnodes, dl)
    devs = [g.copy_to_gpu() for g in graphs]
    try:
        outs, st = sssp_graphs(devs, srcs, Options(delta=32, single_block=True))
        for k in range(len(graphs)):
            assert np.array_equal(outs[k], refs[k]), k
    finally:
        for d in devs:
            d.close()


def test_single_block_falls_back_when_the_piles_outgrow_lds():
    """A complete 4-ary tree of depth 8: its deepest levels are wider than the
    LDS piles, so the call runs the persistent kernel instead and is still exact."""
    n = sum(4 ** k for k in range(9))  # 87381 vertices
    deg = np.zeros(n, dtype=np.int64)
    internal = (n - 4 ** 8)
    deg[:internal] = 4
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(deg)
    dst = np.arange(1, n)  # children of v are 4v+1 .. 4v+4
    w = np.random.default_rng(2).integers(1, 100, n - 1)
    g = CSRGraph.from_arrays(row, dst, w)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    say("single block tree fallback")
    d, st = run(g, 0, delta=32)
    assert np.array_equal(d, ref)
    assert st["grid_blocks"] > 1  # the persistent kernel ran
    lref, _ = oracle.bfs(g.row_start, g.edge_dst, 0)
    b, bst = run(g, 0, bfs=True)
    assert np.array_equal(b, lref)
    assert bst["grid_blocks"] > 1


def test_single_block_overflow_continues_in_the_persistent_kernel():
    """Graphs whose piles outgrow LDS after many rounds or after a few: the persistent kernel
    continues from the distances the workgroup left (every reached vertex queued again), exact;
    the rounds of both launches are counted.  Grids at wide deltas, road-like graphs at their
    delta, RMAT from random sources, BFS on each, zero weights and delta 1 among them."""
    rng = np.random.default_rng(91)
    cases = [(CSRGraph.grid(512, 512, 1, 100, seed=2), 0, 256),
             (CSRGraph.roadlike(256, 256, 0.10, 0.001, 1, 10000, seed=3), 0, 10000)]
    for i in range(10):
        kind = i % 3
        if kind == 0:
            g = CSRGraph.grid(int(rng.integers(200, 400)), int(rng.integers(200, 400)), int(rng.integers(0, 2)),
                              100, seed=500 + i)
            dl = int(rng.choice([512, 2048, 1 << 20]))
        elif kind == 1:
            n = int(rng.integers(150, 300))
            g = CSRGraph.roadlike(n, n, 0.10, 0.001, 1, 10000, seed=500 + i)
            dl = int(rng.choice([1, 5000, 20000]))
        else:
            g = CSRGraph.rmat(int(rng.integers(12, 16)), 8, seed=500 + i)
            dl = int(rng.choice([1, 32, 1 << 20]))
        cases.append((g, int(rng.integers(0, g.nnodes)), dl))
    fell_back = 0
    for k, (g, src, dl) in enumerate(cases):
        ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
        say("single block overflow case", k, g.nnodes, src, dl)
        d, st = run(g, src, delta=dl)
        assert np.array_equal(d, ref), (k, g.nnodes, src, dl)
        fell_back += st["grid_blocks"] > 1
        if st["grid_blocks"] > 1:
            assert st["near_iters"] >= 2
        lref, _ = oracle.bfs(g.row_start, g.edge_dst, src)
        b, _ = run(g, src, bfs=True)
        assert np.array_equal(b, lref), (k, "bfs")
    assert fell_back >= len(cases) // 2, fell_back


def test_single_block_trace_holds_the_persistent_rounds_only():
    """A traced one-workgroup call leaves no round trace when it completes in the workgroup (not
    the previous call's), and the persistent kernel's rounds when it continued there."""
    from galoisgpu_amd.apps import last_trace
    g = CSRGraph.grid(64, 64, 1, 100, seed=5)
    say("single block trace")
    run(g, 0, delta=32, single_block=False, trace=True)  # a persistent-kernel trace first
    assert len(last_trace()[0]) > 0
    d, st = run(g, 0, delta=32, trace=True)
    assert st["grid_blocks"] == 1 and len(last_trace()[0]) == 0
    assert np.array_equal(d, oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0))
    r = CSRGraph.rmat(14, 8, seed=9)
    src = r.max_degree_vertex()
    d, st = run(r, src, delta=32, trace=True)
    assert st["grid_blocks"] > 1 and len(last_trace()[0]) > 0
    assert np.array_equal(d, oracle.dijkstra(r.row_start, r.edge_dst, r.edge_data, src))


def test_single_block_overflow_continues_under_learned_schedules():
    """After an overflow the persistent kernel may run the learned scheduler (single_block with
    sched): under every forced per-phase action, random schedules and an untrained network, the
    distances equal the oracle's."""
    from galoisgpu_amd import sched as lsched
    base = lsched.new_model(0, kind="plain")
    rng = np.random.default_rng(5)
    schedules = [[a] for a in range(lsched.N_ACT)] + [list(rng.integers(0, lsched.N_ACT, 300)) for _ in range(2)]
    cases = [(CSRGraph.roadlike(256, 256, 0.10, 0.001, 1, 10000, seed=3), 0, 10000),
             (CSRGraph.rmat(14, 8, seed=9), None, 32)]
    try:
        for g, src, delta in cases:
            src = g.max_degree_vertex() if src is None else src
            ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
            for sch in schedules:
                lsched.set_model(lsched.forced(base, sch), kind="plain")
                say("single block then learned", g.nnodes, sch[:6])
                d, st = run(g, src, delta=delta, sched=1)
                assert np.array_equal(d, ref), sch
                assert st["grid_blocks"] > 1  # the workgroup overflowed and the persistent kernel went on
            lsched.set_model(base, kind="plain")
            d, _ = run(g, src, delta=delta, sched=1)
            assert np.array_equal(d, ref)
    finally:
        lsched.set_model(base, kind="plain")
