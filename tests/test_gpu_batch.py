"""GPU parity of the batched (multi-source) path and the learned-scheduler kernels.

Every column of a batched run must equal the oracle's distances bit for bit,
whatever delta width, far-pile mode or forced per-phase schedule is used.
"""
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError, Options
from galoisgpu_amd import sched as lsched
from galoisgpu_amd.batch import last_features, sssp_batch, sssp_multi

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def oracle_cols(g, srcs):
    return np.stack([oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, int(s)) for s in srcs], axis=1)


def test_multi_small_golden(small_graphs):
    for name, gg in small_graphs.items():
        g = CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"])
        n = g.nnodes
        srcs = [gg["src"]] + [int(x) for x in np.random.default_rng(1).integers(0, n, 7)]
        ref = oracle_cols(g, srcs)
        assert np.array_equal(ref[:, 0], gg["dist"])
        dev = g.copy_to_gpu()
        for delta in (1, 5, 32):
            for far in (0, 1, 2):
                say("multi small", name, delta, far)
                d, _ = sssp_multi(dev, srcs, Options(delta=delta, far_mode=far))
                assert np.array_equal(d, ref), (name, delta, far)
        dev.close()


def degree_boundary_graph():
    """Vertices whose out-degrees sit on both sides of every boundary of the
    batched kernel's expansion paths: the 8-edge relax window (7, 8, 9, 15,
    16, 17), and chunk sizes heavy_deg = 8, 16, 17 and 64 (a chunk of one edge,
    a chunk shorter than a window).  The last vertex has 37 edges, so its last
    window ends at the last edge of the array.  Weights include zeros."""
    rng = np.random.default_rng(11)
    V = 3000
    deg = rng.integers(0, 6, V)
    deg[:50] = np.maximum(deg[:50], 2)
    special = [1, 7, 8, 9, 15, 16, 17, 18, 31, 33, 63, 64, 65, 127, 128, 129, 136, 255, 257, 1025]
    hubs = rng.choice(np.arange(1, V - 1), len(special), replace=False)
    deg[hubs] = special
    deg[V - 1] = 37
    row = np.zeros(V + 1, dtype=np.int64)
    row[1:] = np.cumsum(deg)
    E = int(row[-1])
    col = rng.integers(0, V, E)
    # every hub is reached from the low vertices the sources are drawn from
    for k, h in enumerate(hubs):
        if deg[k % 50] > 0:
            col[row[k % 50]] = h
    w = rng.integers(0, 100, E)
    w[rng.random(E) < 0.6] += 1
    return CSRGraph.from_arrays(row, col, w), hubs


def test_multi_degree_boundaries():
    g, hubs = degree_boundary_graph()
    srcs = np.concatenate([np.arange(0, 50, 3), hubs[:8]]).astype(np.uint32)
    srcs = np.resize(srcs, 64)
    ref = oracle_cols(g, srcs)
    dev = g.copy_to_gpu()
    for hd in (8, 16, 17, 64, 0):
        for S in (1, 7, 64):
            for far, delta in ((0, 1 << 30), (2, 40), (0, 13)):
                o = Options(delta=delta, far_mode=far, heavy_deg=hd, block_size=256)
                say("degree boundaries", hd, S, far, delta)
                d, _ = sssp_multi(dev, srcs[:S], o)
                assert np.array_equal(d, ref[:, :S]), (hd, S, far, delta)
    with pytest.raises(LsgError, match="heavy_deg"):
        sssp_multi(dev, srcs[:4], Options(delta=32, heavy_deg=7))
    dev.close()


@pytest.fixture(scope="module")
def rmat16():
    g = CSRGraph.rmat(16, 16, seed=3)
    srcs = np.random.default_rng(5).integers(0, g.nnodes, 64)
    srcs[10] = srcs[3]  # a duplicated source
    srcs[0] = g.max_degree_vertex()
    return g, srcs, oracle_cols(g, srcs)


@pytest.mark.parametrize("opts", [dict(), dict(far_mode=1), dict(far_mode=2), dict(far_mode=2, delta=1),
                                  dict(block_size=256, heavy_deg=256), dict(block_size=1024, delta=7),
                                  dict(delta=10000)])
def test_multi_rmat16(rmat16, opts):
    g, srcs, ref = rmat16
    dev = g.copy_to_gpu()
    o = dict(delta=32)
    o.update(opts)
    say("multi rmat16", o)
    d, st = sssp_multi(dev, srcs, Options(**o))
    assert np.array_equal(d, ref), o
    assert st["launches"] == 1 and st["rounds"] > 0
    dev.close()


def test_multi_partial_width_and_batch(rmat16):
    g, srcs, ref = rmat16
    dev = g.copy_to_gpu()
    for S in (1, 5, 32, 63):
        say("multi width", S)
        d, _ = sssp_multi(dev, srcs[:S], Options(delta=32))
        assert np.array_equal(d, ref[:, :S]), S
    many = np.concatenate([srcs, srcs[:37]])
    say("batch 101 sources")
    outs, st = sssp_batch(dev, many, Options(delta=32))
    assert st["launches"] == 2
    for k, s in enumerate(many):
        assert np.array_equal(outs[k], ref[:, k % 64]), k
    dev.close()


def test_multi_sources_without_out_edges(rmat16):
    """Sources without out-edges take no lane (the others run in a compact matrix that is
    then expanded): every column must still be the oracle's, for a mix of such sources and
    others, for only such sources, and with one repeated."""
    g, srcs, _ = rmat16
    deg = np.diff(g.row_start.astype(np.int64))
    dead = np.flatnonzero(deg == 0)[:20]
    live = np.flatnonzero(deg > 0)[:30]
    assert len(dead) == 20 and len(live) == 30
    mixed = np.empty(50, dtype=np.int64)
    mixed[0::5] = dead[:10]
    mixed[np.arange(50) % 5 != 0] = np.concatenate([live, dead[10:20]])
    mixed[7] = mixed[0]  # a repeated source without out-edges
    dev = g.copy_to_gpu()
    for name, ss in (("mixed", mixed), ("only without out-edges", dead[:9]), ("one", dead[:1]),
                     ("64 mixed", np.concatenate([mixed, live[:14]]))):
        say("multi dead sources", name, len(ss))
        d, st = sssp_multi(dev, ss, Options(delta=32))
        assert np.array_equal(d, oracle_cols(g, ss)), name
    dev.close()


def test_partial_width_rows_up_to_the_last_vertex():
    """Fewer than 64 sources on a graph whose last vertex is reached: rows of
    D are S words, and relaxing edges into the last row exercises the row
    loads at the end of D (the grid's corner vertex is every source's
    destination)."""
    g = CSRGraph.grid(96, 80, 1, 100, seed=8)
    V = g.nnodes
    srcs = np.random.default_rng(8).integers(0, V, 64)
    srcs[0] = V - 1
    ref = np.stack([oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, int(s)) for s in srcs], axis=1)
    dev = g.copy_to_gpu()
    for S in (1, 7, 32, 64):
        for far in (0, 2):
            say("width to last vertex", S, far)
            d, _ = sssp_multi(dev, srcs[:S], Options(delta=64, far_mode=far))
            assert np.array_equal(d, ref[:, :S]), (S, far)
    dev.close()


def test_forced_schedules_keep_distances(rmat16):
    """Distances are identical whatever schedule the scheduler imposes."""
    g, srcs, ref = rmat16
    dev = g.copy_to_gpu()
    base = lsched.new_model(0)
    rng = np.random.default_rng(9)
    schedules = [[a] for a in range(lsched.N_ACT)] + [list(rng.integers(0, lsched.N_ACT, 40)) for _ in range(4)]
    for sch in schedules:
        lsched.set_model(lsched.forced(base, sch))
        say("forced", sch[:6])
        d, st = sssp_multi(dev, srcs, Options(delta=32, sched=1, trace=True))
        assert np.array_equal(d, ref), sch
        feats = last_features()
        assert len(feats) == st["phases"] - 1
        acts = feats[:, 16].astype(int)
        assert all(acts[p] == sch[min(p, len(sch) - 1)] for p in range(len(acts)))
    # an untrained (random) network as the scheduler: still exact
    lsched.set_model(base)
    d, _ = sssp_multi(dev, srcs, Options(delta=32, sched=1))
    assert np.array_equal(d, ref)
    dev.close()


def numpy_forward(m, X):
    o = lsched
    x = (X - m[o.OFF_MEAN:o.OFF_MEAN + 16]) * m[o.OFF_INVSTD:o.OFF_INVSTD + 16]
    W1 = m[o.OFF_W1:o.OFF_B1].reshape(16, 64)
    W2 = m[o.OFF_W2:o.OFF_B2].reshape(64, 64)
    W3 = m[o.OFF_W3:o.OFF_B3].reshape(64, 8)
    h1 = np.maximum(x @ W1 + m[o.OFF_B1:o.OFF_B1 + 64], 0)
    h2 = np.maximum(h1 @ W2 + m[o.OFF_B2:o.OFF_B2 + 64], 0)
    return h2 @ W3 + m[o.OFF_B3:o.OFF_B3 + 8]


def test_mlp_train_and_infer_kernels():
    """The HIP trainer learns a separable rule; HIP inference matches an fp32 numpy forward."""
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (3000, 16)).astype(np.float32) * 3 + 1
    y = ((X[:, 0] > 1).astype(int) + 2 * (X[:, 3] > 1).astype(int) + 4 * (X[:, 7] > 1).astype(int)).astype(np.int32)
    say("mlp train")
    m, loss = lsched.train(X, y, steps=3000, lr=0.05)
    pred = lsched.infer(m, X)
    acc = float((pred == y).mean())
    assert acc > 0.95 and loss < 0.3, (acc, loss)
    logits = numpy_forward(m.astype(np.float64), X.astype(np.float64))
    top2 = np.sort(logits, axis=1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) > 1e-3  # skip near-ties, where fp32 order may differ
    assert np.array_equal(pred[clear], np.argmax(logits, axis=1)[clear])


def test_many_graphs_one_launch():
    """Config 4B in small: graphs of uneven sizes, one source each, one persistent launch."""
    from galoisgpu_amd.batch import sssp_graphs
    rng = np.random.default_rng(21)
    graphs, devs, srcs = [], [], []
    for i in range(40):
        g = CSRGraph.rmat(int(rng.integers(6, 13)), 8, seed=100 + i)
        graphs.append(g)
        devs.append(g.copy_to_gpu())
        srcs.append(g.max_degree_vertex() if i % 3 else int(rng.integers(0, g.nnodes)))
    tiny = CSRGraph.from_arrays([0, 0], [], [])  # a one-node graph among them
    graphs.append(tiny)
    devs.append(tiny.copy_to_gpu())
    srcs.append(0)
    for far, graph_launch in ((0, False), (1, False), (0, True)):
        say("many graphs", far, graph_launch)
        outs, st = sssp_graphs(devs, srcs, Options(delta=32, far_mode=far, graph_launch=graph_launch))
        assert st["launches"] == 1
        for g, s, d in zip(graphs, srcs, outs):
            assert np.array_equal(d, oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, s))
    for d in devs:
        d.close()


def test_graph_launch_replays_match_eager(rmat16):
    """Captured hipGraph launches replayed several times, with eager work and
    synchronisation in between, each equal to the eager result."""
    from galoisgpu_amd.apps import sssp_into
    from galoisgpu_amd.batch import DeviceArray, sssp_multi_into
    g, srcs, ref = rmat16
    dev = g.copy_to_gpu()
    D = DeviceArray(g.nnodes * 64 * 4)
    single = DeviceArray(g.nnodes * 4)
    for rep in range(4):
        say("graph replay", rep)
        sssp_multi_into(dev, srcs, Options(delta=32, graph_launch=True), D)
        assert np.array_equal(D.to_numpy(np.int32, (g.nnodes, 64)), ref), rep
        # eager work between replays: a plain single-source run
        sssp_into(dev, single.ptr, Options(delta=32, start_node=int(srcs[rep])))
        assert np.array_equal(single.to_numpy(np.int32, g.nnodes), ref[:, rep])
        sssp_into(dev, single.ptr, Options(delta=32, start_node=int(srcs[rep + 1]), graph_launch=True))
        assert np.array_equal(single.to_numpy(np.int32, g.nnodes), ref[:, rep + 1])
    D.close()
    single.close()
    dev.close()
