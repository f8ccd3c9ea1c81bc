"""GPU parity of the -l path (host-driven rounds with the edge-balanced pass,
apps/sssp/kernel.cu:78-176,318-354, apps/bfs/kernel.cu:33-120,286-311) and of
the single-source split and bucket far piles, against the oracle.

The reference's DEGREE_LIMIT is the sweep's thread count (2^19 on 256 CUs),
which no vertex of these graphs reaches, so the tests also lower the
threshold (lb_deg) until most frontier vertices take the edge-balanced pass.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import ROOT
from galoisgpu_amd import CSRGraph, LsgError, Options
from galoisgpu_amd.apps import bfs_into, sssp_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def run(dev, n, fn, **kw):
    buf = DeviceArray(n * 4)
    st = fn(dev, buf.ptr, Options(**kw))
    d = buf.to_numpy(np.int32, n)
    buf.close()
    return d, st


def test_lb_small_golden(small_graphs):
    for name, gg in small_graphs.items():
        g = CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"])
        dev = g.copy_to_gpu()
        for lb_deg in (1, 2, 5, 0):
            for delta in (1, 32, 10000):
                say("lb small", name, lb_deg, delta)
                d, _ = run(dev, g.nnodes, sssp_into, delta=delta, start_node=gg["src"], enable_lb=1, lb_deg=lb_deg)
                assert np.array_equal(d, gg["dist"]), (name, lb_deg, delta)
            d, _ = run(dev, g.nnodes, bfs_into, start_node=gg["src"], enable_lb=1, lb_deg=lb_deg)
            assert np.array_equal(d, gg["bfs"]), (name, lb_deg)
        dev.close()


def test_lb_degree_boundaries():
    """Thresholds on both sides of hub degrees and block boundaries of the flattened pass
    (a block of 256 edges spanning several vertices, vertices spanning several blocks),
    with zero weights, under both sweep expansion modes and the split / bucket far piles."""
    from test_gpu_batch import degree_boundary_graph
    g, hubs = degree_boundary_graph()
    dev = g.copy_to_gpu()
    for src in (0, int(hubs[0]), int(hubs[5])):
        ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
        lv, _ = oracle.bfs(g.row_start, g.edge_dst, src)
        for lb_deg in (1, 8, 64, 128, 129, 256, 1025, 1026):
            for kw in (dict(), dict(expand_mode=0), dict(far_mode=1), dict(far_mode=2, delta=7)):
                o = dict(delta=40, start_node=src, enable_lb=1, lb_deg=lb_deg)
                o.update(kw)
                say("lb boundaries", src, lb_deg, kw)
                d, st = run(dev, g.nnodes, sssp_into, **o)
                assert np.array_equal(d, ref), (src, lb_deg, kw)
            d, _ = run(dev, g.nnodes, bfs_into, start_node=src, enable_lb=1, lb_deg=lb_deg)
            assert np.array_equal(d, lv), (src, lb_deg)
    dev.close()


@pytest.fixture(scope="module")
def rmat20():
    g = CSRGraph.rmat(20, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    dev = g.copy_to_gpu()
    yield g, dev
    dev.close()


def test_lb_rmat20_config2(rmat20):
    """C2 with -l: SSSP (delta 32) and BFS from the max-degree vertex.  With the reference's
    threshold no vertex is heavy (the sweep does everything); lower thresholds send the hubs,
    then most of the frontier, through the edge-balanced pass."""
    g, dev = rmat20
    src = g.max_degree_vertex()
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
    lv, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    for lb_deg in (0, 4096, 256, 16):
        say("lb rmat20 sssp", lb_deg)
        d, st = run(dev, g.nnodes, sssp_into, delta=32, start_node=src, enable_lb=1, lb_deg=lb_deg)
        assert np.array_equal(d, ref), lb_deg
        if lb_deg == 0:
            assert st["heavy_chunks"] == 0  # no edge went through the edge-balanced pass
        if lb_deg == 256:
            assert st["heavy_chunks"] > 0
        say("lb rmat20 bfs", lb_deg)
        d, st = run(dev, g.nnodes, bfs_into, start_node=src, enable_lb=1, lb_deg=lb_deg)
        assert np.array_equal(d, lv), lb_deg


def test_lb_more_than_2_24_edges_in_one_round():
    """The reference bounds the edge-balanced loop by ceilf((float)total / nthreads), which
    loses edges once a round's heavy edges exceed 2^24.  RMAT-22 BFS from the hub with every
    frontier vertex in the edge-balanced pass has such a round (checked from the oracle's
    levels); the levels must still be exact."""
    g = CSRGraph.rmat(22, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    src = g.max_degree_vertex()
    lv, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    deg = np.diff(g.row_start.astype(np.int64))
    fin = lv != INF
    per_level = np.bincount(lv[fin], weights=deg[fin])
    assert per_level.max() > 2 ** 24, per_level.max()
    dev = g.copy_to_gpu()
    say("lb rmat22 bfs lb_deg=1")
    d, st = run(dev, g.nnodes, bfs_into, start_node=src, enable_lb=1, lb_deg=1)
    assert np.array_equal(d, lv)
    assert st["heavy_chunks"] == int(deg[fin].sum())  # every reached vertex's edges, once
    dev.close()


def test_lb_rejects_bad_threshold(rmat20):
    g, dev = rmat20
    with pytest.raises(LsgError):
        run(dev, g.nnodes, sssp_into, delta=32, enable_lb=1, lb_deg=-1)


def test_cli_l_flag(tmp_path):
    """./sssp -l and ./bfs -l: the DiffChecker contract (apps/sssp/bmktest2.py:19) on the -l path."""
    g = CSRGraph.rmat(14, 16, seed=7)
    gr = str(tmp_path / "g.gr")
    g.write(gr)
    src = g.max_degree_vertex()
    for app, ref in (("sssp", oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)),
                     ("bfs", oracle.bfs(g.row_start, g.edge_dst, src)[0])):
        out = str(tmp_path / f"{app}.txt")
        exe = os.path.join(ROOT, "galoisgpu_amd", "bin", app)
        args = [exe, "-l", "-s", str(src), "-o", out, gr]
        if app == "sssp":
            args[1:1] = ["-d", "32"]
        say("cli -l", app)
        r = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "Total time:" in r.stderr
        want = "".join(("%d INF\n" % i) if x == INF else ("%d %d\n" % (i, x)) for i, x in enumerate(ref))
        assert open(out).read() == want, app
