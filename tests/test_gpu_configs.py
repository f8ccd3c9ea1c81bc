"""GPU parity at the benchmark configurations' full sizes.

Each test runs one configuration exactly as the benchmark and the measurement
script run it (same generator, seeds, sources and options) and checks the
distances against the oracle: every column through the O(E) shortest-path
certificate (schedule-independent, exact for integer weights) and a sample
of columns bit for bit against the oracle's Dijkstra.

  C3   road-like 2048^2 grid, delta 10000, learned plain-kernel schedule
  C4A  RMAT-22 ef 16, 64 uniform sources in one batched launch (learned and
       autotuned fixed schedules, hipGraph launch)
  C4B  256 RMAT graphs of scales U[10,16], ef 8, one launch
  C5   RMAT-24 ef 16, 32 sources in one batched launch
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, Options
from galoisgpu_amd import autotune
from galoisgpu_amd import sched as lsched
from galoisgpu_amd.apps import sssp_into
from galoisgpu_amd.batch import DeviceArray, sssp_graphs, sssp_multi_into

pytestmark = pytest.mark.gpu
INF = 0x7FFFFFFF


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def certify_columns(g, srcs, D, threads=16):
    """Every column of D [V][S] passes the certificate for its source."""
    cols = [np.ascontiguousarray(D[:, k]) for k in range(D.shape[1])]

    def one(k):
        return oracle.certificate(g.row_start, g.edge_dst, g.edge_data, int(srcs[k]), cols[k])

    with ThreadPoolExecutor(max_workers=threads) as ex:
        bad = [k for k, r in enumerate(ex.map(one, range(len(cols)))) if r != 0]
    assert not bad, f"columns failing the certificate: {bad[:8]}"


def live_sources(g, srcs):
    deg = np.diff(g.row_start.astype(np.int64))
    return [k for k, s in enumerate(srcs) if deg[s] > 0]


@pytest.fixture(scope="module")
def rmat22():
    g = CSRGraph.rmat(22, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    srcs = np.random.default_rng(1001).integers(0, g.nnodes, 64).astype(np.uint32)  # bench.py's sources
    dev = g.copy_to_gpu()
    yield g, srcs, dev
    dev.close()


def test_c4a_rmat22_64_sources(rmat22):
    """C4A: the headline workload, one launch for 64 sources, under the schedules bench.py
    times (learned, the autotuned fixed one) and as a captured hipGraph."""
    g, srcs, dev = rmat22
    V = g.nnodes
    D = DeviceArray(V * 64 * 4)
    tuned = autotune.tuned_options("rmat22_64src", "batched") or {"delta": 1 << 30}
    variants = [("fixed", Options(**tuned)), ("fixed graph_launch", Options(graph_launch=True, **tuned)),
                ("reference delta 32", Options(delta=32))]
    model = lsched.default_model_path("batched")
    if os.path.exists(model):
        lsched.load_model(model, device=0, kind="batched")
        launch = {k: v for k, v in tuned.items() if k in ("heavy_deg", "block_size")}
        variants.insert(0, ("learned", Options(delta=32, sched=1, **launch)))
    first = None
    for name, o in variants:
        say("C4A", name, o)
        st = sssp_multi_into(dev, srcs, o, D)
        Dh = D.to_numpy(np.int32, (V, 64))
        assert st["launches"] == 1
        if first is None:
            certify_columns(g, srcs, Dh)
            live = live_sources(g, srcs)
            for k in live[:4] + [k for k in range(64) if k not in live][:1]:
                ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, int(srcs[k]))
                assert np.array_equal(Dh[:, k], ref), (name, k)
            first = Dh
        else:
            assert np.array_equal(Dh, first), name
    D.close()


def test_c4b_256_graphs_one_launch():
    """C4B at its specified size: 256 RMAT graphs, scales uniform in [10,16], edge factor 8,
    the max-degree vertex of each as its source, all in one persistent launch."""
    rng = np.random.default_rng(44)
    scales = rng.integers(10, 17, 256)
    assert scales.min() == 10 and scales.max() == 16
    graphs = [CSRGraph.rmat(int(s), 8, 0.57, 0.19, 0.19, 1, 255, seed=100 + i) for i, s in enumerate(scales)]
    srcs = [g.max_degree_vertex() for g in graphs]
    devs = [g.copy_to_gpu() for g in graphs]

    def ref(k):
        return oracle.dijkstra(graphs[k].row_start, graphs[k].edge_dst, graphs[k].edge_data, srcs[k])

    with ThreadPoolExecutor(max_workers=16) as ex:
        refs = list(ex.map(ref, range(len(graphs))))
    for name, o in (("reference delta 32", Options(delta=32)), ("graph_launch", Options(delta=32, graph_launch=True)),
                    ("bucket", Options(delta=32, far_mode=2)),
                    ("one workgroup per graph", Options(delta=32, single_block=True))):
        say("C4B", name)
        outs, st = sssp_graphs(devs, srcs, o)
        assert st["launches"] == 1
        if o.single_block:
            assert st["grid_blocks"] == len(graphs)  # no fallback to the union launch
        bad = [k for k in range(len(graphs)) if not np.array_equal(outs[k], refs[k])]
        assert not bad, (name, bad[:8])
    for d in devs:
        d.close()


def test_c5_rmat24_32_sources():
    """C5: RMAT-24 (16.8 M vertices, ~2.2 GB of CSR), 32 sources in one launch; every column
    certified, two against Dijkstra."""
    g = CSRGraph.rmat(24, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    V = g.nnodes
    srcs = np.random.default_rng(1024).integers(0, V, 32).astype(np.uint32)
    dev = g.copy_to_gpu()
    D = DeviceArray(V * 32 * 4)
    model = lsched.default_model_path("batched")
    if os.path.exists(model):
        lsched.load_model(model, device=0, kind="batched")
        o = Options(delta=32, sched=1, heavy_deg=256, block_size=256)
    else:
        o = Options(delta=1 << 30, heavy_deg=256, block_size=256)
    say("C5", o)
    st = sssp_multi_into(dev, srcs, o, D)
    assert st["launches"] == 1
    Dh = D.to_numpy(np.int32, (V, 32))
    certify_columns(g, srcs, Dh)
    for k in live_sources(g, srcs)[:2]:
        ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, int(srcs[k]))
        assert np.array_equal(Dh[:, k], ref), k
    D.close()
    dev.close()


def test_c3_road2048_learned_and_bucket():
    """C3 at full size: the 2048^2 road-like grid from vertex 0, delta 10000, under the learned
    plain-kernel schedule and the split and bucket far piles, against Dijkstra."""
    g = CSRGraph.roadlike(2048, 2048, 0.10, 0.001, 1, 10000, seed=1)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    dev = g.copy_to_gpu()
    buf = DeviceArray(g.nnodes * 4)
    variants = [("split", Options(delta=10000, far_mode=1)), ("bucket", Options(delta=10000, far_mode=2)),
                ("bucket delta 32", Options(delta=32, far_mode=2))]
    model = lsched.default_model_path("plain")
    if os.path.exists(model):
        lsched.load_model(model, device=0, kind="plain")
        variants.insert(0, ("learned", Options(delta=10000, sched=1)))
    for name, o in variants:
        say("C3", name)
        sssp_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), ref), name
    buf.close()
    dev.close()
