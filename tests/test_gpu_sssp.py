"""GPU parity: the HIP path (through the C ABI) against the oracle and golden vectors.

Integer work, so every comparison is bit-exact.  Sizes up to RMAT-20 compare
full distance arrays with the oracle's Dijkstra / BFS; larger graphs are
checked through the O(E) shortest-path certificate.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import ROOT
from galoisgpu_amd import CSRGraph, LsgError, Options
from galoisgpu_amd import apps

pytestmark = pytest.mark.gpu

INF = 0x7FFFFFFF


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def run_sssp(g, dev, src, delta, **kw):
    apps.sssp_gg_main(g, dev, Options(delta=delta, start_node=src, **kw))
    dev.copy_to_cpu(g)
    return g.node_data.copy()


def run_bfs(g, dev, src, **kw):
    apps.bfs_gg_main(g, dev, Options(start_node=src, **kw))
    dev.copy_to_cpu(g)
    return g.node_data.copy()


MODES = [
    dict(),
    dict(expand_mode=1),
    dict(block_size=512),
    dict(block_size=1024, expand_mode=1),
    dict(heavy_deg=256, thread_deg=2),
    dict(heavy_deg=256, expand_mode=1, grid_blocks=7),
    dict(far_mode=1),
    dict(far_mode=2, expand_mode=0),
    dict(far_mode=2, block_size=512, heavy_deg=512),
    dict(enable_lb=1, lb_deg=3),
]


def test_small_golden(small_graphs):
    for name, gg in small_graphs.items():
        g = CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"])
        dev = g.copy_to_gpu()
        for mode in MODES:
            for delta in (1, 5, 32, 10000):
                say("small", name, mode, delta)
                d = run_sssp(g, dev, gg["src"], delta, **mode)
                assert np.array_equal(d, gg["dist"]), (name, mode, delta)
            d = run_bfs(g, dev, gg["src"], **mode)
            assert np.array_equal(d, gg["bfs"]), (name, mode)
        dev.close()


def test_grid256_config1(grid256_golden):
    """Config 1: 256x256 grid, w in [1,100], source 0, delta 32 -- bit-exact against
    scipy's golden vector and the near/far restatement."""
    g = CSRGraph.grid(256, 256, 1, 100, 1)
    dev = g.copy_to_gpu()
    nf, _ = oracle.nearfar(g.row_start, g.edge_dst, g.edge_data, 0, 32)
    assert np.array_equal(nf, grid256_golden["dist"])
    for mode in MODES:
        say("grid256", mode)
        d = run_sssp(g, dev, 0, 32, **mode)
        assert np.array_equal(d, grid256_golden["dist"]), mode
    b = run_bfs(g, dev, 0)
    assert np.array_equal(b, grid256_golden["bfs"])


@pytest.fixture(scope="module")
def rmat20():
    g = CSRGraph.rmat(20, 16, seed=1)
    return g


def test_rmat20_config2_sssp(rmat20):
    g = rmat20
    src = g.max_degree_vertex()
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
    dev = g.copy_to_gpu()
    for delta, mode in ((32, dict()), (32, dict(expand_mode=1)), (32, dict(block_size=512, heavy_deg=1024)),
                        (32, dict(far_mode=1)), (32, dict(far_mode=2)), (7, dict(far_mode=2, expand_mode=0))):
        say("rmat20 sssp", delta, mode)
        d = run_sssp(g, dev, src, delta, **mode)
        assert np.array_equal(d, ref), (delta, mode)
    # another source, reference default delta
    src2 = int(np.flatnonzero(g.out_degree() > 0)[12345])
    ref2 = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src2)
    say("rmat20 sssp src2", src2)
    assert np.array_equal(run_sssp(g, dev, src2, 10000), ref2)
    dev.close()


def test_rmat20_config2_bfs(rmat20):
    g = rmat20
    src = g.max_degree_vertex()
    ref, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    dev = g.copy_to_gpu()
    for mode in (dict(), dict(expand_mode=1)):
        say("rmat20 bfs", mode)
        assert np.array_equal(run_bfs(g, dev, src, **mode), ref), mode
    dev.close()


def test_roadlike_config3():
    g = CSRGraph.roadlike(2048, 2048, 0.10, 0.001, 1, 10000, seed=1)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    dev = g.copy_to_gpu()
    for delta, mode in ((32, dict()), (10000, dict(expand_mode=1))):
        say("roadlike", delta, mode)
        d = run_sssp(g, dev, 0, delta, **mode)
        assert np.array_equal(d, ref), (delta, mode)
    dev.close()


def test_plain_learned_schedules_keep_distances():
    """The single-source kernel under forced and learned per-phase schedules
    (delta width x expansion mode) gives exactly the oracle's distances."""
    from galoisgpu_amd import sched as lsched
    cases = [(CSRGraph.roadlike(192, 192, 0.10, 0.001, 1, 10000, seed=5), 2000),
             (CSRGraph.rmat(14, 16, seed=5), 32)]
    base = lsched.new_model(0, kind="plain")
    rng = np.random.default_rng(3)
    schedules = [[a] for a in range(lsched.N_ACT)] + [list(rng.integers(0, lsched.N_ACT, 300)) for _ in range(2)]
    for g, delta in cases:
        src = g.max_degree_vertex()
        ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
        dev = g.copy_to_gpu()
        for sch in schedules:
            lsched.set_model(lsched.forced(base, sch), kind="plain")
            say("plain forced", g.nnodes, sch[:6])
            d = run_sssp(g, dev, src, delta, sched=1, trace=True)
            assert np.array_equal(d, ref), sch
            from galoisgpu_amd.batch import last_features
            feats = last_features()
            assert len(feats) > 0
            acts = feats[:, 16].astype(int)
            assert all(acts[p] == sch[min(p, len(sch) - 1)] for p in range(len(acts)))
            modes = feats[:, 18].astype(int)
            assert all(modes[p] == lsched.ACTIONS_BY_KIND["plain"][acts[p]][1] for p in range(len(acts)))
        lsched.set_model(base, kind="plain")  # an untrained network as the scheduler: still exact
        assert np.array_equal(run_sssp(g, dev, src, delta, sched=1), ref)
        with pytest.raises(LsgError, match="BFS"):
            run_bfs(g, dev, src, sched=1)
        dev.close()


def test_rmat22_sources_certificate():
    g = CSRGraph.rmat(22, 16, seed=1)
    dev = g.copy_to_gpu()
    rng = np.random.default_rng(4)
    for src in [g.max_degree_vertex()] + rng.integers(0, g.nnodes, 3).tolist():
        say("rmat22 src", src)
        d = run_sssp(g, dev, int(src), 32)
        assert oracle.certificate(g.row_start, g.edge_dst, g.edge_data, int(src), d) == 0
    dev.close()


def test_edge_cases(tmp_path):
    # single node, no edges
    g = CSRGraph.from_arrays([0, 0], [], [])
    dev = g.copy_to_gpu()
    assert run_sssp(g, dev, 0, 32).tolist() == [0]
    assert run_bfs(g, dev, 0).tolist() == [0]
    dev.close()
    # source with no out-edges; zero weights (sizeEdgeTy = 0 file)
    h = CSRGraph.rmat(10, 8, seed=2)
    sink = int(np.flatnonzero(h.out_degree() == 0)[0])
    dev = h.copy_to_gpu()
    d = run_sssp(h, dev, sink, 32)
    assert d[sink] == 0 and (np.delete(d, sink) == INF).all()
    with pytest.raises(LsgError) as e:
        run_sssp(h, dev, h.nnodes, 32)
    assert e.value.name == "LSG_ERR_ARG"
    with pytest.raises(LsgError):
        run_sssp(h, dev, 0, 0)
    dev.close()
    p = str(tmp_path / "z.gr")
    h.write(p, with_weights=False)
    z = CSRGraph.read(p)
    dz = z.copy_to_gpu()
    src = z.max_degree_vertex()
    d = run_sssp(z, dz, src, 1)
    b, _ = oracle.bfs(z.row_start, z.edge_dst, src)
    assert np.array_equal(d == 0, b != INF) and (d[b == INF] == INF).all()
    dz.close()


def test_workspace_reuse_shrinking_inputs():
    """Calls in a row on graphs of different sizes share the cached device workspace."""
    for scale in (16, 12, 14, 10):
        g = CSRGraph.rmat(scale, 8, seed=scale)
        src = g.max_degree_vertex()
        say("reuse", scale)
        dev = g.copy_to_gpu()
        d = run_sssp(g, dev, src, 32)
        assert np.array_equal(d, oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src))
        dev.close()


def test_barrier_timeout_is_reported(monkeypatch):
    """A zero spin limit, set from the host, makes the grid barrier give up at once:
    every wave still exits and the call reports LSG_ERR_TIMEOUT instead of hanging."""
    g = CSRGraph.rmat(12, 8, seed=5)
    dev = g.copy_to_gpu()
    monkeypatch.setenv("LSG_BARRIER_TIMEOUT_S", "0")
    say("forced barrier timeout")
    with pytest.raises(LsgError) as e:
        run_sssp(g, dev, g.max_degree_vertex(), 32)
    assert e.value.name == "LSG_ERR_TIMEOUT"
    monkeypatch.delenv("LSG_BARRIER_TIMEOUT_S")
    d = run_sssp(g, dev, g.max_degree_vertex(), 32)
    assert oracle.certificate(g.row_start, g.edge_dst, g.edge_data, g.max_degree_vertex(), d) == 0
    dev.close()


def test_cli_output_matches_oracle_text(tmp_path):
    """./sssp -d 32 -o out graph.gr: the DiffChecker contract (apps/sssp/bmktest2.py:19)."""
    g = CSRGraph.grid(64, 64, 1, 100, 3)
    gr = str(tmp_path / "g.gr")
    g.write(gr)
    out = str(tmp_path / "out.txt")
    exe = os.path.join(ROOT, "galoisgpu_amd", "bin", "sssp")
    say("cli sssp")
    r = subprocess.run([exe, "-d", "32", "-o", out, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Using GPU:" in r.stdout and "delta: 32" in r.stdout and "iterations:" in r.stdout
    assert "Total time:" in r.stderr
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    want = "".join(("%d INF\n" % i) if x == INF else ("%d %d\n" % (i, x)) for i, x in enumerate(ref))
    assert open(out).read() == want
    out2 = str(tmp_path / "bfs.txt")
    exe = os.path.join(ROOT, "galoisgpu_amd", "bin", "bfs")
    r = subprocess.run([exe, "-s", "5", "-o", out2, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lv, _ = oracle.bfs(g.row_start, g.edge_dst, 5)
    want = "".join(("%d INF\n" % i) if x == INF else ("%d %d\n" % (i, x)) for i, x in enumerate(lv))
    assert open(out2).read() == want
    say("cli -w")
    out4 = str(tmp_path / "bfs_w.txt")
    r = subprocess.run([exe, "-w", "-s", "5", "-o", out4, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert open(out4).read() == want
    out5 = str(tmp_path / "sssp_w.txt")
    r = subprocess.run([os.path.join(ROOT, "galoisgpu_amd", "bin", "sssp"), "-w", "-d", "32", "-o", out5, gr],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert open(out5).read() == "".join(("%d INF\n" % i) if x == INF else ("%d %d\n" % (i, x))
                                        for i, x in enumerate(ref))
    say("cli bfs -D")
    out3 = str(tmp_path / "bfs_do.txt")
    r = subprocess.run([exe, "-D", "-s", "5", "-o", out3, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "transpose:" in r.stderr and "Total time:" in r.stderr
    assert open(out3).read() == want


def test_bmk_harness_runs_the_cli(tmp_path):
    """The bmk2-style harness runs bin/sssp and bin/bfs, parses Total time and
    diff-checks the output against a reference written from the oracle."""
    from galoisgpu_amd import bmk
    g = CSRGraph.grid(48, 40, 1, 100, seed=12)
    path = str(tmp_path / "g.gr")
    g.write(path)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 5)
    apps.output(ref, str(tmp_path / "g.sssp.out"))
    lvl, _ = oracle.bfs(g.row_start, g.edge_dst, 5)
    apps.output(lvl, str(tmp_path / "g.bfs.out"))
    say("bmk sssp")
    r = bmk.run("sssp", path, delta=32, source=5, runs=2)
    assert r["check"] == "pass", r
    assert len(r["times_ns"]) == 2 and all(t and t > 0 for t in r["times_ns"])
    say("bmk bfs")
    assert bmk.run("bfs", path, source=5)["check"] == "pass"
    bad = str(tmp_path / "bad.out")
    ref[7] += 1
    apps.output(ref, bad)
    r = bmk.run("sssp", path, delta=32, source=5, ref=bad)
    assert r["check"] == "fail" and r["line"] == 8


def test_wrap_device_validates_the_csr():
    """lsg_graph_wrap_device checks caller-owned device arrays on the device: an edge out of
    range, decreasing offsets, a row_start that does not end at nedges and a negative weight
    are LSG_ERR_FORMAT; a good CSR runs and gives the oracle's distances."""
    from galoisgpu_amd.batch import DeviceArray
    from galoisgpu_amd.graph import DeviceGraph
    g = CSRGraph.rmat(12, 8, seed=9)
    row, dst, w = g.row_start.copy(), g.edge_dst.copy(), g.edge_data.copy()
    V, E = g.nnodes, g.nedges

    def wrap(r, d, x):
        bufs = [DeviceArray.from_numpy(np.ascontiguousarray(a)) for a in (r, d, x)]
        try:
            return DeviceGraph.wrap(V, E, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr), bufs
        except LsgError:
            for b in bufs:
                b.close()
            raise

    bad_dst = dst.copy()
    bad_dst[E // 2] = V
    bad_row = row.copy()
    k = V // 2
    bad_row[k] = row[k + 1] + 1  # offsets decrease from k to k + 1
    end_row = row.copy()
    end_row[V] = E - 1
    neg = w.copy()
    neg[E - 1] = -5
    for name, r, d, x in (("dst", row, bad_dst, w), ("offsets", bad_row, dst, w), ("end", end_row, dst, w),
                          ("weight", row, dst, neg)):
        say("wrap rejects", name)
        with pytest.raises(LsgError) as e:
            wrap(r, d, x)
        assert e.value.name == "LSG_ERR_FORMAT", name
    dg, bufs = wrap(row, dst, w)
    out = DeviceArray(V * 4)
    src = g.max_degree_vertex()
    apps.sssp_into(dg, out.ptr, Options(delta=32, start_node=src))
    assert np.array_equal(out.to_numpy(np.int32, V), oracle.dijkstra(row, dst, w, src))
    for b in bufs + [out]:
        b.close()


def test_workspace_release_then_allocating_call():
    """After lsg_workspace_release the next call allocates its buffers again and still gives
    the oracle's distances (the with-allocation timing of the reference's gg_main)."""
    from galoisgpu_amd._lib import check, lib
    g = CSRGraph.rmat(14, 16, seed=11)
    src = g.max_degree_vertex()
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
    dev = g.copy_to_gpu()
    for rep in range(2):
        check(lib().lsg_workspace_release(0), "release")
        assert np.array_equal(run_sssp(g, dev, src, 32), ref), rep
        assert np.array_equal(run_sssp(g, dev, src, 32, enable_lb=1, lb_deg=64), ref), rep
    dev.close()
