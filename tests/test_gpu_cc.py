"""GPU parity of connected components (apps/cc) through the C ABI and the CLI.

Labels must equal the oracle's (smallest vertex id of each weakly connected
component) bit for bit; the component count is the reference's
"components: %d" line (apps/cc/kernel.cu:719), and the CLI writes the
reference's output text (apps/cc/support.cu:17-36).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError
from galoisgpu_amd.apps import cc, cc_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def check(g, name):
    ref, k = oracle.cc(g.row_start, g.edge_dst)
    dev = g.copy_to_gpu()
    lab = DeviceArray(g.nnodes * 4)
    try:
        say("cc", name, g.nnodes, g.nedges)
        st = cc_into(dev, lab.ptr)
        got = lab.to_numpy(np.int32, g.nnodes)
        assert np.array_equal(got, ref), name
        assert st["items"] == k, (name, st["items"], k)
        # a second call on the same buffers gives the same answer
        cc_into(dev, lab.ptr)
        assert np.array_equal(lab.to_numpy(np.int32, g.nnodes), ref), name
    finally:
        lab.close()
        dev.close()


def random_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(np.bincount(src, minlength=n))
    return CSRGraph.from_arrays(row, dst[order], rng.integers(1, 100, m))


def test_cc_golden(small_graphs):
    for name, gg in small_graphs.items():
        check(CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"]), name)


def test_cc_many_components():
    # sparse random graphs: thousands of components of every size
    for seed, (n, m) in enumerate([(1000, 400), (20000, 12000), (100000, 90000)]):
        check(random_graph(n, m, seed), f"random {n} {m}")


def test_cc_generated_graphs():
    check(CSRGraph.grid(50, 80, 1, 100, seed=2), "grid")
    check(CSRGraph.rmat(12, 4, seed=5), "rmat12")
    check(CSRGraph.rmat(16, 8, seed=6), "rmat16")
    check(CSRGraph.roadlike(128, 128, 0.3, 0.0005, 1, 100, seed=3), "road with cuts")


def test_cc_edge_cases():
    check(CSRGraph.from_arrays(np.array([0, 0]), np.array([], dtype=np.uint32)), "single vertex")
    check(CSRGraph.from_arrays(np.zeros(11, dtype=np.int64), np.array([], dtype=np.uint32)), "no edges")
    # edges only from larger to smaller ids: one component labelled 0
    n = 500
    row = np.concatenate([[0, 0], np.arange(1, n)])
    check(CSRGraph.from_arrays(row, np.arange(0, n - 1)), "descending chain")
    dev = CSRGraph.rmat(4, 2, seed=1).copy_to_gpu()
    with pytest.raises(LsgError):
        cc_into(dev, 0)
    dev.close()


def test_cc_gg_main_and_cli(tmp_path):
    g = random_graph(3000, 2500, 9)
    ref, k = oracle.cc(g.row_start, g.edge_dst)
    lab, st = cc(g)
    assert np.array_equal(lab, ref) and st["items"] == k
    gr = str(tmp_path / "g.gr")
    g.write(gr)
    out = str(tmp_path / "cc.txt")
    exe = os.path.join(ROOT, "galoisgpu_amd", "bin", "cc")
    say("cli cc")
    r = subprocess.run([exe, "-o", out, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"components: {k}" in r.stdout and "Total time:" in r.stderr
    row, col, w = g.row_start, g.edge_dst, g.edge_data
    want = []
    for i in range(g.nnodes):
        want.append("%d %d\n" % (i, ref[i]))
        for j in range(row[i], row[i + 1]):
            want.append("\te %d: %d %d %d\n" % (j, col[j], w[j], ref[col[j]]))
    assert open(out).read() == "".join(want)


def test_cc_barrier_timeout_is_reported(monkeypatch):
    """A zero spin limit, set from the host, makes the persistent launch's grid barriers give
    up at once: every wave still exits and the call reports LSG_ERR_TIMEOUT instead of hanging;
    the next call with the default limit gives the oracle's labels."""
    g = random_graph(4096, 16384, 11)
    ref, _ = oracle.cc(g.row_start, g.edge_dst)
    dev = g.copy_to_gpu()
    lab = DeviceArray(g.nnodes * 4)
    try:
        monkeypatch.setenv("LSG_BARRIER_TIMEOUT_S", "0")
        say("cc forced barrier timeout")
        with pytest.raises(LsgError) as e:
            cc_into(dev, lab.ptr)
        assert e.value.name == "LSG_ERR_TIMEOUT"
        monkeypatch.delenv("LSG_BARRIER_TIMEOUT_S")
        cc_into(dev, lab.ptr)
        assert np.array_equal(lab.to_numpy(np.int32, g.nnodes), ref)
    finally:
        lab.close()
        dev.close()
