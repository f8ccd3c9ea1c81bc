"""GPU parity of PageRank (apps/pr) through the C ABI and the CLI.

Floating point: the GPU computes in float32 and stops when no rank changes by
more than 1e-6 of itself; the oracle is the double-precision fixpoint.  The
tolerance is relative 1e-4 per vertex (ranks are at least 1 - ALPHA = 0.15):
the stopping rule leaves at most 1e-6 * ALPHA / (1 - ALPHA) ~ 6e-6, and a
float sum over a hub's in-edges (up to ~40k terms, 64 lanes plus a butterfly)
rounds by a few 1e-5 at worst.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError
from galoisgpu_amd.apps import pagerank_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RTOL = 1e-4


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def run(g, **kw):
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    out = DeviceArray(g.nnodes * 4)
    try:
        st = pagerank_into(dev, rev, out.ptr, **kw)
        return out.to_numpy(np.float32, g.nnodes).astype(np.float64), st
    finally:
        out.close()
        rev.close()
        dev.close()


def check(g, name):
    ref, _ = oracle.pagerank(g.row_start, g.edge_dst)
    say("pagerank", name, g.nnodes, g.nedges)
    got, st = run(g)
    err = np.abs(got - ref) / ref
    assert err.max() <= RTOL, (name, float(err.max()), int(err.argmax()))
    return st


def random_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, dtype=np.int64)
    row[1:] = np.cumsum(np.bincount(src, minlength=n))
    return CSRGraph.from_arrays(row, dst[order], rng.integers(1, 100, m))


def test_pagerank_golden(small_graphs):
    for name, gg in small_graphs.items():
        check(CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"]), name)


def test_pagerank_closed_forms():
    n = 1000  # directed cycle: every rank is exactly 1
    got, _ = run(CSRGraph.from_arrays(np.arange(n + 1), (np.arange(n) + 1) % n))
    assert np.allclose(got, 1.0, rtol=RTOL, atol=0)
    got, st = run(CSRGraph.from_arrays(np.zeros(8, dtype=np.int64), np.array([], dtype=np.uint32)))
    assert np.allclose(got, 0.15, rtol=1e-6, atol=0) and st["iterations"] == 1


def test_pagerank_generated_graphs():
    check(random_graph(5000, 20000, 1), "random with dangling vertices")
    check(CSRGraph.grid(60, 90, 1, 100, seed=2), "grid")
    check(CSRGraph.rmat(12, 8, seed=5), "rmat12")
    st = check(CSRGraph.rmat(16, 16, seed=6), "rmat16")
    assert st["items"] > 0  # rmat16 has vertices summed by whole wavefronts


def test_pagerank_arguments():
    g = CSRGraph.rmat(6, 4, seed=1)
    with pytest.raises(LsgError):
        run(g, alpha=1.0)
    with pytest.raises(LsgError):
        run(g, epsilon=0.0)
    got, st = run(g, max_iterations=3)
    assert st["iterations"] == 3


def test_pagerank_cli(tmp_path):
    g = random_graph(3000, 9000, 4)
    ref, _ = oracle.pagerank(g.row_start, g.edge_dst)
    gr = str(tmp_path / "g.gr")
    g.write(gr)
    out = str(tmp_path / "pr.txt")
    exe = os.path.join(ROOT, "galoisgpu_amd", "bin", "pr")
    say("cli pr")
    r = subprocess.run([exe, "-o", out, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "PR took" in r.stdout and "Total time:" in r.stderr
    s = [l for l in r.stdout.splitlines() if l.startswith("sum: ")][0]
    assert s.endswith(f"({g.nnodes})") and abs(float(s.split()[1]) - ref.sum()) <= RTOL * ref.sum()
    lines = open(out).read().splitlines()
    assert len(lines) == g.nnodes
    ids = np.array([int(l.split()[0]) for l in lines])
    vals = np.array([float(l.split()[1]) for l in lines])
    assert np.array_equal(ids, np.arange(g.nnodes))
    assert np.all(np.abs(vals - ref) <= RTOL * ref + 1e-6)  # "%f" keeps 6 decimals
    r = subprocess.run([exe, "-n", "-x", "5", "-o", out, gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PR took 5 iterations" in r.stdout
    assert open(out).read() == "".join(f"{i}\n" for i in range(g.nnodes))
