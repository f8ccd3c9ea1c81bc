"""GPU parity of the direction-optimising BFS and of the reverse CSR it reads.

Levels must equal the oracle's level-synchronous BFS (apps/bfs/kernel.cu:
312-334) bit for bit under every direction policy: automatic, top-down only
and bottom-up only.  The reverse CSR is checked against a numpy stable sort of the
edges by destination, array for array.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError, lib
from galoisgpu_amd.apps import BFS_AUTO, BFS_BOTTOM_UP, BFS_TOP_DOWN, bfs_do_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu
D2H = 1


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def fetch(ptr, n, device=0):
    out = np.empty(n, dtype=np.uint32)
    if n:
        assert lib().lsg_dev_copy(device, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes,
                                  D2H) == 0
    return out


def check_levels(g, srcs, dirs=(BFS_AUTO, BFS_TOP_DOWN, BFS_BOTTOM_UP)):
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    lv = DeviceArray(max(1, g.nnodes) * 4)
    try:
        for s in srcs:
            ref, _ = oracle.bfs(g.row_start, g.edge_dst, int(s))
            for d in dirs:
                st = bfs_do_into(dev, rev, lv.ptr, int(s), d)
                got = lv.to_numpy(np.int32, g.nnodes)
                assert np.array_equal(got, ref), (s, d)
                if d == BFS_BOTTOM_UP:
                    assert st["near_iters"] == st["iterations"] + 1  # every launch bottom-up
    finally:
        lv.close()
        rev.close()
        dev.close()


def test_transpose_matches_numpy():
    g = CSRGraph.rmat(12, 8, seed=4)
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    V, E = g.nnodes, g.nedges
    row = fetch(ctypes.cast(rev._g.row_start, ctypes.c_void_p).value, V + 1)
    src = fetch(ctypes.cast(rev._g.edge_dst, ctypes.c_void_p).value, E)
    w = fetch(ctypes.cast(rev._g.edge_data, ctypes.c_void_p).value, E)
    r = g.row_start.astype(np.int64)
    col = g.edge_dst.astype(np.int64)
    sources = np.repeat(np.arange(V), np.diff(r))
    indeg = np.bincount(col, minlength=V)
    assert np.array_equal(row, np.concatenate([[0], np.cumsum(indeg)]).astype(np.uint32))
    # each row lists its in-edges in forward order (a stable sort of the edges by destination)
    order = np.argsort(col, kind="stable")
    assert np.array_equal(src, sources[order].astype(np.uint32))
    assert np.array_equal(w, g.edge_data[order].astype(np.uint32))
    # the same arrays on a second call
    rev2 = dev.transpose()
    assert np.array_equal(fetch(ctypes.cast(rev2._g.edge_dst, ctypes.c_void_p).value, E), src)
    rev2.close()
    rev.close()
    dev.close()


@pytest.mark.parametrize("scale,ef", [(8, 2), (17, 8)])
def test_transpose_larger_graphs(scale, ef):
    # two and three radix passes, many tiles, hub rows of thousands of in-edges
    g = CSRGraph.rmat(scale, ef, seed=scale)
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    V, E = g.nnodes, g.nedges
    say("transpose rmat", scale, ef)
    row = fetch(ctypes.cast(rev._g.row_start, ctypes.c_void_p).value, V + 1)
    src = fetch(ctypes.cast(rev._g.edge_dst, ctypes.c_void_p).value, E)
    col = g.edge_dst.astype(np.int64)
    order = np.argsort(col, kind="stable")
    sources = np.repeat(np.arange(V), np.diff(g.row_start.astype(np.int64)))
    assert np.array_equal(row, np.concatenate([[0], np.cumsum(np.bincount(col, minlength=V))]).astype(np.uint32))
    assert np.array_equal(src, sources[order].astype(np.uint32))
    rev.close()
    dev.close()


def test_bfs_do_golden(small_graphs):
    for name, gg in small_graphs.items():
        g = CSRGraph.from_arrays(gg["row"], gg["col"], gg["w"])
        say("bfs_do golden", name)
        check_levels(g, [gg["src"], 0, g.nnodes - 1])


def test_bfs_do_grid_closed_form():
    # unit-weight grid: the level of (r, c) from (0, 0) is r + c
    g = CSRGraph.grid(40, 70, 1, 1, seed=1)
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    lv = DeviceArray(g.nnodes * 4)
    for d in (BFS_AUTO, BFS_TOP_DOWN, BFS_BOTTOM_UP):
        bfs_do_into(dev, rev, lv.ptr, 0, d)
        got = lv.to_numpy(np.int32, g.nnodes).reshape(40, 70)
        assert np.array_equal(got, np.add.outer(np.arange(40), np.arange(70))), d
    lv.close()
    rev.close()
    dev.close()


@pytest.mark.parametrize("scale,ef", [(10, 4), (14, 16), (16, 8)])
def test_bfs_do_rmat(scale, ef):
    g = CSRGraph.rmat(scale, ef, seed=scale)
    srcs = [g.max_degree_vertex()] + [int(x) for x in np.random.default_rng(scale).integers(0, g.nnodes, 3)]
    say("bfs_do rmat", scale, ef)
    check_levels(g, srcs)


def test_bfs_do_roadlike():
    g = CSRGraph.roadlike(96, 96, 0.1, 0.001, 1, 100, seed=2)
    say("bfs_do roadlike")
    check_levels(g, [0, 4000])


def test_bfs_do_edge_cases():
    # a single vertex, a sink source, and an isolated vertex beside a component
    one = CSRGraph.from_arrays(np.array([0, 0]), np.array([], dtype=np.uint32))
    check_levels(one, [0])
    g = CSRGraph.from_arrays(np.array([0, 1, 2, 2, 2]), np.array([1, 2]))
    check_levels(g, [0, 2, 3])
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    lv = DeviceArray(g.nnodes * 4)
    with pytest.raises(LsgError):
        bfs_do_into(dev, rev, lv.ptr, g.nnodes, BFS_AUTO)
    with pytest.raises(LsgError):
        bfs_do_into(dev, rev, lv.ptr, 0, 3)
    with pytest.raises(LsgError, match="reverse graph"):
        other = CSRGraph.rmat(4, 2, seed=1).copy_to_gpu()
        bfs_do_into(dev, other, lv.ptr, 0, BFS_AUTO)
    lv.close()
    rev.close()
    dev.close()


def test_bfs_do_barrier_timeout_is_reported(monkeypatch):
    """A zero spin limit, set from the host, makes the persistent launch's grid barrier give
    up at once: every wave still exits and the call reports LSG_ERR_TIMEOUT instead of hanging;
    the next call with the default limit is exact."""
    g = CSRGraph.rmat(12, 8, seed=5)
    src = g.max_degree_vertex()
    ref, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    dev = g.copy_to_gpu()
    rev = dev.transpose()
    lv = DeviceArray(g.nnodes * 4)
    monkeypatch.setenv("LSG_BARRIER_TIMEOUT_S", "0")
    say("bfs_do forced barrier timeout")
    with pytest.raises(LsgError) as e:
        bfs_do_into(dev, rev, lv.ptr, src, BFS_AUTO)
    assert e.value.name == "LSG_ERR_TIMEOUT"
    monkeypatch.delenv("LSG_BARRIER_TIMEOUT_S")
    bfs_do_into(dev, rev, lv.ptr, src, BFS_AUTO)
    assert np.array_equal(lv.to_numpy(np.int32, g.nnodes), ref)
    lv.close()
    rev.close()
    dev.close()
