"""GPU test of lsg_sssp_dist_noalloc, the single-source entry point that
allocates nothing: every buffer beyond the caller's graph and distance array
lives in a static device array.  Its output must equal lsg_sssp_dist's, bit for
bit, on two calls in a row, the second on a smaller graph (the static array
keeps the first call's contents), and it must refuse graphs beyond its limit.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError, Options
from galoisgpu_amd._lib import LsgStats, check, lib
from galoisgpu_amd.apps import sssp_into
from galoisgpu_amd.batch import DeviceArray

pytestmark = pytest.mark.gpu

MAX_NODES, MAX_EDGES = 1 << 20, 1 << 24


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def noalloc(dev, out, opts):
    st = LsgStats()
    check(lib().lsg_sssp_dist_noalloc(ctypes.byref(dev._g), ctypes.byref(opts.to_c()), ctypes.c_void_p(out.ptr),
                                      ctypes.byref(st)), "sssp_dist_noalloc")
    return st.as_dict()


def test_noalloc_matches_sssp_dist_on_two_calls():
    cases = [(CSRGraph.rmat(20, 16, 0.57, 0.19, 0.19, 1, 255, seed=1), 32, dict()),  # C2-sized: at the limits
             (CSRGraph.grid(64, 48, 1, 100, seed=4), 16, dict(expand_mode=0)),      # smaller, second
             (CSRGraph.rmat(12, 8, seed=7), 5, dict(far_mode=2, block_size=512))]
    for g, delta, kw in cases:
        assert g.nnodes <= MAX_NODES and g.nedges <= MAX_EDGES
        src = g.max_degree_vertex()
        dev = g.copy_to_gpu()
        want = DeviceArray(g.nnodes * 4)
        got = DeviceArray(g.nnodes * 4)
        o = Options(delta=delta, start_node=src, **kw)
        sssp_into(dev, want.ptr, o)
        ref = want.to_numpy(np.int32, g.nnodes)
        assert np.array_equal(ref, oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src))
        for rep in range(2):
            say("noalloc", g.nnodes, g.nedges, delta, kw, rep)
            st = noalloc(dev, got, o)
            assert np.array_equal(got.to_numpy(np.int32, g.nnodes), ref), (g.nnodes, rep)
            assert st["iterations"] > 0
        want.close()
        got.close()
        dev.close()


def test_noalloc_rejects_graphs_beyond_its_limit():
    g = CSRGraph.grid(1025, 1024, 1, 10, seed=2)  # 1,049,600 vertices > 2^20
    dev = g.copy_to_gpu()
    out = DeviceArray(g.nnodes * 4)
    with pytest.raises(LsgError) as e:
        noalloc(dev, out, Options(delta=8))
    assert e.value.name == "LSG_ERR_ARG" and "limit" in str(e.value)
    out.close()
    dev.close()
    g = CSRGraph.grid(32, 32, 1, 10, seed=2)
    dev = g.copy_to_gpu()
    out = DeviceArray(g.nnodes * 4)
    for kw in (dict(trace=True), dict(graph_launch=True), dict(enable_lb=1)):
        with pytest.raises(LsgError) as e:
            noalloc(dev, out, Options(delta=8, **kw))
        assert e.value.name == "LSG_ERR_ARG", kw
    out.close()
    dev.close()
