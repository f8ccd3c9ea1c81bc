"""CPU oracle for the SSSP/BFS/CC/PR paths -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker.  It wraps oracle/liboracle.so
(oracle/oracle.c, a plain-C restatement of apps/sssp and apps/bfs, each
function citing the reference lines it follows) and adds a numpy restatement
of the .gr reader (rt/src/csr_graph.cu:187-224).

Parity pins for every function are in tests/test_oracle.py; none is unpinned.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
INF = 0x7FFFFFFF

_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "oracle.c")):
        build()
    l = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, i32, i64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64
    l.oracle_dijkstra.argtypes = [u32, P, P, P, u32, P]
    l.oracle_dijkstra.restype = ctypes.c_int
    l.oracle_nearfar.argtypes = [u32, P, P, P, u32, i32, P, P]
    l.oracle_nearfar.restype = ctypes.c_int
    l.oracle_bfs.argtypes = [u32, P, P, u32, P, P]
    l.oracle_bfs.restype = ctypes.c_int
    l.oracle_sssp_certificate.argtypes = [u32, P, P, P, u32, P]
    l.oracle_sssp_certificate.restype = i64
    l.oracle_reach.argtypes = [u32, P, P, P, P]
    l.oracle_reach.restype = None
    l.oracle_pagerank.argtypes = [u32, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    l.oracle_pagerank.restype = ctypes.c_int
    l.oracle_cc.argtypes = [u32, P, P, P]
    l.oracle_cc.restype = i64
    l.oracle_plant_bug.argtypes = [ctypes.c_int]
    l.oracle_plant_bug.restype = None
    _lib = l
    return l


def _arrs(row_start, edge_dst, edge_data=None):
    r = np.ascontiguousarray(row_start, dtype=np.uint32)
    c = np.ascontiguousarray(edge_dst, dtype=np.uint32)
    w = None if edge_data is None else np.ascontiguousarray(edge_data, dtype=np.int32)
    return r, c, w


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def dijkstra(row_start, edge_dst, edge_data, src: int) -> np.ndarray:
    r, c, w = _arrs(row_start, edge_dst, edge_data)
    n = len(r) - 1
    d = np.empty(n, dtype=np.int32)
    rc = lib().oracle_dijkstra(n, _p(r), _p(c), _p(w), src, _p(d))
    if rc:
        raise ValueError(f"oracle_dijkstra failed ({rc})")
    return d


def nearfar(row_start, edge_dst, edge_data, src: int, delta: int) -> Tuple[np.ndarray, dict]:
    """The reference's near/far pile algorithm (apps/sssp/kernel.cu:355-391), serial."""
    r, c, w = _arrs(row_start, edge_dst, edge_data)
    n = len(r) - 1
    d = np.empty(n, dtype=np.int32)
    st = np.zeros(5, dtype=np.uint64)
    rc = lib().oracle_nearfar(n, _p(r), _p(c), _p(w), src, delta, _p(d), _p(st))
    if rc:
        raise ValueError(f"oracle_nearfar failed ({rc})")
    return d, {"phases": int(st[0]), "near_rounds": int(st[1]), "items": int(st[2]), "edges": int(st[3]),
               "curdelta": int(st[4])}


def bfs(row_start, edge_dst, src: int) -> Tuple[np.ndarray, dict]:
    """Level-synchronous BFS (apps/bfs/kernel.cu:312-334), serial."""
    r, c, _ = _arrs(row_start, edge_dst)
    n = len(r) - 1
    d = np.empty(n, dtype=np.int32)
    st = np.zeros(3, dtype=np.uint64)
    rc = lib().oracle_bfs(n, _p(r), _p(c), src, _p(d), _p(st))
    if rc:
        raise ValueError(f"oracle_bfs failed ({rc})")
    return d, {"levels": int(st[0]), "items": int(st[1]), "edges": int(st[2])}


def certificate(row_start, edge_dst, edge_data, src: int, dist) -> int:
    """Number of violated shortest-path conditions (0 = dist is exact).  Sound for
    positive weights; edge_data=None means unit weights (BFS levels)."""
    r, c, w = _arrs(row_start, edge_dst, edge_data)
    d = np.ascontiguousarray(dist, dtype=np.int32)
    return int(lib().oracle_sssp_certificate(len(r) - 1, _p(r), _p(c), _p(w), src, _p(d)))


def reach(row_start, dist) -> Tuple[int, int]:
    """(V_r, E_r): vertices with a finite value and the sum of their out-degrees."""
    r = np.ascontiguousarray(row_start, dtype=np.uint32)
    d = np.ascontiguousarray(dist, dtype=np.int32)
    vr, er = ctypes.c_uint64(), ctypes.c_uint64()
    lib().oracle_reach(len(r) - 1, _p(r), _p(d), ctypes.byref(vr), ctypes.byref(er))
    return int(vr.value), int(er.value)


def cc(row_start, edge_dst) -> Tuple[np.ndarray, int]:
    """Weakly connected components (apps/cc): (label = smallest id of each
    vertex's component, number of components)."""
    r, c, _ = _arrs(row_start, edge_dst)
    n = len(r) - 1
    lab = np.empty(n, dtype=np.int32)
    k = lib().oracle_cc(n, _p(r), _p(c), _p(lab))
    if k < 0:
        raise ValueError("oracle_cc failed")
    return lab, int(k)


def pagerank(row_start, edge_dst, alpha: float = 0.85, tol: float = 1e-13, max_iter: int = 100000) -> Tuple[np.ndarray, int]:
    """PageRank fixpoint of apps/pr in double: (ranks, Jacobi iterations)."""
    r, c, _ = _arrs(row_start, edge_dst)
    n = len(r) - 1
    p = np.empty(n, dtype=np.float64)
    it = lib().oracle_pagerank(n, _p(r), _p(c), alpha, tol, max_iter, _p(p))
    if it < 0:
        raise ValueError("oracle_pagerank failed")
    return p, int(it)


def plant_bug(which: int) -> None:
    """Plants a known mistake in the oracle (0 clears) so tests can show their pins catch it."""
    lib().oracle_plant_bug(which)


def read_gr(path: str):
    """numpy restatement of readFromGR's byte layout (rt/src/csr_graph.cu:187-224).

    Returns (row_start u32[V+1], edge_dst u32[E], edge_data i32[E]); weights are 0
    when sizeEdgeTy is 0 (the reference's calloc, :37)."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:32].view("<u8")
    version, size_edge, nn, ne = (int(x) for x in hdr)
    if version != 1:
        raise ValueError("version")
    off = 32
    out_idx = raw[off:off + 8 * nn].view("<u8")
    off += 8 * nn
    outs = raw[off:off + 4 * ne].view("<u4")
    off += 4 * ne + (4 if ne % 2 else 0)
    row = np.zeros(nn + 1, dtype=np.uint32)
    row[1:] = out_idx.astype(np.uint32)
    if size_edge:
        w = raw[off:off + 4 * ne].view("<u4").astype(np.uint32).view(np.int32)
    else:
        w = np.zeros(ne, dtype=np.int32)
    return row, outs.astype(np.uint32), w.copy()


def write_gr(path: str, row_start, edge_dst, edge_data=None) -> None:
    """Independent numpy writer of the same layout, used to make hand-built test files."""
    row = np.asarray(row_start, dtype=np.uint64)
    dst = np.asarray(edge_dst, dtype="<u4")
    nn, ne = len(row) - 1, len(dst)
    parts = [np.array([1, 4 if edge_data is not None else 0, nn, ne], dtype="<u8").tobytes(),
             row[1:].astype("<u8").tobytes(), dst.tobytes()]
    if ne % 2:
        parts.append(b"\0\0\0\0")
    if edge_data is not None:
        parts.append(np.asarray(edge_data, dtype=np.int64).astype("<u4").tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))
