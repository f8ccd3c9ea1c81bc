"""The sssp and bfs apps: options, gg_main and output, over the C ABI.

Mirrors the reference app surface:
  * ``Options`` / ``process_prog_opt``  -- apps/sssp/support.cu:6-31, apps/bfs/support.cu:5-25
  * ``sssp_gg_main`` / ``bfs_gg_main``  -- apps/sssp/kernel.cu:455-480, apps/bfs/kernel.cu:380-393
  * ``output``                         -- apps/sssp/support.cu:34-53, apps/bfs/support.cu:27-46
``sssp`` and ``bfs`` run the whole skeleton sequence of skelapp/skel.cu:45-78
(copy to GPU, gg_main, copy back) and return the distance array.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields
from typing import Optional, Tuple

import numpy as np

from ._lib import EXPAND_LB, EXPAND_TWB, FAR_REFERENCE, FAR_SPLIT, INF, LsgOpts, LsgStats, check, lib
from .graph import CSRGraph, DeviceGraph


@dataclass
class Options:
    delta: int = 10000          # -d, apps/sssp/support.cu:10
    start_node: int = 0         # -s, apps/sssp/support.cu:12
    enable_lb: bool = False     # -l
    expand_mode: int = EXPAND_LB
    far_mode: int = FAR_REFERENCE
    block_size: int = 0
    grid_blocks: int = 0
    thread_deg: int = 0
    heavy_deg: int = 0
    sched: int = 0
    trace: bool = False
    graph_launch: bool = False
    lb_deg: int = 0             # -l threshold; 0 = the sweep's thread count (DEGREE_LIMIT)
    single_block: bool = False  # near/far rounds in one workgroup (small high-diameter graphs; per graph in sssp_graphs)

    def to_c(self) -> LsgOpts:
        o = LsgOpts()
        lib().lsg_opts_default(ctypes.byref(o))
        for f in fields(self):
            setattr(o, f.name, int(getattr(self, f.name)))
        return o


def process_prog_opt(opts: Options, c: str, optarg: Optional[str] = None, app: str = "sssp") -> None:
    """Applies one command-line option the way support.cu does; bad values raise ValueError
    where the reference asserts (apps/sssp/support.cu:22,29)."""
    if c == "d" and app == "sssp":
        v = int(optarg)
        if v <= 0:
            raise ValueError(f"delta must be > 0 (got {optarg})")
        opts.delta = v
    elif c == "l":
        opts.enable_lb = True
    elif c == "w":  # no reference counterpart: the one-workgroup first rounds (single_block)
        opts.single_block = True
    elif c == "s":
        v = int(optarg)
        if v < 0:
            raise ValueError(f"start node must be >= 0 (got {optarg})")
        opts.start_node = v
    else:
        raise ValueError(f"unknown option -{c} for {app}")


def sssp_gg_main(host: Optional[CSRGraph], dev: DeviceGraph, opts: Options) -> dict:
    st = LsgStats()
    hg = ctypes.byref(host._g) if host is not None else None
    check(lib().lsg_sssp_gg_main(hg, ctypes.byref(dev._g), ctypes.byref(opts.to_c()), ctypes.byref(st)),
          "sssp gg_main")
    return st.as_dict()


def bfs_gg_main(host: Optional[CSRGraph], dev: DeviceGraph, opts: Options) -> dict:
    st = LsgStats()
    hg = ctypes.byref(host._g) if host is not None else None
    check(lib().lsg_bfs_gg_main(hg, ctypes.byref(dev._g), ctypes.byref(opts.to_c()), ctypes.byref(st)),
          "bfs gg_main")
    return st.as_dict()


def sssp_into(dev: DeviceGraph, dist_ptr: int, opts: Options) -> dict:
    """gg_main writing into a caller-owned device array of dev.nnodes int32."""
    st = LsgStats()
    check(lib().lsg_sssp_dist(ctypes.byref(dev._g), ctypes.byref(opts.to_c()), ctypes.c_void_p(dist_ptr),
                              ctypes.byref(st)), "sssp")
    return st.as_dict()


def bfs_into(dev: DeviceGraph, level_ptr: int, opts: Options) -> dict:
    st = LsgStats()
    check(lib().lsg_bfs_dist(ctypes.byref(dev._g), ctypes.byref(opts.to_c()), ctypes.c_void_p(level_ptr),
                             ctypes.byref(st)), "bfs")
    return st.as_dict()


def cc_into(dev: DeviceGraph, label_ptr: int) -> dict:
    """Connected components (lsg_cc): label[v] = smallest id of v's component
    (edge directions ignored); stats['items'] is the number of components."""
    st = LsgStats()
    n = ctypes.c_uint32()
    check(lib().lsg_cc(ctypes.byref(dev._g), ctypes.c_void_p(label_ptr), ctypes.byref(n), ctypes.byref(st)), "cc")
    return st.as_dict()


def cc(graph: CSRGraph, device: int = 0) -> Tuple[np.ndarray, dict]:
    """apps/cc gg_main on a host graph: (labels, stats)."""
    def run(g, dev, _opts):
        st = LsgStats()
        check(lib().lsg_cc_gg_main(ctypes.byref(g._g), ctypes.byref(dev._g), ctypes.byref(st)), "cc")
        return st.as_dict()
    return _run(graph, Options(), device, run)


PR_ALPHA, PR_EPSILON = 0.85, 1e-6  # apps/pr/kernel.cu:19-20


def pagerank_into(dev: DeviceGraph, rev: DeviceGraph, rank_ptr: int, alpha: float = PR_ALPHA,
                  epsilon: float = PR_EPSILON, max_iterations: int = 2 ** 31 - 1) -> dict:
    """PageRank (lsg_pagerank) into a device float array; `rev` is dev.transpose()."""
    st = LsgStats()
    check(lib().lsg_pagerank(ctypes.byref(dev._g), ctypes.byref(rev._g), alpha, epsilon, max_iterations,
                             ctypes.c_void_p(rank_ptr), ctypes.byref(st)), "pagerank")
    return st.as_dict()


BFS_AUTO, BFS_TOP_DOWN, BFS_BOTTOM_UP = 0, 1, 2


def bfs_do_into(dev: DeviceGraph, rev: DeviceGraph, level_ptr: int, start_node: int = 0,
                direction: int = BFS_AUTO) -> dict:
    """Direction-optimising BFS (lsg_bfs_do); `rev` is dev.transpose()."""
    st = LsgStats()
    check(lib().lsg_bfs_do(ctypes.byref(dev._g), ctypes.byref(rev._g), start_node, direction,
                           ctypes.c_void_p(level_ptr), ctypes.byref(st)), "bfs_do")
    return st.as_dict()


def _run(graph: CSRGraph, opts: Options, device: int, fn) -> Tuple[np.ndarray, dict]:
    dev = graph.copy_to_gpu(device)
    try:
        stats = fn(graph, dev, opts)
        dev.copy_to_cpu(graph)
    finally:
        dev.close()
    return graph.node_data.copy(), stats


def sssp(graph: CSRGraph, start_node: int = 0, delta: int = 10000, device: int = 0,
         **tuning) -> Tuple[np.ndarray, dict]:
    """Skeleton sequence for apps/sssp: copy_to_gpu, gg_main, copy_to_cpu.  Returns (dist, stats)."""
    return _run(graph, Options(delta=delta, start_node=start_node, **tuning), device, sssp_gg_main)


def bfs(graph: CSRGraph, start_node: int = 0, device: int = 0, **tuning) -> Tuple[np.ndarray, dict]:
    return _run(graph, Options(start_node=start_node, **tuning), device, bfs_gg_main)


TRACE_DTYPE = np.dtype([("round", "<u4"), ("phase", "<u4"), ("in_items", "<u4"), ("heavy_chunks", "<u4"),
                        ("near_out", "<u4"), ("far_out", "<u4"), ("heavy_out", "<u4"), ("curdelta_lo", "<u4"),
                        ("curdelta_hi", "<u4"), ("delta", "<u4"), ("t_end", "<u8")])


def last_trace(device: int = 0) -> Tuple[np.ndarray, float]:
    """Round records of the last call made with Options(trace=True): (records, tick_hz)."""
    from ._lib import LsgRoundTrace
    n = ctypes.c_uint32()
    hz = ctypes.c_double()
    check(lib().lsg_last_trace(device, None, 0, ctypes.byref(n), ctypes.byref(hz)), "last_trace")
    buf = (LsgRoundTrace * max(1, n.value))()
    check(lib().lsg_last_trace(device, buf, n.value, ctypes.byref(n), ctypes.byref(hz)), "last_trace")
    arr = np.frombuffer(bytes(buf), dtype=TRACE_DTYPE)[: n.value].copy()
    return arr, hz.value


def output(dist: np.ndarray, path: Optional[str]) -> None:
    """Writes "%d %d" / "%d INF" lines (apps/sssp/support.cu:45-51)."""
    d = np.ascontiguousarray(dist, dtype=np.int32)
    check(lib().lsg_output_dist(d.ctypes.data_as(ctypes.c_void_p), len(d),
                                None if path is None else str(path).encode()), "output")


__all__ = ["Options", "process_prog_opt", "sssp_gg_main", "bfs_gg_main", "sssp_into", "bfs_into", "sssp",
           "bfs", "output", "INF", "EXPAND_TWB", "EXPAND_LB", "FAR_REFERENCE", "FAR_SPLIT"]
