"""Host and device CSR graphs over liblsg (rt/include/csr_graph.h, rt/src/csr_graph.cu).

``CSRGraph`` is the host graph the reference's skeleton reads from a .gr file
(``CSRGraph::read``, rt/src/csr_graph.cu:239); ``DeviceGraph`` is its device
copy (``copy_to_gpu``, :250).  Arrays are exposed as numpy views of the
library's own buffers.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from ._lib import INF, LsgGraph, check, lib


def _view(ptr, n: int, dtype) -> np.ndarray:
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    ctype = ctypes.c_uint32 if dtype == np.uint32 else ctypes.c_int32
    buf = ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents
    return np.frombuffer(buf, dtype=dtype, count=n)


class CSRGraph:
    """Host CSR graph: row_start[V+1], edge_dst[E] (u32), edge_data[E] and node_data[V] (i32)."""

    def __init__(self):
        self._g = LsgGraph()
        self._owned = False

    # -- construction ------------------------------------------------------
    @classmethod
    def read(cls, path: str, quiet: bool = True) -> "CSRGraph":
        """readFromGR (rt/src/csr_graph.cu:152-237); malformed files raise LsgError(LSG_ERR_FORMAT)."""
        g = cls()
        check(lib().lsg_graph_read(ctypes.byref(g._g), str(path).encode(), int(quiet)), f"read {path}")
        g._owned = True
        return g

    @classmethod
    def from_arrays(cls, row_start, edge_dst, edge_data=None) -> "CSRGraph":
        row_start = np.ascontiguousarray(row_start, dtype=np.uint32)
        edge_dst = np.ascontiguousarray(edge_dst, dtype=np.uint32)
        n = len(row_start) - 1
        m = len(edge_dst)
        g = cls()
        check(lib().lsg_graph_alloc_host(ctypes.byref(g._g), n, m), "alloc_host")
        g._owned = True
        g.row_start[:] = row_start
        g.edge_dst[:] = edge_dst
        if edge_data is not None:
            g.edge_data[:] = np.ascontiguousarray(edge_data, dtype=np.int64).astype(np.int32)
        return g

    @classmethod
    def grid(cls, rows: int, cols: int, wlo: int = 1, whi: int = 100, seed: int = 1) -> "CSRGraph":
        g = cls()
        check(lib().lsg_gen_grid(ctypes.byref(g._g), rows, cols, wlo, whi, seed), "gen_grid")
        g._owned = True
        return g

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             wlo: int = 1, whi: int = 255, seed: int = 1) -> "CSRGraph":
        g = cls()
        check(lib().lsg_gen_rmat(ctypes.byref(g._g), scale, edge_factor, a, b, c, wlo, whi, seed), "gen_rmat")
        g._owned = True
        return g

    @classmethod
    def roadlike(cls, rows: int, cols: int, drop: float = 0.10, shortcut_frac: float = 0.001,
                 wlo: int = 1, whi: int = 10000, seed: int = 1) -> "CSRGraph":
        g = cls()
        check(lib().lsg_gen_roadlike(ctypes.byref(g._g), rows, cols, drop, shortcut_frac, wlo, whi, seed),
              "gen_roadlike")
        g._owned = True
        return g

    # -- accessors -----------------------------------------------------------
    @property
    def nnodes(self) -> int:
        return self._g.nnodes

    @property
    def nedges(self) -> int:
        return self._g.nedges

    @property
    def row_start(self) -> np.ndarray:
        return _view(self._g.row_start, self.nnodes + 1, np.uint32)

    @property
    def edge_dst(self) -> np.ndarray:
        return _view(self._g.edge_dst, self.nedges, np.uint32)

    @property
    def edge_data(self) -> np.ndarray:
        return _view(self._g.edge_data, self.nedges, np.int32)

    @property
    def node_data(self) -> np.ndarray:
        return _view(self._g.node_data, self.nnodes, np.int32)

    def out_degree(self) -> np.ndarray:
        return np.diff(self.row_start.astype(np.int64))

    def max_degree_vertex(self) -> int:
        return int(lib().lsg_max_degree_vertex(ctypes.byref(self._g)))

    # -- io ------------------------------------------------------------------
    def write(self, path: str, with_weights: bool = True) -> None:
        check(lib().lsg_graph_write(ctypes.byref(self._g), str(path).encode(), int(with_weights)), f"write {path}")

    def output(self, path: Optional[str]) -> None:
        """output() of the apps (apps/sssp/support.cu:34-53) for node_data."""
        check(lib().lsg_output(ctypes.byref(self._g), None if path is None else str(path).encode()), "output")

    def copy_to_gpu(self, device: int = 0) -> "DeviceGraph":
        d = DeviceGraph()
        check(lib().lsg_graph_copy_to_gpu(ctypes.byref(self._g), ctypes.byref(d._g), device), "copy_to_gpu")
        d._owned = True
        return d

    def close(self) -> None:
        if self._owned:
            lib().lsg_graph_dealloc(ctypes.byref(self._g))
            self._owned = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """Device copy of a CSRGraph (CSRGraph::copy_to_gpu / copy_to_cpu)."""

    def __init__(self):
        self._g = LsgGraph()
        self._owned = False

    @classmethod
    def wrap(cls, nnodes: int, nedges: int, row_start_ptr: int, edge_dst_ptr: int, edge_data_ptr: int,
             node_data_ptr: int = 0, device: int = 0) -> "DeviceGraph":
        """Borrows caller-owned device arrays (e.g. torch tensors' data_ptr())."""
        d = cls()
        check(lib().lsg_graph_wrap_device(ctypes.byref(d._g), nnodes, nedges, row_start_ptr, edge_dst_ptr,
                                          edge_data_ptr, node_data_ptr or None, device), "wrap_device")
        return d

    @property
    def nnodes(self) -> int:
        return self._g.nnodes

    @property
    def nedges(self) -> int:
        return self._g.nedges

    @property
    def device(self) -> int:
        return self._g.device

    def copy_to_cpu(self, host: CSRGraph) -> None:
        check(lib().lsg_graph_copy_to_cpu(ctypes.byref(self._g), ctypes.byref(host._g)), "copy_to_cpu")

    def transpose(self) -> "DeviceGraph":
        """Reverse CSR of this graph, built on its device (lsg_graph_transpose)."""
        t = DeviceGraph()
        check(lib().lsg_graph_transpose(ctypes.byref(self._g), ctypes.byref(t._g)), "transpose")
        t._owned = True
        return t

    def close(self) -> None:
        if self._owned:
            lib().lsg_graph_dealloc(ctypes.byref(self._g))
            self._owned = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["CSRGraph", "DeviceGraph", "INF"]
