"""Batched SSSP and device buffers over liblsg.

``DeviceArray`` is a device allocation made by liblsg itself (no second GPU
runtime in the process).  ``sssp_multi`` runs up to 64 sources of one graph in
one persistent launch (vertex-major distances ``[V][S]``); ``sssp_batch`` runs
any number of sources, 64 per launch, into one array per source.  Both give
the same distances, bit for bit, as the single-source path with the same
delta.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import LsgBatchStats, LsgGraph, check, lib
from .apps import Options
from .graph import DeviceGraph

H2D, D2H, D2D = 0, 1, 2


class DeviceArray:
    """A device buffer owned by this object (freed on close/__del__)."""

    def __init__(self, nbytes: int, device: int = 0):
        self.device = device
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().lsg_dev_malloc(device, self.nbytes, ctypes.byref(p)), "dev_malloc")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, device)
        check(lib().lsg_dev_copy(device, ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D),
              "dev_copy")
        return d

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().lsg_dev_copy(self.device, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                 out.nbytes, D2H), "dev_copy")
        return out

    def close(self) -> None:
        if self.ptr:
            lib().lsg_dev_free(self.device, ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sssp_multi_into(dev: DeviceGraph, srcs: Sequence[int], opts: Options, dist_vs: DeviceArray) -> dict:
    """Up to 64 sources in one launch into dist_vs ([V][len(srcs)] int32 on the device)."""
    s = np.ascontiguousarray(np.asarray(srcs, dtype=np.uint32))
    if dist_vs.nbytes < dev.nnodes * len(s) * 4:
        raise ValueError("distance matrix too small")
    st = LsgBatchStats()
    check(lib().lsg_sssp_multi(ctypes.byref(dev._g), s.ctypes.data_as(ctypes.c_void_p), len(s),
                               ctypes.byref(opts.to_c()), ctypes.c_void_p(dist_vs.ptr), ctypes.byref(st)),
          "sssp_multi")
    return st.as_dict()


def sssp_multi(dev: DeviceGraph, srcs: Sequence[int], opts: Optional[Options] = None) -> Tuple[np.ndarray, dict]:
    """Returns (dist [V][S] int32 on the host, stats)."""
    opts = opts or Options()
    d = DeviceArray(dev.nnodes * len(srcs) * 4, dev.device)
    try:
        st = sssp_multi_into(dev, srcs, opts, d)
        return d.to_numpy(np.int32, (dev.nnodes, len(srcs))), st
    finally:
        d.close()


def sssp_batch(dev: DeviceGraph, srcs: Sequence[int], opts: Optional[Options] = None) -> Tuple[List[np.ndarray], dict]:
    """Any number of sources, 64 per launch; returns one distance array per source."""
    opts = opts or Options()
    s = np.ascontiguousarray(np.asarray(srcs, dtype=np.uint32))
    outs = [DeviceArray(dev.nnodes * 4, dev.device) for _ in s]
    try:
        ptrs = (ctypes.c_void_p * len(s))(*[o.ptr for o in outs])
        st = LsgBatchStats()
        check(lib().lsg_sssp_batch(ctypes.byref(dev._g), s.ctypes.data_as(ctypes.c_void_p), len(s),
                                   ctypes.byref(opts.to_c()), ptrs, ctypes.byref(st)), "sssp_batch")
        return [o.to_numpy(np.int32, dev.nnodes) for o in outs], st.as_dict()
    finally:
        for o in outs:
            o.close()


def sssp_graphs(devs: Sequence[DeviceGraph], srcs: Sequence[int], opts: Optional[Options] = None
                ) -> Tuple[List[np.ndarray], dict]:
    """Many graphs, one source each, in one persistent launch (disjoint-union multi-seed run)."""
    opts = opts or Options()
    n = len(devs)
    if len(srcs) != n:
        raise ValueError("one source per graph")
    arr = (LsgGraph * n)(*[d._g for d in devs])
    s = np.ascontiguousarray(np.asarray(srcs, dtype=np.uint32))
    outs = [DeviceArray(d.nnodes * 4, d.device) for d in devs]
    try:
        ptrs = (ctypes.c_void_p * n)(*[o.ptr for o in outs])
        st = LsgBatchStats()
        check(lib().lsg_sssp_graphs(arr, n, s.ctypes.data_as(ctypes.c_void_p), ctypes.byref(opts.to_c()), ptrs,
                                    ctypes.byref(st)), "sssp_graphs")
        return [o.to_numpy(np.int32, d.nnodes) for o, d in zip(outs, devs)], st.as_dict()
    finally:
        for o in outs:
            o.close()


def last_features(device: int = 0) -> np.ndarray:
    """Per-phase records of the last traced batched call: [nphases][20] float32."""
    n = ctypes.c_uint32()
    check(lib().lsg_last_features(device, None, 0, ctypes.byref(n)), "last_features")
    out = np.zeros((n.value, 20), dtype=np.float32)
    if n.value:
        check(lib().lsg_last_features(device, out.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)),
              "last_features")
    return out


__all__ = ["DeviceArray", "sssp_multi", "sssp_multi_into", "sssp_batch", "sssp_graphs", "last_features"]
