"""ctypes binding of liblsg.so, the C ABI declared in include/lsg.h.

The library is the product: there is no Python or CPU fallback.  If
liblsg.so is missing or cannot be loaded, every entry point raises
``LsgError`` naming the build step (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C galoisgpu_amd/csrc``).
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_size_t, c_uint32, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsg.so")

INF = 0x7FFFFFFF  # INT_MAX, apps/sssp/kernel.cu:18

LSG_OK = 0
ERRORS = {
    -1: "LSG_ERR_ARG",
    -2: "LSG_ERR_IO",
    -3: "LSG_ERR_FORMAT",
    -4: "LSG_ERR_NOMEM",
    -5: "LSG_ERR_HIP",
    -6: "LSG_ERR_CAPACITY",
    -7: "LSG_ERR_TIMEOUT",
    -8: "LSG_ERR_NODEV",
}
EXPAND_TWB, EXPAND_LB = 0, 1
FAR_REFERENCE, FAR_SPLIT, FAR_BUCKET = 0, 1, 2
SCHED_BATCHED, SCHED_PLAIN = 0, 1  # scheduler model kinds


class LsgError(RuntimeError):
    """A negative return code from liblsg, with the library's message."""

    def __init__(self, code: int, where: str, message: str):
        self.code = code
        self.name = ERRORS.get(code, f"code {code}")
        super().__init__(f"{where}: {self.name}: {message}")


class LsgGraph(ctypes.Structure):
    """struct lsg_graph, mirroring CSRGraph (rt/include/csr_graph.h:96-101)."""

    _fields_ = [
        ("nnodes", c_uint32),
        ("nedges", c_uint32),
        ("row_start", POINTER(c_uint32)),
        ("edge_dst", POINTER(c_uint32)),
        ("edge_data", POINTER(c_int32)),
        ("node_data", POINTER(c_int32)),
        ("device_graph", c_int32),
        ("device", c_int32),
    ]


class LsgOpts(ctypes.Structure):
    _fields_ = [
        ("delta", c_int32),
        ("start_node", c_int32),
        ("enable_lb", c_int32),
        ("expand_mode", c_int32),
        ("far_mode", c_int32),
        ("block_size", c_int32),
        ("grid_blocks", c_int32),
        ("thread_deg", c_int32),
        ("heavy_deg", c_int32),
        ("sched", c_int32),
        ("trace", c_int32),
        ("graph_launch", c_int32),
        ("lb_deg", c_int32),
        ("single_block", c_int32),
        ("reserved", c_int32 * 2),
    ]


class LsgRoundTrace(ctypes.Structure):
    _fields_ = [
        ("round", c_uint32),
        ("phase", c_uint32),
        ("in_items", c_uint32),
        ("heavy_chunks", c_uint32),
        ("near_out", c_uint32),
        ("far_out", c_uint32),
        ("heavy_out", c_uint32),
        ("curdelta_lo", c_uint32),
        ("curdelta_hi", c_uint32),
        ("delta", c_uint32),
        ("t_end", c_uint64),
    ]


class LsgStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", c_double),
        ("ms_kernel", c_double),
        ("iterations", c_uint64),
        ("near_iters", c_uint64),
        ("items", c_uint64),
        ("heavy_chunks", c_uint64),
        ("last_delta", c_int32),
        ("grid_blocks", c_int32),
        ("block_size", c_int32),
        ("reserved", c_int32 * 5),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class LsgBatchStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", c_double),
        ("ms_kernel", c_double),
        ("rounds", c_uint64),
        ("phases", c_uint64),
        ("items", c_uint64),
        ("heavy_chunks", c_uint64),
        ("lane_items", c_uint64),
        ("edges_expanded", c_uint64),
        ("improved_rows", c_uint64),
        ("grid_blocks", c_int32),
        ("block_size", c_int32),
        ("launches", c_int32),
        ("reserved", c_int32 * 5),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


# name -> (restype, argtypes); every symbol include/lsg.h declares.
SIGNATURES = {
    "lsg_abi_version": (c_int, []),
    "lsg_last_error": (c_char_p, []),
    "lsg_opts_default": (None, [POINTER(LsgOpts)]),
    "lsg_graph_read": (c_int, [POINTER(LsgGraph), c_char_p, c_int]),
    "lsg_graph_write": (c_int, [POINTER(LsgGraph), c_char_p, c_int]),
    "lsg_graph_alloc_host": (c_int, [POINTER(LsgGraph), c_uint32, c_uint32]),
    "lsg_graph_copy_to_gpu": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), c_int]),
    "lsg_graph_copy_to_cpu": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph)]),
    "lsg_graph_dealloc": (None, [POINTER(LsgGraph)]),
    "lsg_graph_wrap_device": (
        c_int,
        [POINTER(LsgGraph), c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_int],
    ),
    "lsg_gen_grid": (c_int, [POINTER(LsgGraph), c_uint32, c_uint32, c_uint32, c_uint32, c_uint64]),
    "lsg_gen_rmat": (
        c_int,
        [POINTER(LsgGraph), c_uint32, c_uint32, c_double, c_double, c_double, c_uint32, c_uint32, c_uint64],
    ),
    "lsg_gen_roadlike": (
        c_int,
        [POINTER(LsgGraph), c_uint32, c_uint32, c_double, c_double, c_uint32, c_uint32, c_uint64],
    ),
    "lsg_max_degree_vertex": (c_uint32, [POINTER(LsgGraph)]),
    "lsg_sssp_gg_main": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), POINTER(LsgOpts), POINTER(LsgStats)]),
    "lsg_bfs_gg_main": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), POINTER(LsgOpts), POINTER(LsgStats)]),
    "lsg_sssp_dist": (c_int, [POINTER(LsgGraph), POINTER(LsgOpts), c_void_p, POINTER(LsgStats)]),
    "lsg_bfs_dist": (c_int, [POINTER(LsgGraph), POINTER(LsgOpts), c_void_p, POINTER(LsgStats)]),
    "lsg_last_trace": (c_int, [c_int, POINTER(LsgRoundTrace), c_uint32, POINTER(c_uint32), POINTER(c_double)]),
    "lsg_graph_transpose": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph)]),
    "lsg_cc": (c_int, [POINTER(LsgGraph), c_void_p, POINTER(c_uint32), POINTER(LsgStats)]),
    "lsg_cc_gg_main": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), POINTER(LsgStats)]),
    "lsg_output_cc": (c_int, [POINTER(LsgGraph), c_char_p]),
    "lsg_pagerank": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), c_float, c_float, c_int32, c_void_p,
                             POINTER(LsgStats)]),
    "lsg_output_pr": (c_int, [c_void_p, c_uint32, c_char_p, c_int]),
    "lsg_bfs_do": (c_int, [POINTER(LsgGraph), POINTER(LsgGraph), c_uint32, c_int, c_void_p, POINTER(LsgStats)]),
    "lsg_output": (c_int, [POINTER(LsgGraph), c_char_p]),
    "lsg_output_dist": (c_int, [c_void_p, c_uint32, c_char_p]),
    "lsg_sssp_multi": (c_int, [POINTER(LsgGraph), c_void_p, c_uint32, POINTER(LsgOpts), c_void_p,
                               POINTER(LsgBatchStats)]),
    "lsg_sssp_batch": (c_int, [POINTER(LsgGraph), c_void_p, c_uint32, POINTER(LsgOpts), c_void_p,
                               POINTER(LsgBatchStats)]),
    "lsg_sssp_graphs": (c_int, [POINTER(LsgGraph), c_uint32, c_void_p, POINTER(LsgOpts), c_void_p,
                                POINTER(LsgBatchStats)]),
    "lsg_last_features": (c_int, [c_int, c_void_p, c_uint32, POINTER(c_uint32)]),
    "lsg_sched_model_floats": (c_int, []),
    "lsg_sched_set_model": (c_int, [c_int, c_int, c_void_p, c_uint32]),
    "lsg_mlp_train": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_float,
                              c_float, c_uint64, c_void_p]),
    "lsg_mlp_infer": (c_int, [c_int, c_void_p, c_void_p, c_uint32, c_void_p]),
    "lsg_workspace_release": (c_int, [c_int]),
    "lsg_sssp_dist_noalloc": (c_int, [POINTER(LsgGraph), POINTER(LsgOpts), c_void_p, POINTER(LsgStats)]),
    "lsg_dev_malloc": (c_int, [c_int, c_size_t, POINTER(c_void_p)]),
    "lsg_dev_free": (c_int, [c_int, c_void_p]),
    "lsg_dev_copy": (c_int, [c_int, c_void_p, c_void_p, c_size_t, c_int]),
    "lsg_stream_copy": (c_int, [c_int, c_size_t, c_int, POINTER(c_double)]),
    "lsg_set_device": (c_int, [c_int]),
    "lsg_device_name": (c_int, [c_int, c_char_p, c_size_t]),
    "lsg_device_sync": (c_int, [c_int]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads liblsg.so once; raises LsgError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LsgError(-8, "load", f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                                   f"make -C galoisgpu_amd/csrc")
    l = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    _lib = l
    return l


def last_error() -> str:
    msg = lib().lsg_last_error()
    return msg.decode() if msg else ""


def check(rc: int, where: str) -> None:
    if rc != LSG_OK:
        raise LsgError(rc, where, last_error())
