"""MI355X-native (gfx950) worklist shortest-path relaxation: LonestarGPU's apps/sssp and apps/bfs path.

The compute path is liblsg.so (hand-written HIP for gfx950 behind the C ABI in
include/lsg.h); this package is its Python host side.
"""
from ._lib import INF, LsgError, lib
from .apps import Options, bfs, output, process_prog_opt, sssp
from .graph import CSRGraph, DeviceGraph

__all__ = ["INF", "LsgError", "lib", "CSRGraph", "DeviceGraph", "Options", "sssp", "bfs", "output",
           "process_prog_opt"]
