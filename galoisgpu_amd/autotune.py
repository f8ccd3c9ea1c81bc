"""Autotuner for the persistent kernels (replaces the reference's tuning/ scripts).

The reference tunes one compile-time knob, BLKSIZE, with OpenTuner by
rewriting a header, rebuilding and timing three runs (tuning/lsgtuner.py:
68-214, tuning/sssptuner.py:7-16, tuning/bfstuner.py:10-23), and keeps one
result set per GPU generation (tuning/configs/{fermi,kepler,lsg2},
tuning/lsgtuner.py:96-115).  Here every knob is a runtime option, so the
tuner times the real kernels directly, and results are kept per device
architecture: ``tuning/<arch>/<workload>.json`` (for example
``tuning/gfx950/rmat22_64src.json``), read back by ``tuned_options``.

Workloads are the benchmark configurations, each with its own search space:

  grid256        plain SSSP, 256^2 grid, source 0, delta 32 (fixed): expansion
                 mode, block size, heavy-chunk size, far-pile mode; and the
                 one-workgroup mode (single_block) over delta
  rmat20         plain SSSP and BFS, RMAT-20 from the max-degree vertex, delta 32
  road2048       plain SSSP, 2048^2 road-like grid, source 0, delta 10000
  rmat22_64src   batched SSSP, RMAT-22, 64 uniform sources: delta, far-pile
                 mode, heavy-chunk size, block size (and the plain kernel on 8
                 of the sources)
  union256       256 RMAT graphs of scales 10-16 in one launch: delta, expansion
                 mode, block size, heavy-chunk size
  rmat24_32src   batched SSSP, RMAT-24, 32 uniform sources

Each candidate is timed by the library's own HIP events (best of `repeats`);
distances do not depend on any knob, and the tuner checks every candidate's
distances against the first candidate's before accepting it.

Run on a GPU:  python -m galoisgpu_amd.autotune --workload all
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time
from typing import Callable, Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING = os.path.join(HERE, "tuning")

BATCH_SPACE = {
    "delta": [32, 256, 1024, 1 << 30],
    "far_mode": [0, 2],
    "heavy_deg": [256, 512, 1024],
    "block_size": [256, 512],
}
PLAIN_SPACE = {
    "delta": [16, 32, 64, 256],
    "expand_mode": [0, 1],
    "heavy_deg": [256, 512, 4096],
    "block_size": [256, 512],
}
# Single-source configurations with a fixed delta: launch knobs only.
PLAIN_LAUNCH_SPACE = {
    "expand_mode": [0, 1],
    "block_size": [256, 512],
    "heavy_deg": [512, 4096],
    "far_mode": [0, 2],
}
# The one-workgroup mode of small high-diameter graphs: its rounds are cheap,
# so delta trades rounds against re-expanded vertices.
SINGLE_SPACE = {
    "delta": [32, 64, 128, 256, 512, 1024],
}
BFS_SPACE = {
    "expand_mode": [0, 1],
    "block_size": [256, 512, 1024],
    "heavy_deg": [256, 512, 4096],
}
UNION_SPACE = {
    "delta": [32, 256, 1 << 30],
    "expand_mode": [0, 1],
    "block_size": [256, 512],
    "heavy_deg": [512, 4096],
}


def grid(space: Dict[str, list]) -> List[dict]:
    keys = list(space)
    return [dict(zip(keys, vals)) for vals in itertools.product(*(space[k] for k in keys))]


def make_graph(spec: str, seed: int = 1):
    """'rmat:22' | 'grid:256' | 'road:2048' -> CSRGraph (the generators used by the benchmarks)."""
    from .graph import CSRGraph
    kind, size = spec.split(":")
    n = int(size)
    if kind == "rmat":
        return CSRGraph.rmat(n, 16, 0.57, 0.19, 0.19, 1, 255, seed=seed)
    if kind == "grid":
        return CSRGraph.grid(n, n, 1, 100, seed=seed)
    if kind == "road":
        return CSRGraph.roadlike(n, n, 0.10, 0.001, 1, 10000, seed=seed)
    raise ValueError(f"unknown graph spec {spec}")


def _sorted(results: List[dict]) -> List[dict]:
    return sorted(results, key=lambda r: r["ms"])


def tune_batched(dev, srcs, space=None, repeats: int = 2, log=print) -> List[dict]:
    from .apps import Options
    from .batch import DeviceArray, sssp_multi_into
    V, S = dev.nnodes, len(srcs)
    D = DeviceArray(V * S * 4, dev.device)
    results, ref = [], None
    try:
        for cand in grid(space or BATCH_SPACE):
            o = Options(**cand)
            ms = min(sssp_multi_into(dev, srcs, o, D)["ms_kernel"] for _ in range(repeats))
            h = D.to_numpy(np.int32, (V, S))
            if ref is None:
                ref = h
            elif not np.array_equal(h, ref):
                raise RuntimeError(f"batched candidate {cand} changed the distances")
            results.append(dict(cand, ms=ms))
            log(f"batched {cand}: {ms:.2f} ms")
    finally:
        D.close()
    return _sorted(results)


def tune_single(dev, srcs, space, fixed: Optional[dict] = None, bfs: bool = False, repeats: int = 1,
                log=print) -> List[dict]:
    """Single-source kernel (SSSP or BFS): total of the best-of-`repeats` times
    over `srcs` for every candidate of `space` (merged with `fixed`)."""
    from .apps import Options, bfs_into, sssp_into
    from .batch import DeviceArray
    run = bfs_into if bfs else sssp_into
    V = dev.nnodes
    d = DeviceArray(V * 4, dev.device)
    results, ref = [], {}
    try:
        for cand in grid(space):
            kw = dict(fixed or {})
            kw.update(cand)
            total = 0.0
            for s in srcs:
                total += min(run(dev, d.ptr, Options(start_node=int(s), **kw))["ms_kernel"] for _ in range(repeats))
                h = d.to_numpy(np.int32, V)
                key = int(s)
                if key not in ref:
                    ref[key] = h
                elif not np.array_equal(h, ref[key]):
                    raise RuntimeError(f"candidate {kw} changed the distances of source {s}")
            results.append(dict(cand, ms=total))
            log(f"{'bfs' if bfs else 'sssp'} {kw}: {total:.3f} ms for {len(srcs)} source(s)")
    finally:
        d.close()
    return _sorted(results)


def tune_plain(dev, srcs, space=None, repeats: int = 1, log=print) -> List[dict]:
    return tune_single(dev, srcs, space or PLAIN_SPACE, repeats=repeats, log=log)


def tune_union(devs, srcs, space=None, repeats: int = 2, log=print) -> List[dict]:
    from .apps import Options
    from .batch import sssp_graphs
    results, ref = [], None
    for cand in grid(space or UNION_SPACE):
        o = Options(**cand)
        best, outs = None, None
        for _ in range(repeats):
            outs, st = sssp_graphs(devs, srcs, o)
            best = st["ms_kernel"] if best is None else min(best, st["ms_kernel"])
        h = np.concatenate(outs)
        if ref is None:
            ref = h
        elif not np.array_equal(h, ref):
            raise RuntimeError(f"union candidate {cand} changed the distances")
        results.append(dict(cand, ms=best))
        log(f"union {cand}: {best:.3f} ms")
    return _sorted(results)


def device_name(device: int = 0) -> str:
    import ctypes
    from ._lib import check, lib
    buf = ctypes.create_string_buffer(256)
    check(lib().lsg_device_name(device, buf, 256), "device_name")
    return buf.value.decode()


def device_arch(device: int = 0) -> str:
    """'gfx950' from a device name such as 'AMD ... (gfx950:sramecc+:xnack-, 256 CUs)'."""
    name = device_name(device)
    if "(" in name:
        return name.rsplit("(", 1)[1].split(":", 1)[0].split(",", 1)[0].strip()
    return "unknown"


def _best(results: List[dict]) -> dict:
    return {k: v for k, v in results[0].items() if k != "ms"}


def _union_graphs(n: int = 256, seed: int = 44):
    from .graph import CSRGraph
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 17, n)
    graphs = [CSRGraph.rmat(int(s), 8, 0.57, 0.19, 0.19, 1, 255, seed=100 + i) for i, s in enumerate(scales)]
    return graphs, [g.max_degree_vertex() for g in graphs]


def _tune_workload(name: str, log) -> dict:
    """Runs the search of one workload; returns its result record."""
    res: dict = {"workload": name}
    if name in ("rmat22_64src", "rmat24_32src"):
        # the sources bench.py and the C5 parity test use
        scale, nsrc, rng_seed = (22, 64, 1001) if name == "rmat22_64src" else (24, 32, 1024)
        g = make_graph(f"rmat:{scale}")
        srcs = np.random.default_rng(rng_seed).integers(0, g.nnodes, nsrc).astype(np.uint32)
        dev = g.copy_to_gpu(0)
        batched = tune_batched(dev, srcs, log=log)
        res.update(batched=_best(batched), batched_ms=batched[0]["ms"], batched_results=batched)
        if scale == 22:
            live = [s for s in srcs if g.row_start[s + 1] > g.row_start[s]][:8]
            plain = tune_plain(dev, np.array(live, dtype=np.uint32), log=log)
            res.update(plain=_best(plain), plain_ms=plain[0]["ms"], plain_sources=[int(s) for s in live],
                       plain_results=plain)
        dev.close()
    elif name in ("grid256", "road2048", "rmat20"):
        spec, delta = {"grid256": ("grid:256", 32), "road2048": ("road:2048", 10000), "rmat20": ("rmat:20", 32)}[name]
        g = make_graph(spec)
        src = g.max_degree_vertex() if name == "rmat20" else 0
        dev = g.copy_to_gpu(0)
        plain = tune_single(dev, [src], PLAIN_LAUNCH_SPACE, fixed={"delta": delta}, repeats=3, log=log)
        res.update(plain=dict(_best(plain), delta=delta), plain_ms=plain[0]["ms"], plain_results=plain,
                   source=int(src))
        if name == "rmat20":
            bfs = tune_single(dev, [src], BFS_SPACE, bfs=True, repeats=3, log=log)
            res.update(bfs=_best(bfs), bfs_ms=bfs[0]["ms"], bfs_results=bfs)
        if name == "grid256":
            single = tune_single(dev, [src], SINGLE_SPACE, fixed={"single_block": True}, repeats=3, log=log)
            res.update(single=dict(_best(single), single_block=True), single_ms=single[0]["ms"],
                       single_results=single)
        dev.close()
    elif name == "union256":
        graphs, srcs = _union_graphs()
        devs = [g.copy_to_gpu(0) for g in graphs]
        union = tune_union(devs, srcs, log=log)
        res.update(union=_best(union), union_ms=union[0]["ms"], union_results=union)
        for d in devs:
            d.close()
    else:
        raise ValueError(f"unknown workload {name}")
    return res


WORKLOADS = ["grid256", "rmat20", "road2048", "rmat22_64src", "union256", "rmat24_32src"]


def run(workload: str, out: Optional[str] = None, log=print, outdir: Optional[str] = None) -> dict:
    """Tunes one workload on device 0 and writes <outdir>/<arch>/<workload>.json
    (outdir defaults to the package's tuning/), or `out`."""
    t0 = time.time()
    res = _tune_workload(workload, log)
    res.update(device=device_name(0), arch=device_arch(0), created=time.strftime("%Y-%m-%d %H:%M:%S"),
               seconds=round(time.time() - t0, 1))
    path = out or os.path.join(outdir or TUNING, res["arch"], f"{workload}.json")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    log(f"{workload}: {({k: res[k] for k in res if k in ('batched', 'plain', 'bfs', 'union', 'single')})} -> {path}")
    return res


def tuned_options(name: str, which: str, device: int = 0, root: Optional[str] = None) -> Optional[dict]:
    """Best options of kind `which` ('batched', 'plain', 'bfs', 'union' or 'single') for
    workload `name` on this device's architecture: tuning/<arch>/<name>.json,
    else tuning/<name>.json; None when neither holds it."""
    base = root or TUNING
    paths = []
    try:
        paths.append(os.path.join(base, device_arch(device), f"{name}.json"))
    except Exception:
        pass
    paths.append(os.path.join(base, f"{name}.json"))
    for p in paths:
        if os.path.exists(p):
            with open(p) as f:
                rec = json.load(f)
            if which in rec:
                return dict(rec[which])
    return None


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description="tune the persistent SSSP / BFS kernels on this GPU")
    ap.add_argument("--workload", default="rmat22_64src", help=f"one of {WORKLOADS} or 'all'")
    ap.add_argument("--out", help="output file (default <outdir>/<arch>/<workload>.json)")
    ap.add_argument("--outdir", help="directory of the per-architecture result sets (default: the package's tuning/)")
    a = ap.parse_args()
    names = WORKLOADS if a.workload == "all" else [a.workload]
    for n in names:
        run(n, a.out if len(names) == 1 else None, log=lambda *x: print(*x, flush=True), outdir=a.outdir)
