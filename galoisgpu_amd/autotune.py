"""Autotuner for the persistent kernels (replaces the reference's tuning/ scripts).

The reference tunes one compile-time knob, BLKSIZE, with OpenTuner by
rewriting a header, rebuilding and timing three runs (tuning/lsgtuner.py:
68-214, tuning/sssptuner.py:7-16).  Here every knob is a runtime option, so
the tuner times the real kernels directly: for the batched kernel it searches
delta, far-pile mode, heavy-vertex chunk size and block size; for the plain
kernel delta, expansion mode and block size.  Each candidate is timed by the
library's own HIP events (best of `repeats`), results are written as JSON
(``tuning/<name>.json``) and ``tuned_options`` reads them back.

Distances do not depend on any of these knobs; the tuner checks that every
candidate's distances equal the first candidate's before accepting it.

Run on a GPU:  python -m galoisgpu_amd.autotune --graph rmat:22 --nsrc 64 --out galoisgpu_amd/tuning/rmat22_64src.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

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
    return sorted(results, key=lambda r: r["ms"])


def tune_plain(dev, srcs, space=None, repeats: int = 1, log=print) -> List[dict]:
    from .apps import Options, sssp_into
    from .batch import DeviceArray
    V = dev.nnodes
    d = DeviceArray(V * 4, dev.device)
    results, ref = [], {}
    try:
        for cand in grid(space or PLAIN_SPACE):
            total = 0.0
            for s in srcs:
                ms = min(sssp_into(dev, d.ptr, Options(start_node=int(s), **cand))["ms_kernel"]
                         for _ in range(repeats))
                total += ms
                h = d.to_numpy(np.int32, V)
                key = int(s)
                if key not in ref:
                    ref[key] = h
                elif not np.array_equal(h, ref[key]):
                    raise RuntimeError(f"plain candidate {cand} changed the distances of source {s}")
            results.append(dict(cand, ms=total))
            log(f"plain {cand}: {total:.2f} ms for {len(srcs)} sources")
    finally:
        d.close()
    return sorted(results, key=lambda r: r["ms"])


def device_name(device: int = 0) -> str:
    import ctypes
    from ._lib import check, lib
    buf = ctypes.create_string_buffer(256)
    check(lib().lsg_device_name(device, buf, 256), "device_name")
    return buf.value.decode()


def run(graph: str, nsrc: int, out: str, seed: int = 1, plain_sources: int = 8, log=print) -> dict:
    g = make_graph(graph, seed)
    rng = np.random.default_rng(1000 + seed)
    srcs = rng.integers(0, g.nnodes, nsrc).astype(np.uint32)
    dev = g.copy_to_gpu(0)
    t0 = time.time()
    batched = tune_batched(dev, srcs, log=log)
    nontrivial = [s for s in srcs if g.row_start[s + 1] > g.row_start[s]][:plain_sources]
    plain = tune_plain(dev, np.array(nontrivial, dtype=np.uint32), log=log)
    dev.close()
    res = {"graph": graph, "seed": seed, "nsrc": nsrc, "device": device_name(0),
           "created": time.strftime("%Y-%m-%d %H:%M:%S"), "seconds": round(time.time() - t0, 1),
           "batched": {k: v for k, v in batched[0].items() if k != "ms"}, "batched_ms": batched[0]["ms"],
           "plain": {k: v for k, v in plain[0].items() if k != "ms"}, "plain_ms": plain[0]["ms"],
           "plain_sources": [int(s) for s in nontrivial],
           "batched_results": batched, "plain_results": plain}
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    log(f"best batched {res['batched']} {res['batched_ms']:.2f} ms; best plain {res['plain']} "
        f"{res['plain_ms']:.2f} ms -> {out}")
    return res


def tuned_options(name: str, which: str) -> Optional[dict]:
    """Best options of kind `which` ('batched' or 'plain') from tuning/<name>.json, or None."""
    p = os.path.join(HERE, "tuning", f"{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return dict(json.load(f)[which])


if __name__ == "__main__":
    ap = argparse.ArgumentParser(description="tune the persistent SSSP kernels on this GPU")
    ap.add_argument("--graph", default="rmat:22")
    ap.add_argument("--nsrc", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    run(a.graph, a.nsrc, a.out, a.seed, log=lambda *x: print(*x, flush=True))
