"""Measures the benchmark configurations other than the headline one on one GPU.

For each configuration: the plain single-source kernel (the drop-in) against
the learned and/or batched form, every distance array checked bit for bit
(against the oracle's Dijkstra where it finishes in seconds, else against the
plain kernel), GTEPS = sum E_r / time, algorithmic bytes 12 E_r + 16 V_r per
traversal over the kernel time as a fraction of the 8 TB/s HBM peak, and the
oracle's serial near/far restatement timed beside it on the host.

  C1  256x256 grid, w in [1,100], source 0, delta 32; also the one-workgroup
      mode (single_block) at delta 32 and at the autotuner's delta, and BFS
  C2  RMAT-20 ef16, w in [1,255], max-degree source, delta 32 (SSSP and BFS)
  C3  2048x2048 road-like grid (10% edges removed, 0.1% shortcuts), w in
      [1,10000], source 0, delta 10000: plain fixed delta vs learned schedule,
      and the first rounds in one workgroup (single_block)
  C4B 256 RMAT graphs, scales uniform in [10,16], ef 8, max-degree source each:
      one union launch vs the plain kernel graph by graph
  C5  RMAT-24 ef16, 32 sources: batched (+ learned) vs plain per source
  CC  connected components (apps/cc) of the RMAT-22 and 2048x2048 road-like
      graphs against the oracle's union-find
  PR  PageRank (apps/pr) of the same graphs against the oracle's double fixpoint

The headline RMAT-22 x 64-source configuration is bench.py.  This file uses
the oracle (test infrastructure) as the checker and CPU baseline, which is why
it lives under tests/.  Usage:

  python tests/measure_configs.py [--only C1,C3] [--repeats 5]

Prints one JSON object per configuration on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from galoisgpu_amd import CSRGraph, Options  # noqa: E402
from galoisgpu_amd import sched as lsched  # noqa: E402
from galoisgpu_amd.apps import bfs_do_into, bfs_into, cc_into, pagerank_into, sssp_into  # noqa: E402
from galoisgpu_amd.batch import DeviceArray, sssp_graphs, sssp_multi_into  # noqa: E402

HBM_PEAK_GBS = 8000.0
INF = 0x7FFFFFFF


def log(*a):
    print("[measure]", *a, file=sys.stderr, flush=True)


def reach(g: CSRGraph, d: np.ndarray):
    fin = d != INF
    deg = np.diff(g.row_start.astype(np.int64))
    return int(deg[fin].sum()), int(fin.sum())


def rate(er: int, vr: int, ms: float, bfs: bool = False) -> dict:
    b = (8 if bfs else 12) * er + 16 * vr
    return {"ms": round(ms, 4), "gteps": round(er / ms / 1e6, 4),
            "hbm_frac": round(b / ms / 1e6 / HBM_PEAK_GBS, 5)}


def timed_plain(dev, buf, opts: Options, repeats: int, bfs: bool = False):
    """Median kernel ms (HIP events) of `repeats` runs after one warm-up."""
    run = bfs_into if bfs else sssp_into
    run(dev, buf.ptr, opts)
    ms = sorted(run(dev, buf.ptr, opts)["ms_kernel"] for _ in range(repeats))
    return ms[len(ms) // 2]


def cpu_nearfar(g: CSRGraph, srcs, delta: int):
    t = time.perf_counter()
    er = 0
    for s in srcs:
        d, _ = oracle.nearfar(g.row_start, g.edge_dst, g.edge_data, int(s), delta)
        er += reach(g, d)[0]
    dt = time.perf_counter() - t
    return {"gteps": round(er / dt / 1e9, 5), "seconds": round(dt, 2), "sources": len(srcs), "cores": 1,
            "kind": "port (oracle/oracle.c near/far restatement)"}


def have_model(kind: str) -> bool:
    p = lsched.default_model_path(kind)
    if os.path.exists(p):
        lsched.load_model(p, kind=kind)
        return True
    return False


def c1(repeats: int) -> dict:
    g = CSRGraph.grid(256, 256, 1, 100, seed=1)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    dev = g.copy_to_gpu()
    buf = DeviceArray(g.nnodes * 4)
    er, vr = reach(g, ref)
    out = {"config": "C1", "graph": "grid256 w[1,100]", "V": g.nnodes, "E": g.nedges, "E_r": er}
    from galoisgpu_amd import autotune
    tuned = autotune.tuned_options("grid256", "single") or {}
    tuned_delta = int(tuned.get("delta", 256))
    variants = [("plain", {})]
    if have_model("plain"):
        variants.append(("learned", {"sched": 1}))
    variants += [("single_block", {"single_block": True}), ("single_block_tuned", {"single_block": True, "delta": tuned_delta})]
    for name, kw in variants:
        kw = dict({"delta": 32}, **kw)
        o = Options(start_node=0, **kw)
        st = sssp_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), ref), f"C1 {name} distances differ from Dijkstra"
        out[name] = dict(rate(er, vr, timed_plain(dev, buf, o, repeats)), phases=st["iterations"],
                         rounds=st["near_iters"], delta=kw["delta"], blocks=st["grid_blocks"])
    lref, _ = oracle.bfs(g.row_start, g.edge_dst, 0)
    for name, kw in (("bfs", {}), ("bfs_single_block", {"single_block": True})):
        o = Options(start_node=0, **kw)
        bfs_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), lref), f"C1 {name} levels differ from BFS"
        out[name] = rate(er, vr, timed_plain(dev, buf, o, repeats, bfs=True), bfs=True)
    out["cpu"] = cpu_nearfar(g, [0], 32)
    buf.close()
    dev.close()
    return out


def c2(repeats: int) -> dict:
    g = CSRGraph.rmat(20, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    src = g.max_degree_vertex()
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, src)
    lref, _ = oracle.bfs(g.row_start, g.edge_dst, src)
    dev = g.copy_to_gpu()
    buf = DeviceArray(g.nnodes * 4)
    er, vr = reach(g, ref)
    out = {"config": "C2", "graph": "rmat20 ef16 w[1,255]", "V": g.nnodes, "E": g.nedges, "E_r": er, "source": src}
    for name, kw in (("twb", {"expand_mode": 0}), ("lb", {"expand_mode": 1})):
        o = Options(delta=32, start_node=src, **kw)
        sssp_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), ref), f"C2 SSSP {name} differs"
        out[f"sssp_{name}"] = rate(er, vr, timed_plain(dev, buf, o, repeats))
        bo = Options(start_node=src, **kw)
        bfs_into(dev, buf.ptr, bo)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), lref), f"C2 BFS {name} differs"
        out[f"bfs_{name}"] = rate(er, vr, timed_plain(dev, buf, bo, repeats, bfs=True), bfs=True)
    if have_model("plain"):
        o = Options(delta=32, start_node=src, expand_mode=1, sched=1)
        st = sssp_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), ref), "C2 SSSP learned differs"
        out["sssp_learned"] = dict(rate(er, vr, timed_plain(dev, buf, o, repeats)), phases=st["iterations"],
                                   rounds=st["near_iters"])
    # direction-optimising BFS; the reverse CSR is built once per graph, timed apart
    t = time.perf_counter()
    rev = dev.transpose()
    t_rev = (time.perf_counter() - t) * 1e3
    bfs_do_into(dev, rev, buf.ptr, src)
    assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), lref), "C2 BFS (direction-optimising) differs"
    ms = sorted(bfs_do_into(dev, rev, buf.ptr, src)["ms_kernel"] for _ in range(repeats))[repeats // 2]
    out["bfs_do"] = dict(rate(er, vr, ms, bfs=True), transpose_ms=round(t_rev, 3))
    rev.close()
    out["cpu"] = cpu_nearfar(g, [src], 32)
    buf.close()
    dev.close()
    return out


def c3(repeats: int) -> dict:
    g = CSRGraph.roadlike(2048, 2048, 0.10, 0.001, 1, 10000, seed=1)
    ref = oracle.dijkstra(g.row_start, g.edge_dst, g.edge_data, 0)
    dev = g.copy_to_gpu()
    buf = DeviceArray(g.nnodes * 4)
    er, vr = reach(g, ref)
    out = {"config": "C3", "graph": "road2048 w[1,10000]", "V": g.nnodes, "E": g.nedges, "E_r": er, "delta": 10000}
    variants = [("plain", {"expand_mode": 1}), ("plain_twb", {"expand_mode": 0})]
    if have_model("plain"):
        variants += [("learned", {"sched": 1, "expand_mode": 1}), ("learned_twb", {"sched": 1, "expand_mode": 0})]
    # the first rounds in one workgroup, the rest in the persistent kernel once the piles overflow
    variants += [("single_block", {"single_block": True})]
    for name, kw in variants:
        o = Options(delta=10000, start_node=0, **kw)
        st = sssp_into(dev, buf.ptr, o)
        assert np.array_equal(buf.to_numpy(np.int32, g.nnodes), ref), f"C3 {name} differs"
        out[name] = dict(rate(er, vr, timed_plain(dev, buf, o, repeats)), phases=st["iterations"],
                         rounds=st["near_iters"])
    if "learned" in out:
        out["learned_speedup"] = round(out["plain"]["ms"] / out["learned"]["ms"], 3)
    out["cpu"] = cpu_nearfar(g, [0], 10000)
    buf.close()
    dev.close()
    return out


def c4b(repeats: int) -> dict:
    rng = np.random.default_rng(44)
    scales = rng.integers(10, 17, 256)
    graphs = [CSRGraph.rmat(int(s), 8, 0.57, 0.19, 0.19, 1, 255, seed=100 + i) for i, s in enumerate(scales)]
    srcs = [g.max_degree_vertex() for g in graphs]
    devs = [g.copy_to_gpu() for g in graphs]
    bufs = [DeviceArray(g.nnodes * 4) for g in graphs]
    er = vr = 0
    refs = []
    t_plain = 0.0
    for g, d, b, s in zip(graphs, devs, bufs, srcs):
        o = Options(delta=32, start_node=s)
        sssp_into(d, b.ptr, o)
        refs.append(b.to_numpy(np.int32, g.nnodes))
        e, v = reach(g, refs[-1])
        er += e
        vr += v
    for _ in range(repeats):
        for d, b, s in zip(devs, bufs, srcs):
            t_plain += sssp_into(d, b.ptr, Options(delta=32, start_node=s))["ms_kernel"]
    t_plain /= repeats
    out = {"config": "C4B", "graphs": 256, "scales": [int(x) for x in np.bincount(scales, minlength=17)[10:]],
           "E_total": int(sum(g.nedges for g in graphs)), "E_r": er, "plain_one_by_one": rate(er, vr, t_plain)}
    variants = [("union", Options(delta=32)), ("union_graph_launch", Options(delta=32, graph_launch=True))]
    if have_model("plain"):
        variants.append(("union_learned", Options(delta=32, sched=1)))
    # one workgroup per graph (single_block), at the reference delta and wider ones
    variants += [(f"one_workgroup_per_graph_d{d}", Options(delta=d, single_block=True)) for d in (32, 128, 512)]
    for name, o in variants:
        dists, st = sssp_graphs(devs, srcs, o)
        for k in range(len(graphs)):
            assert np.array_equal(dists[k], refs[k]), f"C4B {name} graph {k} differs"
        ms = sorted(sssp_graphs(devs, srcs, o)[1]["ms_kernel"] for _ in range(repeats))[repeats // 2]
        out[name] = dict(rate(er, vr, ms), blocks=st["grid_blocks"], rounds=st["rounds"])
    out["speedup_union_vs_plain"] = round(t_plain / out["union"]["ms"], 3)
    out["cpu"] = cpu_nearfar(graphs[int(np.argmax(scales))], [srcs[int(np.argmax(scales))]], 32)
    for b in bufs:
        b.close()
    for d in devs:
        d.close()
    return out


def c5(repeats: int) -> dict:
    g = CSRGraph.rmat(24, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)
    V = g.nnodes
    srcs = np.random.default_rng(1024).integers(0, V, 32).astype(np.uint32)
    dev = g.copy_to_gpu()
    D = DeviceArray(V * 32 * 4)
    buf = DeviceArray(V * 4)
    learned = have_model("batched")
    bo = Options(delta=32, sched=1, heavy_deg=256, block_size=256) if learned else \
        Options(delta=1 << 30, far_mode=0, heavy_deg=256, block_size=256)
    sssp_multi_into(dev, srcs, bo, D)
    Dh = D.to_numpy(np.int32, (V, 32))
    er = vr = 0
    t_plain = 0.0
    for k, s in enumerate(srcs):
        o = Options(delta=32, start_node=int(s), expand_mode=1)
        sssp_into(dev, buf.ptr, o)
        d = buf.to_numpy(np.int32, V)
        assert np.array_equal(d, Dh[:, k]), f"C5 source {k} differs between batched and plain"
        e, v = reach(g, d)
        er += e
        vr += v
        t_plain += sssp_into(dev, buf.ptr, o)["ms_kernel"]
    ms = sorted(sssp_multi_into(dev, srcs, bo, D)["ms_kernel"] for _ in range(repeats))[repeats // 2]
    out = {"config": "C5", "graph": "rmat24 ef16 w[1,255]", "V": V, "E": g.nedges, "sources": 32, "E_r": er,
           "plain": rate(er, vr, t_plain), "batched": dict(rate(er, vr, ms), learned=learned, options=str(bo)),
           "speedup": round(t_plain / ms, 3)}
    if learned:
        # the learned schedule against the fixed Bellman-Ford-round schedule, median of >= 5 launches each
        fixed = Options(delta=1 << 30, far_mode=0, heavy_deg=256, block_size=256)
        n = max(5, repeats)
        ml = sorted(sssp_multi_into(dev, srcs, bo, D)["ms_kernel"] for _ in range(n))
        mf = sorted(sssp_multi_into(dev, srcs, fixed, D)["ms_kernel"] for _ in range(n))
        assert np.array_equal(D.to_numpy(np.int32, (V, 32)), Dh), "C5 fixed schedule differs"
        out["learned_vs_fixed"] = {"launches": n, "learned_ms_median": round(ml[n // 2], 3),
                                   "fixed_ms_median": round(mf[n // 2], 3), "fixed": str(fixed),
                                   "learned_over_fixed": round(mf[n // 2] / ml[n // 2], 3)}
    nz = [int(s) for s in srcs if g.row_start[s + 1] > g.row_start[s]][:1]
    out["cpu"] = cpu_nearfar(g, nz, 32)
    D.close()
    buf.close()
    dev.close()
    return out


def cc_graphs(repeats: int) -> dict:
    """Connected components (apps/cc) on the RMAT-22 and road-like inputs of the
    SSSP configurations; the CPU side is the oracle's union-find, one core."""
    out = {"config": "CC"}
    for name, g in (("rmat22", CSRGraph.rmat(22, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)),
                    ("road2048", CSRGraph.roadlike(2048, 2048, 0.10, 0.001, 1, 10000, seed=1))):
        t = time.perf_counter()
        ref, k = oracle.cc(g.row_start, g.edge_dst)
        t_cpu = time.perf_counter() - t
        dev = g.copy_to_gpu()
        lab = DeviceArray(g.nnodes * 4)
        st = cc_into(dev, lab.ptr)
        assert np.array_equal(lab.to_numpy(np.int32, g.nnodes), ref), f"CC {name} differs"
        ms = sorted(cc_into(dev, lab.ptr)["ms_kernel"] for _ in range(repeats))[repeats // 2]
        out[name] = {"V": g.nnodes, "E": g.nedges, "components": k, "passes": st["iterations"], "ms": round(ms, 4),
                     "gedges_per_s": round(g.nedges / ms / 1e6, 3),
                     "cpu": {"ms": round(t_cpu * 1e3, 2), "cores": 1, "kind": "port (oracle/oracle.c union-find)"}}
        lab.close()
        dev.close()
    return out


def pr_graphs(repeats: int) -> dict:
    """PageRank (apps/pr, alpha 0.85, epsilon 1e-6) on the RMAT-22 and road-like
    inputs: GPU float32 against the oracle's double fixpoint (largest relative
    error reported); the CPU side is the oracle's Jacobi iteration timed per
    iteration over 3 iterations, one core."""
    out = {"config": "PR"}
    for name, g in (("rmat22", CSRGraph.rmat(22, 16, 0.57, 0.19, 0.19, 1, 255, seed=1)),
                    ("road2048", CSRGraph.roadlike(2048, 2048, 0.10, 0.001, 1, 10000, seed=1))):
        dev = g.copy_to_gpu()
        rev = dev.transpose()
        rank = DeviceArray(g.nnodes * 4)
        st = pagerank_into(dev, rev, rank.ptr)
        got = rank.to_numpy(np.float32, g.nnodes).astype(np.float64)
        ms = sorted(pagerank_into(dev, rev, rank.ptr)["ms_kernel"] for _ in range(repeats))[repeats // 2]
        ref, _ = oracle.pagerank(g.row_start, g.edge_dst, tol=1e-9, max_iter=400)
        err = float((np.abs(got - ref) / ref).max())
        assert err <= 1e-4, f"PR {name} relative error {err}"
        t = time.perf_counter()
        _, it = oracle.pagerank(g.row_start, g.edge_dst, max_iter=3)
        cpu_iter = (time.perf_counter() - t) / it
        out[name] = {"V": g.nnodes, "E": g.nedges, "iterations": st["iterations"], "ms": round(ms, 3),
                     "ms_per_iteration": round(ms / st["iterations"], 4), "max_rel_error": err,
                     "cpu": {"ms_per_iteration": round(cpu_iter * 1e3, 2), "cores": 1,
                             "kind": "port (oracle/oracle.c Jacobi, double)"}}
        rank.close()
        rev.close()
        dev.close()
    return out


CONFIGS = {"C1": c1, "C2": c2, "C3": c3, "C4B": c4b, "C5": c5, "CC": cc_graphs, "PR": pr_graphs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=",".join(CONFIGS))
    ap.add_argument("--repeats", type=int, default=5)
    a = ap.parse_args()
    for name in a.only.split(","):
        t = time.time()
        log(f"{name} ...")
        r = CONFIGS[name](a.repeats)
        log(f"{name} done in {time.time() - t:.1f}s")
        print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
