"""Headline benchmark: SSSP GTEPS on RMAT-22, batched (+ learned schedule) vs the plain kernel.

Workload (one "step"): SSSP from 64 sources sampled uniformly from a seeded
RMAT scale-22 graph (edge factor 16, a/b/c/d = .57/.19/.19/.05, self loops and
duplicates removed, weights uniform in [1,255]), delta 32.  The batched path
runs the 64 sources in one persistent launch, replayed as a captured hipGraph;
the plain path runs them one after another through the single-source kernel.
Distances of the two are compared bit for bit before anything is timed.

Besides the headline line the JSON carries the learned schedule against the
autotuned fixed one (median and spread of per-launch kernel times over the
same number of launches each), and a with-allocation step: the library's
cached device buffers are freed first, so that the step allocates them as the
reference's gg_main allocates its worklists on every call.

GTEPS = sum over sources of E_r (out-degree sum of reached vertices) / time.
Algorithmic bytes per source B = 12 E_r + 16 V_r; roofline.achieved = sum B /
mean kernel time of the dominant launch (HIP events on the library stream).

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--nsrc", type=int, default=64)
    ap.add_argument("--delta", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain kernel")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="at most this many of the sources with out-edges in the CPU baseline sample")
    ap.add_argument("--sched", choices=["auto", "fixed", "learned"], default="auto")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(sources would be split round-robin across GPUs with no exchange).", file=sys.stderr)
        return 2

    import numpy as np

    from galoisgpu_amd import CSRGraph, Options
    from galoisgpu_amd.apps import sssp_into
    from galoisgpu_amd.batch import DeviceArray, sssp_multi_into
    from galoisgpu_amd import sched as lsched

    t0 = time.time()
    g = CSRGraph.rmat(args.scale, 16, 0.57, 0.19, 0.19, 1, 255, seed=args.seed)
    V, E = g.nnodes, g.nedges
    rng = np.random.default_rng(1000 + args.seed)
    srcs = rng.integers(0, V, args.nsrc).astype(np.uint32)
    log(f"rmat{args.scale}: V={V} E={E} generated in {time.time() - t0:.1f}s; {args.nsrc} sources")
    dev = g.copy_to_gpu(0)
    row = g.row_start.astype(np.int64)
    deg = np.diff(row)

    from galoisgpu_amd import autotune
    tname = f"rmat{args.scale}_{args.nsrc}src"
    tuned_b = autotune.tuned_options(tname, "batched") or {}
    tuned_p = autotune.tuned_options(tname, "plain") or {}
    launch_b = {k: v for k, v in tuned_b.items() if k in ("heavy_deg", "block_size")}
    launch_p = {k: v for k, v in tuned_p.items() if k in ("expand_mode", "heavy_deg", "block_size")}
    use_learned = False
    model_path = lsched.default_model_path("batched")
    if args.sched == "learned" or (args.sched == "auto" and os.path.exists(model_path)):
        lsched.load_model(model_path, device=0, kind="batched")
        use_learned = True
    # fixed batched schedule: the autotuner's choice (or the plain delta if never tuned)
    fixed_kw = dict(tuned_b) if tuned_b else dict(delta=args.delta)
    bopts = Options(delta=args.delta, sched=1, graph_launch=True, **launch_b) if use_learned else \
        Options(graph_launch=True, **fixed_kw)
    popts = dict(delta=args.delta, **launch_p)
    log(f"batched options: {bopts}; plain options: {popts}")

    # ---- correctness gate (outside the timed region) ------------------------
    D = DeviceArray(V * args.nsrc * 4, 0)
    st_b = sssp_multi_into(dev, srcs, bopts, D)
    Dh = D.to_numpy(np.int32, (V, args.nsrc))
    plain_dist = DeviceArray(V * 4, 0)
    er = np.zeros(args.nsrc, dtype=np.int64)
    vr = np.zeros(args.nsrc, dtype=np.int64)
    for k, s in enumerate(srcs):
        sssp_into(dev, plain_dist.ptr, Options(start_node=int(s), **popts))
        d = plain_dist.to_numpy(np.int32, V)
        if not np.array_equal(d, Dh[:, k]):
            bad = int(np.count_nonzero(d != Dh[:, k]))
            log(f"MISMATCH source {k} ({s}): {bad} distances differ between batched and plain")
            return 1
        fin = d != 0x7FFFFFFF
        vr[k] = int(fin.sum())
        er[k] = int(deg[fin].sum())
    log(f"batched == plain for all {args.nsrc} sources; sum E_r = {er.sum()}, sources with E_r>0: "
        f"{int((er > 0).sum())}; learned={use_learned}; batched stats {st_b}")
    bytes_per_step = float((12 * er + 16 * vr).sum())
    edges_per_step = float(er.sum())

    # ---- batched: warmup + timed --------------------------------------------
    for _ in range(args.warmup):
        sssp_multi_into(dev, srcs, bopts, D)
    kms = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        st = sssp_multi_into(dev, srcs, bopts, D)
        kms.append(st["ms_kernel"])
    wall = time.perf_counter() - t1
    ms_step = wall * 1000.0 / args.steps
    gteps = edges_per_step * args.steps / wall / 1e9
    kernel_ms = float(np.mean(kms))
    achieved = bytes_per_step / (kernel_ms / 1000.0) / 1e9
    log(f"batched: {ms_step:.3f} ms/step wall, kernel {kernel_ms:.3f} ms, {gteps:.2f} GTEPS")

    # ---- the autotuned fixed-schedule batched run, for comparison -----------
    # Both schedules are timed over the same number of launches, alternating,
    # so that drift affects them alike; median and spread of the kernel times.
    fixed = None
    if use_learned:
        fo = Options(graph_launch=True, **fixed_kw)
        sssp_multi_into(dev, srcs, fo, D)
        assert np.array_equal(D.to_numpy(np.int32, (V, args.nsrc)), Dh), "fixed schedule changed distances"
        n = max(5, args.steps)
        lk, fk = [], []
        for _ in range(n):
            lk.append(sssp_multi_into(dev, srcs, bopts, D)["ms_kernel"])
            fk.append(sssp_multi_into(dev, srcs, fo, D)["ms_kernel"])
        lk, fk = np.array(lk), np.array(fk)
        fixed = {"options": fixed_kw, "launches": n,
                 "learned_kernel_ms_median": round(float(np.median(lk)), 3),
                 "learned_kernel_ms_min_max": [round(float(lk.min()), 3), round(float(lk.max()), 3)],
                 "fixed_kernel_ms_median": round(float(np.median(fk)), 3),
                 "fixed_kernel_ms_min_max": [round(float(fk.min()), 3), round(float(fk.max()), 3)],
                 "learned_over_fixed": round(float(np.median(fk) / np.median(lk)), 4)}
        log(f"learned vs fixed {fixed_kw}: median {fixed['learned_kernel_ms_median']} vs "
            f"{fixed['fixed_kernel_ms_median']} ms over {n} launches each")

    # ---- with allocation: the step after the cached buffers are freed --------
    from galoisgpu_amd._lib import check as _check, lib as _lib
    import dataclasses
    wopts = dataclasses.replace(bopts, graph_launch=False)  # allocation, not graph capture
    walloc = []
    for _ in range(3):
        _check(_lib().lsg_workspace_release(0), "workspace_release")
        walloc.append(sssp_multi_into(dev, srcs, wopts, D)["ms_total"])
    with_alloc = {"ms_per_step_median": round(float(np.median(walloc)), 3),
                  "gteps": round(edges_per_step / (float(np.median(walloc)) / 1000.0) / 1e9, 3),
                  "note": "host time of one batched call made right after lsg_workspace_release, so that it "
                          "allocates its queues, masks and distance matrix as the reference's gg_main does"}
    log(f"with allocation: {with_alloc['ms_per_step_median']} ms/step")

    # ---- plain kernel on the same GPU and sources ---------------------------
    plain = None
    if not args.no_plain:
        for s in srcs[: min(4, len(srcs))]:
            sssp_into(dev, plain_dist.ptr, Options(start_node=int(s), **popts))
        pk = 0.0
        t2 = time.perf_counter()
        for _ in range(args.steps):
            for s in srcs:
                st = sssp_into(dev, plain_dist.ptr, Options(start_node=int(s), **popts))
                pk += st["ms_kernel"]
        pwall = time.perf_counter() - t2
        plain = {"gteps": edges_per_step * args.steps / pwall / 1e9, "ms_per_step": pwall * 1000.0 / args.steps,
                 "kernel_ms_per_step": pk / args.steps}
        log(f"plain: {plain['ms_per_step']:.3f} ms/step wall, {plain['gteps']:.2f} GTEPS")

    # ---- CPU baseline: the oracle's restatement of the reference near/far ---
    # One source per host thread (the C restatement runs with the GIL released),
    # at most 16 threads: the host share of one GPU.
    cpu = None
    if not args.no_cpu:
        import oracle
        from concurrent.futures import ThreadPoolExecutor
        sample = [int(s) for s, e in zip(srcs, er) if e > 0][: args.cpu_sample]
        threads = max(1, min(len(sample), os.cpu_count() or 1, 16))
        w = g.edge_data

        def one(s):
            d, _ = oracle.nearfar(g.row_start, g.edge_dst, w, s, args.delta)
            k = int(np.flatnonzero(srcs == s)[0])
            return bool(np.array_equal(d, Dh[:, k])), int(er[k])

        tc = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            res = list(ex.map(one, sample))
        tcpu = time.perf_counter() - tc
        assert all(ok for ok, _ in res), "CPU restatement disagrees with the GPU"
        ce = sum(e for _, e in res)
        cpu = {"value": ce / tcpu / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} of the {args.nsrc} sources (those with E_r>0), one per thread on {threads} "
                         f"of the host's {os.cpu_count()} logical cores, serial near/far restatement of apps/sssp "
                         f"(oracle/oracle.c), {tcpu:.1f}s wall"}
        log(f"cpu: {cpu['value']:.4f} GTEPS on {len(sample)} sources with {threads} threads in {tcpu:.1f}s")

    traffic = lsched.profiled_traffic(ROOT, "sssp_multi_rmat22")
    out = {
        "metric": "SSSP GTEPS on RMAT-22, learned+batched vs plain kernel, bit-exact dists",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"rmat{args.scale}_ef16_{args.nsrc}src_delta{args.delta}", "scale": args.scale,
                   "edge_factor": 16, "nsrc": args.nsrc, "delta": args.delta, "weights": "U[1,255]",
                   "seed": args.seed, "nnodes": V, "nedges": E, "sum_E_r": int(er.sum()),
                   "scheduler": "learned" if use_learned else "fixed", "batched_options": str(bopts),
                   "plain_options": popts, "parallelism": "single GPU"},
        "plain": plain,
        "learned_vs_fixed_schedule": fixed,
        "with_allocation": with_alloc,
        "speedup_vs_plain": round(gteps / plain["gteps"], 3) if plain else None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    for o in (D, plain_dist):
        o.close()
    dev.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
