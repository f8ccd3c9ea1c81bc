"""Learned phase scheduler: model files, forced schedules, trace sweeps and training.

A model (layout in csrc/sched_mlp.h) maps 16 frontier features, taken by a
persistent kernel's barrier leader at every phase boundary, to one of 8
actions: a delta multiplier in {1/4, 1, 8, 2^24 (unbounded)} times a mode.
There is one model per kernel kind:

* ``"plain"`` -- the single-source kernel (the drop-in for apps/sssp); the mode
  is the expansion mode, degree-binned (TWB) or tile-flattened (LB), the
  reference's np_schedulers / -l choice (apps/sssp/kernel.cu:10,434-453).
* ``"batched"`` -- the multi-source kernel; the mode is the far-pile mode,
  reference (one shared threshold) or bucket (per-source thresholds).

Training data comes from the kernel itself: for each training instance it is
run under a sweep of forced per-phase schedules, the fastest one is kept, and
the features it recorded at each phase are labelled with the action it took
there (``sweep_instance``).  ``train`` fits the MLP with the HIP training
kernel (lsg_mlp_train).  Whatever the model picks, distances are identical:
the actions only reorder the same u32 relaxations.
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import EXPAND_LB, EXPAND_TWB, SCHED_BATCHED, SCHED_PLAIN, check, lib

HERE = os.path.dirname(os.path.abspath(__file__))

N_FEAT, N_HID, N_ACT, MAX_FORCED = 16, 64, 8, 256
DELTA_MULTS = [0.25, 1.0, 8.0, float(1 << 24)]  # the last one makes a phase plain Bellman-Ford rounds
FAR_REFERENCE, FAR_SPLIT, FAR_BUCKET = 0, 1, 2
KINDS = {"batched": SCHED_BATCHED, "plain": SCHED_PLAIN}
# action id = 4 * (mode index) + delta multiplier index
ACTIONS_BY_KIND = {
    "batched": [(m, far) for far in (FAR_REFERENCE, FAR_BUCKET) for m in DELTA_MULTS],
    "plain": [(m, ex) for ex in (EXPAND_TWB, EXPAND_LB) for m in DELTA_MULTS],
}
ACTIONS = ACTIONS_BY_KIND["batched"]
# The fixed schedule (delta x1 with each kernel's default mode: reference far
# pile for the batched kernel, tile-flattened expansion for the plain one).
FIXED_BY_KIND = {"batched": 1, "plain": 5}
FIXED_ACTION = FIXED_BY_KIND["batched"]

OFF_MEAN = 0
OFF_INVSTD = OFF_MEAN + N_FEAT
OFF_W1 = OFF_INVSTD + N_FEAT
OFF_B1 = OFF_W1 + N_FEAT * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID * N_ACT
OFF_DELTA = OFF_B3 + N_ACT
OFF_FAR = OFF_DELTA + N_ACT
OFF_FORCED = OFF_FAR + N_ACT
MODEL_FLOATS = OFF_FORCED + MAX_FORCED


def model_floats() -> int:
    n = int(lib().lsg_sched_model_floats())
    assert n == MODEL_FLOATS, (n, MODEL_FLOATS)
    return n


def default_model_path(kind: str = "batched") -> str:
    return os.path.join(HERE, "models", f"sched_{kind}.npz")


def new_model(seed: int = 0, kind: str = "batched") -> np.ndarray:
    """Randomly initialised network (He init), action tables, no forced schedule."""
    rng = np.random.default_rng(seed)
    m = np.zeros(MODEL_FLOATS, dtype=np.float32)
    m[OFF_INVSTD:OFF_INVSTD + N_FEAT] = 1.0
    m[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / N_FEAT), N_FEAT * N_HID)
    m[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / N_HID), N_HID * N_HID)
    m[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / N_HID), N_HID * N_ACT)
    for a, (mult, mode) in enumerate(ACTIONS_BY_KIND[kind]):
        m[OFF_DELTA + a] = mult
        m[OFF_FAR + a] = mode
    m[OFF_FORCED:OFF_FORCED + MAX_FORCED] = -1.0
    return m


def forced(model: np.ndarray, actions: Sequence[int]) -> np.ndarray:
    """Copy of `model` whose phase p takes actions[min(p, len-1)] regardless of
    the MLP (the kernel holds the last of the MAX_FORCED entries for later phases)."""
    m = model.copy()
    acts = list(actions) or [FIXED_ACTION]
    for p in range(MAX_FORCED):
        m[OFF_FORCED + p] = acts[min(p, len(acts) - 1)]
    return m


def set_model(model: np.ndarray, device: int = 0, kind: str = "batched") -> None:
    m = np.ascontiguousarray(model, dtype=np.float32)
    check(lib().lsg_sched_set_model(device, KINDS[kind], m.ctypes.data_as(ctypes.c_void_p), len(m)),
          "sched_set_model")


def save_model(path: str, model: np.ndarray, meta: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, model=np.asarray(model, dtype=np.float32), meta=np.array(json.dumps(meta)))


def load_model(path: Optional[str] = None, device: int = 0, kind: Optional[str] = None
               ) -> Tuple[np.ndarray, dict]:
    """Loads a saved model onto `device` for its kind (from the file's metadata
    unless `kind` is given)."""
    z = np.load(path or default_model_path(kind or "batched"), allow_pickle=False)
    model = z["model"].astype(np.float32)
    meta = json.loads(str(z["meta"]))
    if len(model) != MODEL_FLOATS:
        raise ValueError(f"{path}: model has {len(model)} floats, expected {MODEL_FLOATS}")
    kind = kind or meta.get("kind", "batched")
    model[OFF_FORCED:OFF_FORCED + MAX_FORCED] = -1.0  # a stored model never forces phases
    set_model(model, device, kind)
    return model, meta


def infer(model: np.ndarray, X: np.ndarray, device: int = 0) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.float32)
    m = np.ascontiguousarray(model, dtype=np.float32)
    out = np.zeros(len(X), dtype=np.int32)
    check(lib().lsg_mlp_infer(device, m.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                              len(X), out.ctypes.data_as(ctypes.c_void_p)), "mlp_infer")
    return out


def train(X: np.ndarray, y: np.ndarray, w: Optional[np.ndarray] = None, steps: int = 4000, lr: float = 0.02,
          momentum: float = 0.9, seed: int = 1, device: int = 0, model: Optional[np.ndarray] = None
          ) -> Tuple[np.ndarray, float]:
    """Fits the MLP on (X raw features, y actions) with the HIP training kernel."""
    m = np.ascontiguousarray(model if model is not None else new_model(seed), dtype=np.float32).copy()
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.int32)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        wp = w.ctypes.data_as(ctypes.c_void_p)
    loss = ctypes.c_float()
    check(lib().lsg_mlp_train(device, m.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                              y.ctypes.data_as(ctypes.c_void_p), wp, len(X), steps, lr, momentum, seed,
                              ctypes.byref(loss)), "mlp_train")
    return m, float(loss.value)


def run_schedule(dev, srcs, base_delta: int, schedule: Sequence[int], model: np.ndarray, device: int = 0,
                 repeats: int = 2, kind: str = "batched", **opt) -> Tuple[float, np.ndarray]:
    """Runs the kernel of `kind` under a forced per-phase schedule (the plain
    kernel from srcs[0]).  Returns the best kernel time (ms) over `repeats`
    runs and the per-phase feature records."""
    from .apps import Options, sssp_into
    from .batch import DeviceArray, last_features, sssp_multi_into
    set_model(forced(model, schedule), device, kind)
    nsrc = len(srcs) if kind == "batched" else 1
    D = DeviceArray(dev.nnodes * nsrc * 4, device)
    try:
        best = float("inf")
        feats = None
        for _ in range(repeats):
            if kind == "batched":
                st = sssp_multi_into(dev, srcs, Options(delta=base_delta, sched=1, trace=True, **opt), D)
            else:
                st = sssp_into(dev, D.ptr, Options(delta=base_delta, start_node=int(srcs[0]), sched=1, trace=True,
                                                   **opt))
            if st["ms_kernel"] < best:
                best = st["ms_kernel"]
                feats = last_features(device)
        return best, feats
    finally:
        D.close()


def sweep_instance(dev, srcs, base_delta: int, model: np.ndarray, max_greedy_phases: int = 12, device: int = 0,
                   log=None, kind: str = "batched", **opt) -> Tuple[np.ndarray, np.ndarray, dict]:
    """Labels one training instance by sweeping schedules of the plain method.

    Every constant action is timed; then, phase by phase (up to
    max_greedy_phases), each action is tried at that phase with the best
    schedule found so far elsewhere, and the faster one is kept.  Returns the
    fastest schedule's per-phase features X [P][16], its actions y [P] and a
    summary (times of the fixed plain schedule, the best constant and the best
    schedule)."""
    times = {}

    def timed(sch):
        key = tuple(sch)
        if key not in times:
            times[key] = run_schedule(dev, srcs, base_delta, sch, model, device, kind=kind, **opt)
        return times[key]

    fixed_ms, _ = timed([FIXED_BY_KIND[kind]])
    const = {a: timed([a])[0] for a in range(N_ACT)}
    best_const = min(const, key=const.get)
    best = [best_const]
    best_ms, feats = timed(best)
    nph = len(feats)
    def extend(sch, n):
        return [sch[min(i, len(sch) - 1)] for i in range(n)]

    for p in range(min(nph, max_greedy_phases)):
        for a in range(N_ACT):
            cand = extend(best, max(len(best), p + 2))
            cand[p] = a
            ms, f = timed(cand)
            if ms < best_ms * 0.995:
                best, best_ms, feats = cand, ms, f
        if log:
            log(f"  phase {p}: best {best_ms:.2f} ms schedule {best[:p + 2]}")
    X = feats[:, :N_FEAT].copy()
    y = feats[:, N_FEAT].astype(np.int32)
    summary = {"fixed_ms": fixed_ms, "best_const": int(best_const), "best_const_ms": const[best_const],
               "best_ms": best_ms, "schedule": [int(a) for a in best], "phases": int(len(y)), "runs": len(times)}
    return X, y, summary


def training_instances(quick: bool = False, kind: str = "batched") -> List[dict]:
    """Graphs the scheduler is trained on (none of them is a benchmark graph)."""
    if kind == "plain":
        # road-like grids across sizes and delta widths, plain grids, RMAT; the
        # benchmark graphs (seed 1) are not among them
        inst = [dict(kind="road", rows=1024, cols=1024, seed=21, nsrc=1, delta=10000),
                dict(kind="road", rows=1536, cols=1536, seed=22, nsrc=1, delta=10000),
                dict(kind="road", rows=2048, cols=1024, seed=23, nsrc=1, delta=5000),
                dict(kind="road", rows=1024, cols=1024, seed=29, nsrc=1, delta=50000),
                dict(kind="road", rows=512, cols=512, seed=30, nsrc=1, delta=2000),
                dict(kind="grid", rows=512, cols=512, seed=24, nsrc=1, delta=32),
                dict(kind="grid", rows=256, cols=1024, seed=25, nsrc=1, delta=100),
                dict(kind="rmat", scale=16, seed=26, nsrc=1, delta=32),
                dict(kind="rmat", scale=18, seed=27, nsrc=1, delta=32),
                dict(kind="rmat", scale=17, seed=28, nsrc=1, delta=512)]
        return inst[:3] if quick else inst
    inst = []
    for scale, seed in ((14, 11), (16, 12), (17, 13), (18, 14)):
        inst.append(dict(kind="rmat", scale=scale, seed=seed, nsrc=64, delta=32))
    inst.append(dict(kind="rmat", scale=16, seed=15, nsrc=32, delta=256))
    inst.append(dict(kind="grid", rows=256, cols=256, seed=16, nsrc=64, delta=32))
    inst.append(dict(kind="road", rows=512, cols=512, seed=17, nsrc=64, delta=1000))
    inst.append(dict(kind="road", rows=384, cols=384, seed=18, nsrc=32, delta=5000))
    return inst[:3] if quick else inst


def build_instance(spec: dict):
    from .graph import CSRGraph
    if spec["kind"] == "rmat":
        g = CSRGraph.rmat(spec["scale"], 16, seed=spec["seed"])
    elif spec["kind"] == "grid":
        g = CSRGraph.grid(spec["rows"], spec["cols"], 1, 100, seed=spec["seed"])
    else:
        g = CSRGraph.roadlike(spec["rows"], spec["cols"], 0.10, 0.001, 1, 10000, seed=spec["seed"])
    rng = np.random.default_rng(spec["seed"] + 7)
    deg = g.out_degree()
    cand = np.flatnonzero(deg > 0)
    srcs = rng.choice(cand, size=spec["nsrc"], replace=False).astype(np.uint32)
    return g, srcs


def train_from_sweeps(out_path: str, quick: bool = False, device: int = 0, steps: int = 6000, log=print,
                      kind: str = "batched") -> dict:
    """Full pipeline: sweep every training instance, train the MLP with the
    HIP trainer, save the model with its provenance."""
    base = new_model(0, kind)
    Xs, ys, ws, summaries = [], [], [], []
    for spec in training_instances(quick, kind):
        g, srcs = build_instance(spec)
        dev = g.copy_to_gpu(device)
        t0 = time.time()
        X, y, summ = sweep_instance(dev, srcs, spec["delta"], base, device=device, log=log, kind=kind)
        dev.close()
        gain = summ["fixed_ms"] / max(summ["best_ms"], 1e-6)
        summ.update(spec=spec, sweep_s=round(time.time() - t0, 1), gain=round(gain, 3))
        log(f"instance {spec}: fixed {summ['fixed_ms']:.2f} ms, best const a{summ['best_const']} "
            f"{summ['best_const_ms']:.2f} ms, best {summ['best_ms']:.2f} ms ({summ['runs']} runs)")
        Xs.append(X)
        ys.append(y)
        ws.append(np.full(len(y), 1.0 + np.log(gain), dtype=np.float32))
        summaries.append(summ)
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    w = np.concatenate(ws)
    model, loss = train(X, y, w, steps=steps, device=device, model=base)
    acc = float((infer(model, X, device) == y).mean())
    meta = {"kind": kind, "instances": summaries, "samples": int(len(y)), "train_loss": loss, "train_acc": acc,
            "actions": ACTIONS_BY_KIND[kind], "steps": steps, "created": time.strftime("%Y-%m-%d %H:%M:%S")}
    save_model(out_path, model, meta)
    log(f"trained on {len(y)} phase samples: loss {loss:.4f}, accuracy {acc:.3f} -> {out_path}")
    return meta


def profiled_traffic(root: str, name: str) -> Optional[float]:
    """HBM bytes per launch of kernel/workload `name` from the committed PMC
    summary profiles/<name>_pmc.txt (line 'traffic_bytes_per_launch: N'), or None."""
    p = os.path.join(root, "profiles", f"{name}_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


__all__ = ["ACTIONS", "ACTIONS_BY_KIND", "KINDS", "FIXED_ACTION", "FIXED_BY_KIND", "MODEL_FLOATS", "new_model", "forced", "set_model",
           "save_model", "load_model", "infer", "train", "default_model_path", "profiled_traffic"]


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="train the learned phase scheduler on the GPU")
    ap.add_argument("command", choices=["train"])
    ap.add_argument("--kind", choices=sorted(KINDS), default="batched")
    ap.add_argument("--out", default=None)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--steps", type=int, default=6000)
    a = ap.parse_args()
    train_from_sweeps(a.out or default_model_path(a.kind), quick=a.quick, steps=a.steps, kind=a.kind,
                      log=lambda *x: print(*x, flush=True))
