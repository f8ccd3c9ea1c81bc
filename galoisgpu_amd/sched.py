"""Learned phase scheduler: model files, forced schedules, trace sweeps and training.

The model (layout in csrc/sched_mlp.h) maps 16 frontier features, taken by the
persistent kernel's barrier leader at every phase boundary, to one of 8
actions: delta multiplier in {1/4, 1, 4, 16} x far-pile mode in {reference
(one shared threshold), bucket (per-source thresholds)}.  Training data comes from the plain method itself: for each training
instance the kernel is run under a sweep of forced per-phase schedules, the
fastest one is kept, and the features it recorded at each phase are labelled
with the action it took there (``sweep_instance``).  ``train`` fits the MLP with
the HIP training kernel (lsg_mlp_train).  Whatever the model picks, distances
are identical: the actions only reorder the same u32 relaxations.
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))

N_FEAT, N_HID, N_ACT, MAX_FORCED = 16, 64, 8, 256
DELTA_MULTS = [0.25, 1.0, 4.0, 16.0]
FAR_REFERENCE, FAR_SPLIT, FAR_BUCKET = 0, 1, 2
# action id = 4 * (far mode is bucket) + delta multiplier index
ACTIONS = [(m, far) for far in (FAR_REFERENCE, FAR_BUCKET) for m in DELTA_MULTS]
FIXED_ACTION = 1  # delta x1, reference far pile: the plain method's schedule

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


def default_model_path() -> str:
    return os.path.join(HERE, "models", "sched_model.npz")


def new_model(seed: int = 0) -> np.ndarray:
    """Randomly initialised network (He init), action tables, no forced schedule."""
    rng = np.random.default_rng(seed)
    m = np.zeros(MODEL_FLOATS, dtype=np.float32)
    m[OFF_INVSTD:OFF_INVSTD + N_FEAT] = 1.0
    m[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2.0 / N_FEAT), N_FEAT * N_HID)
    m[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2.0 / N_HID), N_HID * N_HID)
    m[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1.0 / N_HID), N_HID * N_ACT)
    for a, (mult, far) in enumerate(ACTIONS):
        m[OFF_DELTA + a] = mult
        m[OFF_FAR + a] = far
    m[OFF_FORCED:OFF_FORCED + MAX_FORCED] = -1.0
    return m


def forced(model: np.ndarray, actions: Sequence[int]) -> np.ndarray:
    """Copy of `model` whose phase p takes actions[min(p, len-1)] regardless of the MLP."""
    m = model.copy()
    acts = list(actions) or [FIXED_ACTION]
    for p in range(MAX_FORCED):
        m[OFF_FORCED + p] = acts[min(p, len(acts) - 1)]
    return m


def set_model(model: np.ndarray, device: int = 0) -> None:
    m = np.ascontiguousarray(model, dtype=np.float32)
    check(lib().lsg_sched_set_model(device, m.ctypes.data_as(ctypes.c_void_p), len(m)), "sched_set_model")


def save_model(path: str, model: np.ndarray, meta: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, model=np.asarray(model, dtype=np.float32), meta=np.array(json.dumps(meta)))


def load_model(path: Optional[str] = None, device: int = 0) -> Tuple[np.ndarray, dict]:
    z = np.load(path or default_model_path(), allow_pickle=False)
    model = z["model"].astype(np.float32)
    if len(model) != MODEL_FLOATS:
        raise ValueError(f"{path}: model has {len(model)} floats, expected {MODEL_FLOATS}")
    model[OFF_FORCED:OFF_FORCED + MAX_FORCED] = -1.0  # a stored model never forces phases
    set_model(model, device)
    return model, json.loads(str(z["meta"]))


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


__all__ = ["ACTIONS", "FIXED_ACTION", "MODEL_FLOATS", "new_model", "forced", "set_model", "save_model",
           "load_model", "infer", "train", "default_model_path", "profiled_traffic"]
