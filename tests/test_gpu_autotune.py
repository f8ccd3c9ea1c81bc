"""GPU test of the autotuner (galoisgpu_amd/autotune.py): it times real kernels,
keeps every candidate whose distances equal the first candidate's, sorts by
time, writes a per-architecture result file and reads it back."""
import json
import os
import sys

import numpy as np
import pytest

from galoisgpu_amd import CSRGraph
from galoisgpu_amd import autotune

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def test_autotune_searches_and_reads_back(tmp_path):
    g = CSRGraph.rmat(14, 16, seed=21)
    srcs = np.random.default_rng(2).integers(0, g.nnodes, 12).astype(np.uint32)
    dev = g.copy_to_gpu()
    say("tune batched")
    b = autotune.tune_batched(dev, srcs, space={"delta": [32, 1 << 30], "far_mode": [0, 2], "heavy_deg": [256],
                                                "block_size": [256]}, log=lambda *x: None)
    assert len(b) == 4 and all(b[i]["ms"] <= b[i + 1]["ms"] for i in range(3))
    say("tune single")
    p = autotune.tune_single(dev, srcs[:3], {"expand_mode": [0, 1], "block_size": [256, 512]}, fixed={"delta": 32},
                             log=lambda *x: None)
    assert len(p) == 4 and {r["expand_mode"] for r in p} == {0, 1}
    q = autotune.tune_single(dev, srcs[:2], {"expand_mode": [0, 1]}, bfs=True, log=lambda *x: None)
    assert len(q) == 2
    dev.close()
    arch = autotune.device_arch(0)
    assert arch.startswith("gfx"), arch
    os.makedirs(tmp_path / arch)
    with open(tmp_path / arch / "w.json", "w") as f:
        json.dump({"batched": {k: v for k, v in b[0].items() if k != "ms"}}, f)
    with open(tmp_path / "w.json", "w") as f:
        json.dump({"batched": {"delta": 7}, "plain": {"delta": 9}}, f)
    assert autotune.tuned_options("w", "batched", root=str(tmp_path)) == {k: v for k, v in b[0].items() if k != "ms"}
    assert autotune.tuned_options("w", "plain", root=str(tmp_path)) == {"delta": 9}  # falls back to the generic set
    assert autotune.tuned_options("nope", "plain", root=str(tmp_path)) is None
