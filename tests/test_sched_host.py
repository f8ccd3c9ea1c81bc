"""Host-side logic of the learned scheduler and the autotuner (no GPU needed)."""
import json
import os

import numpy as np
import pytest

from galoisgpu_amd import autotune
from galoisgpu_amd import sched as lsched


def test_model_layout_matches_library():
    assert lsched.model_floats() == lsched.MODEL_FLOATS
    # 16 -> 64 -> 64 -> 8 plus normalisation, action tables and the forced table
    assert lsched.MODEL_FLOATS == 32 + 16 * 64 + 64 + 64 * 64 + 64 + 64 * 8 + 8 + 8 + 8 + 256


def test_new_model_action_tables():
    m = lsched.new_model(3)
    for a, (mult, far) in enumerate(lsched.ACTIONS):
        assert m[lsched.OFF_DELTA + a] == np.float32(mult)
        assert m[lsched.OFF_FAR + a] == far
    assert (m[lsched.OFF_FORCED:lsched.OFF_FORCED + lsched.MAX_FORCED] == -1).all()
    assert lsched.ACTIONS[lsched.FIXED_ACTION] == (1.0, 0)  # the plain method's schedule
    assert np.isfinite(m).all()


def test_plain_model_action_tables():
    m = lsched.new_model(3, kind="plain")
    for a, (mult, mode) in enumerate(lsched.ACTIONS_BY_KIND["plain"]):
        assert m[lsched.OFF_DELTA + a] == np.float32(mult)
        assert m[lsched.OFF_FAR + a] == mode  # expansion mode: 0 TWB, 1 LB
    # delta x1 with the default, tile-flattened expansion
    assert lsched.ACTIONS_BY_KIND["plain"][lsched.FIXED_BY_KIND["plain"]] == (1.0, 1)
    assert sorted({mode for _, mode in lsched.ACTIONS_BY_KIND["plain"]}) == [0, 1]


def test_set_model_rejects_bad_kind_and_size():
    from galoisgpu_amd._lib import LSG_OK, lib
    import ctypes
    m = np.zeros(lsched.MODEL_FLOATS, dtype=np.float32)
    p = m.ctypes.data_as(ctypes.c_void_p)
    assert lib().lsg_sched_set_model(0, 7, p, len(m)) != LSG_OK
    assert lib().lsg_sched_set_model(0, 1, p, len(m) - 1) != LSG_OK
    bad = m.copy()
    bad[5] = np.nan
    assert lib().lsg_sched_set_model(0, 1, bad.ctypes.data_as(ctypes.c_void_p), len(m)) != LSG_OK


def test_forced_schedule_extends_last_action():
    m = lsched.forced(lsched.new_model(0), [5, 2, 7])
    f = m[lsched.OFF_FORCED:lsched.OFF_FORCED + lsched.MAX_FORCED]
    assert list(f[:3]) == [5, 2, 7] and (f[3:] == 7).all()
    assert (lsched.forced(lsched.new_model(0), [])[lsched.OFF_FORCED:][:4] == lsched.FIXED_ACTION).all()


def test_save_model_round_trip(tmp_path):
    m = lsched.new_model(1)
    p = str(tmp_path / "m.npz")
    lsched.save_model(p, m, {"note": "x", "samples": 3})
    z = np.load(p, allow_pickle=False)
    assert np.array_equal(z["model"], m) and json.loads(str(z["meta"]))["samples"] == 3


def test_autotune_grid_and_specs():
    g = autotune.grid({"a": [1, 2], "b": [3, 4, 5]})
    assert len(g) == 6 and {"a": 2, "b": 5} in g
    assert len(autotune.grid(autotune.BATCH_SPACE)) == 4 * 2 * 3 * 2
    assert autotune.make_graph("grid:8").nnodes == 64
    with pytest.raises(ValueError):
        autotune.make_graph("cube:3")


def test_tuned_options_file():
    p = os.path.join(autotune.HERE, "tuning", "rmat22_64src.json")
    if not os.path.exists(p):
        pytest.skip("no tuning file committed yet")
    b = autotune.tuned_options("rmat22_64src", "batched")
    pl = autotune.tuned_options("rmat22_64src", "plain")
    assert set(b) <= set(autotune.BATCH_SPACE) and set(pl) <= set(autotune.PLAIN_SPACE)
    assert autotune.tuned_options("no_such_workload", "batched") is None


def test_profiled_traffic_parser(tmp_path):
    os.makedirs(tmp_path / "profiles")
    (tmp_path / "profiles" / "k_pmc.txt").write_text("command: x\ntraffic_bytes_per_launch: 12345.5\n")
    assert lsched.profiled_traffic(str(tmp_path), "k") == 12345.5
    assert lsched.profiled_traffic(str(tmp_path), "missing") is None
