import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path through the C ABI")


def load_small_graphs():
    z = np.load(os.path.join(GOLDEN, "small_graphs.npz"))
    out = {}
    for name in z["names"]:
        name = str(name)
        out[name] = {k: z[f"{name}_{k}"] for k in ("row", "col", "w", "dist", "bfs")}
        out[name]["src"] = int(z[f"{name}_src"][0])
    return out


@pytest.fixture(scope="session")
def small_graphs():
    return load_small_graphs()


@pytest.fixture(scope="session")
def grid256_golden():
    return np.load(os.path.join(GOLDEN, "grid256.npz"))
