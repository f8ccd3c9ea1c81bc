"""Measured HBM copy bandwidth (the stream-copy reference figure).

python -m galoisgpu_amd.copybench [--bytes N] [--iters K] prints one JSON line:
bytes copied per launch, ms per launch and GB/s counting read + write.  It is
also the known-size kernel that rocprofv3's FETCH_SIZE / WRITE_SIZE counters
are calibrated against.
"""
from __future__ import annotations

import argparse
import ctypes
import json

from ._lib import check, lib


def copy_bandwidth(nbytes: int = 1 << 31, iters: int = 20, device: int = 0) -> dict:
    ms = ctypes.c_double()
    check(lib().lsg_stream_copy(device, nbytes, iters, ctypes.byref(ms)), "stream_copy")
    return {"bytes": nbytes, "iters": iters, "ms_per_copy": ms.value,
            "gbs_read_plus_write": 2.0 * nbytes / (ms.value / 1e3) / 1e9}


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 31)
    ap.add_argument("--iters", type=int, default=20)
    a = ap.parse_args()
    print(json.dumps(copy_bandwidth(a.bytes, a.iters)), flush=True)
