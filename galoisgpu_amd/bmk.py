"""bmk2-style run-and-check harness for the sssp / bfs CLIs.

The reference's benchmark spec (apps/sssp/bmktest2.py:8-20) runs the app
binary as ``test INPUT -d DELTA -o @output``, checks ``@output`` against the
input's reference output with bmk2's DiffChecker, and takes the time from the
app's ``Total time:`` lines (skelapp/skel.cu:65-66) with a PerfRE.  This module
does the same against ``galoisgpu_amd/bin/sssp`` and ``bin/bfs``:

  python -m galoisgpu_amd.bmk sssp INPUT [--delta D] [--source S] [--ref REF] [--runs N]
  python -m galoisgpu_amd.bmk bfs  INPUT [--source S] [--ref REF] [--runs N]

INPUT is a .gr path or a reference input name resolved in $LSG_INPUTS
(``grtool.resolve_input``); REF defaults to ``<input>.<app>.out`` beside it.
Prints one JSON record: the command, per-run times in ns, and the check result
(``pass``, ``fail`` with the first differing line, or ``no-reference``).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
from typing import List, Optional

from .grtool import resolve_input

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "bin")
# The app prints "Total time: <n> ms" and "Total time: <n> ns" on stderr (skelapp/skel.cu:65-66).
PERF_RE = re.compile(r"^Total time: (\d+) ns$", re.MULTILINE)


def parse_time_ns(stderr: str) -> Optional[int]:
    m = PERF_RE.findall(stderr)
    return int(m[-1]) if m else None


def diff_check(output: str, reference: str) -> dict:
    """DiffChecker: the output file must equal the reference file byte for byte."""
    with open(output, "rb") as a, open(reference, "rb") as b:
        out, ref = a.read(), b.read()
    if out == ref:
        return {"check": "pass"}
    ol, rl = out.split(b"\n"), ref.split(b"\n")
    for i, (x, y) in enumerate(zip(ol, rl)):
        if x != y:
            return {"check": "fail", "line": i + 1, "output": x.decode(errors="replace"),
                    "reference": y.decode(errors="replace")}
    return {"check": "fail", "line": min(len(ol), len(rl)) + 1, "output_lines": len(ol), "reference_lines": len(rl)}


def command(app: str, inp: str, out: str, delta: Optional[int] = None, source: Optional[int] = None,
            extra: Optional[List[str]] = None) -> List[str]:
    cmd = [os.path.join(BIN, app)]
    if app == "sssp" and delta is not None:
        cmd += ["-d", str(delta)]
    if source is not None:
        cmd += ["-s", str(source)]
    cmd += list(extra or []) + ["-o", out, inp]
    return cmd


def run(app: str, input_name: str, delta: Optional[int] = None, source: Optional[int] = None,
        ref: Optional[str] = None, runs: int = 1, extra: Optional[List[str]] = None) -> dict:
    if app not in ("sssp", "bfs"):
        raise ValueError("app must be sssp or bfs")
    inp = resolve_input(input_name)
    if not os.path.exists(os.path.join(BIN, app)):
        raise FileNotFoundError(f"{os.path.join(BIN, app)} is not built (make -C galoisgpu_amd/csrc)")
    ref = ref or os.path.splitext(inp)[0] + f".{app}.out"
    rec = {"app": app, "input": inp, "times_ns": []}
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "output")
        cmd = command(app, inp, out, delta, source, extra)
        rec["command"] = " ".join(cmd)
        for _ in range(runs):
            p = subprocess.run(cmd, capture_output=True, text=True)
            if p.returncode != 0:
                rec.update(check="error", returncode=p.returncode, stderr=p.stderr[-2000:])
                return rec
            rec["times_ns"].append(parse_time_ns(p.stderr))
        rec.update(diff_check(out, ref) if os.path.exists(ref) else {"check": "no-reference", "reference": ref})
    return rec


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m galoisgpu_amd.bmk")
    ap.add_argument("app", choices=["sssp", "bfs"])
    ap.add_argument("input")
    ap.add_argument("--delta", type=int, default=None)
    ap.add_argument("--source", type=int, default=None)
    ap.add_argument("--ref", default=None)
    ap.add_argument("--runs", type=int, default=1)
    a = ap.parse_args(argv)
    rec = run(a.app, a.input, a.delta, a.source, a.ref, a.runs)
    print(json.dumps(rec))
    return 0 if rec.get("check") in ("pass", "no-reference") else 1


if __name__ == "__main__":
    sys.exit(main())
