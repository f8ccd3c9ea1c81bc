"""`.gr` tooling: convert edge files to the binary CSR format, generate inputs, inspect files.

The `.gr` v1 layout (rt/src/csr_graph.cu:187-199) is read and written by the
library (lsg_graph_read / lsg_graph_write); this module builds the CSR from
text inputs and names the reference's benchmark inputs.

  python -m galoisgpu_amd.grtool convert IN OUT [--format auto|dimacs|mtx|edgelist]
                                            [--symmetric] [--dedupe] [--weights LO HI] [--seed S]
  python -m galoisgpu_amd.grtool gen rmat SCALE OUT [--ef 16] [--seed 1] [--weights 1 255]
  python -m galoisgpu_amd.grtool gen grid ROWS COLS OUT [--seed 1] [--weights 1 100]
  python -m galoisgpu_amd.grtool gen road ROWS COLS OUT [--seed 1] [--weights 1 10000]
  python -m galoisgpu_amd.grtool info FILE.gr
  python -m galoisgpu_amd.grtool inputs [--dir DIR]

Input formats for `convert`:
  dimacs    9th DIMACS shortest-path files ("p sp N M", "a U V W", 1-based), the
            USA-road-d.* inputs of the reference's benchmark list
  mtx       Matrix Market coordinate files (1-based; a third column is the weight;
            "symmetric" headers add both directions)
  edgelist  whitespace-separated "U V [W]" lines, 0-based, '#' or '%' comments
Edges are ordered by source (stable within a row, i.e. in input order) as
CSRGraph requires; --dedupe drops self loops and keeps the lightest of
duplicate edges.  Without weights in the input the file gets none
(sizeEdgeTy 0, read back as zero weights) unless --weights draws them.
"""
from __future__ import annotations

import argparse
import os
import sys
from typing import Optional, Tuple

import numpy as np

from .graph import CSRGraph

# Benchmark inputs of the reference's SSSP run list (apps/sssp/bmktest2.py:23).
REFERENCE_INPUTS = ["USA-road-d.NY", "r4-2e23", "rmat22", "USA-road-d.USA", "USA-road-d.CAL"]


def inputs_dir(d: Optional[str] = None) -> str:
    return d or os.environ.get("LSG_INPUTS", os.path.join(os.getcwd(), "inputs"))


def resolve_input(name: str, d: Optional[str] = None) -> str:
    """Path of input `name` (a file, or a reference input name looked up as
    <dir>/<name>.gr in $LSG_INPUTS or ./inputs)."""
    if os.path.exists(name):
        return name
    p = os.path.join(inputs_dir(d), name if name.endswith(".gr") else name + ".gr")
    if not os.path.exists(p):
        raise FileNotFoundError(f"input {name!r} not found (looked for {p}; set LSG_INPUTS or pass a path)")
    return p


def _sniff(path: str) -> str:
    with open(path, "r", errors="replace") as f:
        for line in f:
            t = line.strip()
            if not t:
                continue
            if t.startswith("%%MatrixMarket"):
                return "mtx"
            if t.startswith("c ") or t == "c" or t.startswith("p "):
                return "dimacs"
            if t[0] in "#%":
                continue
            return "edgelist"
    return "edgelist"


def _table(path: str, comment: str, ncols: int):
    """Whitespace-separated table as a pandas frame of strings (C parser)."""
    import pandas as pd
    return pd.read_csv(path, sep=r"\s+", header=None, comment=comment, names=list(range(ncols)), dtype=str,
                       engine="c", skip_blank_lines=True)


def _ints(col) -> np.ndarray:
    return col.astype(float).astype(np.int64).to_numpy() if col.dtype == object else col.to_numpy(np.int64)


def read_edges(path: str, fmt: str = "auto"):
    """Returns (nnodes, src, dst, weight or None, symmetric) with 0-based ids."""
    fmt = _sniff(path) if fmt == "auto" else fmt
    if fmt == "dimacs":
        t = _table(path, "c", 4)
        prob = t[t[0] == "p"]
        if prob.empty:
            raise ValueError(f"{path}: DIMACS file without a 'p' line")
        n = int(prob.iloc[0][2])
        arcs = t[t[0] == "a"]
        w = _ints(arcs[3]) if arcs[3].notna().all() and len(arcs) else np.ones(len(arcs), dtype=np.int64)
        return n, _ints(arcs[1]) - 1, _ints(arcs[2]) - 1, w, False
    if fmt == "mtx":
        with open(path, "r") as f:
            header = f.readline().lower()
        if "coordinate" not in header:
            raise ValueError(f"{path}: only coordinate Matrix Market files are supported")
        t = _table(path, "%", 3)
        nr, nc = int(t.iloc[0][0]), int(t.iloc[0][1])
        e = t.iloc[1:]
        w = _ints(e[2]) if "pattern" not in header and e[2].notna().all() and len(e) else None
        return max(nr, nc), _ints(e[0]) - 1, _ints(e[1]) - 1, w, "symmetric" in header
    if fmt == "edgelist":
        lines = []
        with open(path, "r") as f:
            for line in f:
                s = line.split("#", 1)[0].split("%", 1)[0].strip()
                if s:
                    lines.append(s)
        if not lines:
            z = np.zeros(0, dtype=np.int64)
            return 0, z, z, None, False
        width = len(lines[0].split())
        if width not in (2, 3) or any(len(ln.split()) != width for ln in lines):
            raise ValueError(f"{path}: every line must be 'U V' or every line 'U V W'")
        a = np.array([ln.split() for ln in lines], dtype=float).astype(np.int64)
        n = int(max(a[:, 0].max(), a[:, 1].max())) + 1
        return n, a[:, 0], a[:, 1], (a[:, 2] if a.shape[1] > 2 else None), False
    raise ValueError(f"unknown format {fmt!r}")


def build_csr(n: int, src, dst, w=None, symmetric: bool = False, dedupe: bool = False) -> CSRGraph:
    """CSRGraph with rows ordered by source, edges within a row kept in input
    order (or sorted and deduplicated with `dedupe`)."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    if len(src) and (src.min() < 0 or dst.min() < 0 or src.max() >= n or dst.max() >= n):
        raise ValueError(f"edge endpoint outside [0, {n})")
    if w is not None:
        w = np.asarray(w, dtype=np.int64)
        if len(w) and (w.min() < 0 or w.max() > 0x7FFFFFFF):
            raise ValueError("weights must lie in [0, INT_MAX]")
    if symmetric:
        off = src != dst
        src, dst = np.concatenate([src, dst[off]]), np.concatenate([dst, src[off]])
        if w is not None:
            w = np.concatenate([w, w[off]])
    if dedupe:
        keep = src != dst
        src, dst = src[keep], dst[keep]
        ww = w[keep] if w is not None else np.zeros(len(src), dtype=np.int64)
        order = np.lexsort((ww, dst, src))
        src, dst, ww = src[order], dst[order], ww[order]
        first = np.ones(len(src), dtype=bool)
        first[1:] = (src[1:] != src[:-1]) | (dst[1:] != dst[:-1])
        src, dst, ww = src[first], dst[first], ww[first]
        w = ww if w is not None else None
    else:
        order = np.argsort(src, kind="stable")
        src, dst = src[order], dst[order]
        if w is not None:
            w = w[order]
    if len(src) >= 1 << 32:
        raise ValueError("more than 2^32 - 1 edges do not fit the u32 CSR")
    row = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row)
    return CSRGraph.from_arrays(row.astype(np.uint32), dst.astype(np.uint32), w)


def convert(inp: str, out: str, fmt: str = "auto", symmetric: bool = False, dedupe: bool = False,
            weights: Optional[Tuple[int, int]] = None, seed: int = 1) -> CSRGraph:
    n, src, dst, w, sym_hint = read_edges(inp, fmt)
    if weights is not None:
        lo, hi = weights
        w = np.random.default_rng(seed).integers(lo, hi + 1, len(src))
    g = build_csr(n, src, dst, w, symmetric or sym_hint, dedupe)
    g.write(out, with_weights=w is not None)
    return g


def info(path: str) -> dict:
    g = CSRGraph.read(path)
    deg = g.out_degree()
    return {"file": path, "nnodes": g.nnodes, "nedges": g.nedges,
            "max_degree": int(deg.max()) if len(deg) else 0,
            "max_degree_vertex": g.max_degree_vertex() if g.nnodes else 0,
            "mean_degree": float(deg.mean()) if len(deg) else 0.0,
            "isolated": int((deg == 0).sum()),
            "min_weight": int(g.edge_data.min()) if g.nedges else 0,
            "max_weight": int(g.edge_data.max()) if g.nedges else 0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m galoisgpu_amd.grtool", description=".gr graph tooling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("convert")
    c.add_argument("input")
    c.add_argument("output")
    c.add_argument("--format", default="auto", choices=["auto", "dimacs", "mtx", "edgelist"])
    c.add_argument("--symmetric", action="store_true")
    c.add_argument("--dedupe", action="store_true")
    c.add_argument("--weights", nargs=2, type=int, metavar=("LO", "HI"))
    c.add_argument("--seed", type=int, default=1)
    gen = sub.add_parser("gen")
    gen.add_argument("kind", choices=["rmat", "grid", "road"])
    gen.add_argument("args", nargs="+", help="rmat: SCALE OUT; grid/road: ROWS COLS OUT")
    gen.add_argument("--ef", type=int, default=16)
    gen.add_argument("--seed", type=int, default=1)
    gen.add_argument("--weights", nargs=2, type=int, metavar=("LO", "HI"))
    i = sub.add_parser("info")
    i.add_argument("file")
    ls = sub.add_parser("inputs")
    ls.add_argument("--dir", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "convert":
        g = convert(a.input, a.output, a.format, a.symmetric, a.dedupe,
                    tuple(a.weights) if a.weights else None, a.seed)
        print(f"{a.output}: {g.nnodes} nodes, {g.nedges} edges")
    elif a.cmd == "gen":
        if a.kind == "rmat":
            if len(a.args) != 2:
                ap.error("gen rmat SCALE OUT")
            lo, hi = a.weights or (1, 255)
            g = CSRGraph.rmat(int(a.args[0]), a.ef, 0.57, 0.19, 0.19, lo, hi, seed=a.seed)
        else:
            if len(a.args) != 3:
                ap.error(f"gen {a.kind} ROWS COLS OUT")
            r, cols = int(a.args[0]), int(a.args[1])
            if a.kind == "grid":
                lo, hi = a.weights or (1, 100)
                g = CSRGraph.grid(r, cols, lo, hi, seed=a.seed)
            else:
                lo, hi = a.weights or (1, 10000)
                g = CSRGraph.roadlike(r, cols, 0.10, 0.001, lo, hi, seed=a.seed)
        g.write(a.args[-1])
        print(f"{a.args[-1]}: {g.nnodes} nodes, {g.nedges} edges")
    elif a.cmd == "info":
        for k, v in info(a.file).items():
            print(f"{k}: {v}")
    else:
        d = inputs_dir(a.dir)
        for name in REFERENCE_INPUTS:
            p = os.path.join(d, name + ".gr")
            print(f"{name}: {p if os.path.exists(p) else 'missing (' + p + ')'}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
