"""The .gr loader/writer and the synthetic generators (host code of liblsg)."""
import os

import numpy as np
import pytest

import oracle
from galoisgpu_amd import CSRGraph, LsgError


def make_csr(rng, n, m, wmax=255):
    src = np.sort(rng.integers(0, n, m))
    dst = rng.integers(0, n, m).astype(np.uint32)
    row = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]).astype(np.uint32)
    return row, dst, rng.integers(1, wmax + 1, m).astype(np.int32)


@pytest.mark.parametrize("m", [0, 1, 7, 100, 101])
def test_read_matches_numpy_restatement(tmp_path, m):
    rng = np.random.default_rng(m)
    row, col, w = make_csr(rng, 20, m)
    p = str(tmp_path / "g.gr")
    oracle.write_gr(p, row, col, w)
    r2, c2, w2 = oracle.read_gr(p)
    g = CSRGraph.read(p)
    assert g.nnodes == 20 and g.nedges == m
    assert np.array_equal(g.row_start, r2) and np.array_equal(g.row_start, row)
    assert np.array_equal(g.edge_dst, c2) and np.array_equal(g.edge_data, w2)
    assert (g.node_data == 0).all()
    # odd edge counts carry one padding word before the weights (rt/src/csr_graph.cu:198)
    assert os.path.getsize(p) == 32 + 8 * 20 + 4 * m + 4 * (m % 2) + 4 * m


def test_write_round_trip(tmp_path):
    g = CSRGraph.rmat(10, 8, seed=3)
    p = str(tmp_path / "r.gr")
    g.write(p)
    h = CSRGraph.read(p)
    for a in ("row_start", "edge_dst", "edge_data"):
        assert np.array_equal(getattr(g, a), getattr(h, a))
    r2, c2, w2 = oracle.read_gr(p)
    assert np.array_equal(r2, g.row_start) and np.array_equal(c2, g.edge_dst) and np.array_equal(w2, g.edge_data)
    # unweighted files (sizeEdgeTy = 0) read back with weights 0
    p0 = str(tmp_path / "r0.gr")
    g.write(p0, with_weights=False)
    h0 = CSRGraph.read(p0)
    assert np.array_equal(h0.edge_dst, g.edge_dst) and (h0.edge_data == 0).all()


def _raw(tmp_path, name, words64, words32=()):
    p = str(tmp_path / name)
    with open(p, "wb") as f:
        f.write(np.array(words64, dtype="<u8").tobytes())
        f.write(np.array(words32, dtype="<u4").tobytes())
    return p


@pytest.mark.parametrize("case", ["version", "edgesize", "truncated", "dst_range", "offsets", "short",
                                  "zero_nodes", "negative_weight", "missing"])
def test_malformed_files_are_rejected(tmp_path, case):
    if case == "version":
        p = _raw(tmp_path, "a.gr", [2, 4, 2, 1, 1, 1], [1, 0, 5, 0])
    elif case == "edgesize":
        p = _raw(tmp_path, "a.gr", [1, 8, 2, 1, 1, 1], [1, 0, 5, 0, 0, 0])
    elif case == "truncated":
        p = _raw(tmp_path, "a.gr", [1, 4, 2, 2, 1, 2], [1, 0])
    elif case == "dst_range":
        p = _raw(tmp_path, "a.gr", [1, 4, 2, 1, 1, 1], [7, 0, 5])
    elif case == "offsets":
        p = _raw(tmp_path, "a.gr", [1, 4, 2, 2, 2, 1], [1, 0, 5, 5])
    elif case == "short":
        p = str(tmp_path / "a.gr")
        open(p, "wb").write(b"\x01\x00")
    elif case == "zero_nodes":
        p = _raw(tmp_path, "a.gr", [1, 4, 0, 0])
    elif case == "negative_weight":
        p = _raw(tmp_path, "a.gr", [1, 4, 2, 1, 1, 1], [1, 0, 0x80000000])
    else:
        p = str(tmp_path / "does_not_exist.gr")
    with pytest.raises(LsgError) as e:
        CSRGraph.read(p)
    assert e.value.name == ("LSG_ERR_IO" if case == "missing" else "LSG_ERR_FORMAT")


def test_grid_generator():
    g = CSRGraph.grid(7, 9, 1, 100, 5)
    assert g.nnodes == 63 and g.nedges == 2 * (7 * 8 + 9 * 6)
    row, col, w = g.row_start, g.edge_dst, g.edge_data
    for v in range(63):
        r, c = divmod(v, 9)
        nb = [(rr * 9 + cc) for rr, cc in ((r - 1, c), (r, c - 1), (r, c + 1), (r + 1, c))
              if 0 <= rr < 7 and 0 <= cc < 9]
        assert list(col[row[v]:row[v + 1]]) == nb
    assert w.min() >= 1 and w.max() <= 100
    assert np.array_equal(CSRGraph.grid(7, 9, 1, 100, 5).edge_data, w)
    assert not np.array_equal(CSRGraph.grid(7, 9, 1, 100, 6).edge_data, w)


def _check_simple(g, wlo, whi):
    row, col, w = g.row_start.astype(np.int64), g.edge_dst.astype(np.int64), g.edge_data
    src = np.repeat(np.arange(g.nnodes), np.diff(row))
    assert (src != col).all(), "self loop"
    key = src * g.nnodes + col
    assert (np.diff(key) > 0).all(), "rows not sorted or duplicate edge"
    assert w.min() >= wlo and w.max() <= whi


def test_rmat_generator():
    g = CSRGraph.rmat(12, 16, seed=9)
    assert g.nnodes == 4096
    assert 0.8 * 16 * 4096 < g.nedges < 16 * 4096
    _check_simple(g, 1, 255)
    h = CSRGraph.rmat(12, 16, seed=9)
    assert np.array_equal(g.edge_dst, h.edge_dst) and np.array_equal(g.edge_data, h.edge_data)
    deg = g.out_degree()
    assert g.max_degree_vertex() == int(np.argmax(deg))
    assert deg.max() > 20 * deg.mean()  # heavy-tailed


def test_roadlike_generator():
    g = CSRGraph.roadlike(64, 64, 0.10, 0.001 * 20, 1, 10000, seed=2)
    _check_simple(g, 1, 10000)
    row, col = g.row_start.astype(np.int64), g.edge_dst.astype(np.int64)
    src = np.repeat(np.arange(g.nnodes), np.diff(row))
    fwd = set(zip(src.tolist(), col.tolist()))
    assert all((v, u) in fwd for u, v in fwd), "edges must come in both directions"
    grid_edges = 2 * (64 * 63 * 2)
    assert 0.85 * grid_edges < g.nedges < grid_edges * 1.05


def test_output_text(tmp_path):
    d = np.array([0, 5, 0x7FFFFFFF, 123456789, 2147483646], dtype=np.int32)
    from galoisgpu_amd import output
    p = str(tmp_path / "o.txt")
    output(d, p)
    want = "".join(("%d INF\n" % i) if x == 0x7FFFFFFF else ("%d %d\n" % (i, x)) for i, x in enumerate(d))
    assert open(p).read() == want
    output(d, None)  # no -o: nothing written, no error
    with pytest.raises(LsgError):
        output(d, str(tmp_path / "no_such_dir" / "o.txt"))
