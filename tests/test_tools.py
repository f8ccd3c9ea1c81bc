"""`.gr` tooling (grtool) and the bmk2-style harness (bmk): host logic, no GPU."""
import json
import os

import numpy as np
import pytest

from galoisgpu_amd import CSRGraph, bmk, grtool


def read_back(path):
    g = CSRGraph.read(path)
    return g.row_start.copy(), g.edge_dst.copy(), g.edge_data.copy()


def test_convert_dimacs(tmp_path):
    src = tmp_path / "g.dimacs"
    src.write_text("c tiny road graph\np sp 4 5\na 1 2 7\na 1 3 2\na 2 4 1\na 3 2 3\na 4 1 5\n")
    out = str(tmp_path / "g.gr")
    g = grtool.convert(str(src), out)
    assert g.nnodes == 4 and g.nedges == 5
    row, dst, w = read_back(out)
    assert list(row) == [0, 2, 3, 4, 5]
    assert list(dst) == [1, 2, 3, 1, 0]
    assert list(w) == [7, 2, 1, 3, 5]


def test_convert_mtx_symmetric_and_pattern(tmp_path):
    src = tmp_path / "g.mtx"
    src.write_text("%%MatrixMarket matrix coordinate integer symmetric\n% comment\n3 3 3\n2 1 4\n3 2 9\n3 3 1\n")
    out = str(tmp_path / "g.gr")
    grtool.convert(str(src), out)
    row, dst, w = read_back(out)
    # edges 1->0 (4), 2->1 (9), self loop 2->2 (1), plus reverses 0->1, 1->2 (loop not doubled)
    assert list(row) == [0, 1, 3, 5]
    assert list(dst) == [1, 0, 2, 1, 2]
    assert list(w) == [4, 4, 9, 9, 1]
    pat = tmp_path / "p.mtx"
    pat.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 2\n")
    grtool.convert(str(pat), str(tmp_path / "p.gr"))
    row, dst, w = read_back(str(tmp_path / "p.gr"))
    assert list(row) == [0, 1, 1] and list(dst) == [1] and list(w) == [0]  # no weights stored: read as 0


def test_convert_edgelist_dedupe_and_weights(tmp_path):
    src = tmp_path / "e.txt"
    src.write_text("# u v w\n0 1 5\n0 1 3\n1 1 2\n2 0 4\n1 2 6\n")
    out = str(tmp_path / "e.gr")
    grtool.convert(str(src), out)
    row, dst, w = read_back(out)
    assert list(row) == [0, 2, 4, 5] and list(dst) == [1, 1, 1, 2, 0] and list(w) == [5, 3, 2, 6, 4]
    grtool.convert(str(src), out, dedupe=True)
    row, dst, w = read_back(out)
    assert list(row) == [0, 1, 2, 3] and list(dst) == [1, 2, 0] and list(w) == [3, 6, 4]
    plain = tmp_path / "u.txt"
    plain.write_text("0 1\n1 2\n2 0\n")
    grtool.convert(str(plain), out, weights=(1, 9), seed=4)
    _, _, w = read_back(out)
    assert len(w) == 3 and w.min() >= 1 and w.max() <= 9


def test_convert_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("0 1 2\n3 4\n")
    with pytest.raises(ValueError):
        grtool.convert(str(bad), str(tmp_path / "x.gr"))
    neg = tmp_path / "neg.txt"
    neg.write_text("0 1 -2\n")
    with pytest.raises(ValueError):
        grtool.convert(str(neg), str(tmp_path / "x.gr"))
    nop = tmp_path / "nop.dimacs"
    nop.write_text("c no problem line\na 1 2 3\n")
    with pytest.raises(ValueError):
        grtool.convert(str(nop), str(tmp_path / "x.gr"), fmt="dimacs")


def test_gen_info_and_inputs(tmp_path, capsys, monkeypatch):
    out = str(tmp_path / "r.gr")
    assert grtool.main(["gen", "rmat", "8", out, "--ef", "4", "--seed", "3"]) == 0
    g = CSRGraph.rmat(8, 4, 0.57, 0.19, 0.19, 1, 255, seed=3)
    row, dst, w = read_back(out)
    assert np.array_equal(row, g.row_start) and np.array_equal(dst, g.edge_dst) and np.array_equal(w, g.edge_data)
    d = grtool.info(out)
    assert d["nnodes"] == 256 and d["nedges"] == g.nedges and d["max_degree_vertex"] == g.max_degree_vertex()
    monkeypatch.setenv("LSG_INPUTS", str(tmp_path))
    os.rename(out, str(tmp_path / "rmat22.gr"))
    assert grtool.resolve_input("rmat22") == str(tmp_path / "rmat22.gr")
    with pytest.raises(FileNotFoundError):
        grtool.resolve_input("USA-road-d.NY")
    assert grtool.main(["inputs"]) == 0
    listing = capsys.readouterr().out
    assert "rmat22: " + str(tmp_path / "rmat22.gr") in listing and "USA-road-d.NY: missing" in listing


def test_bmk_perf_and_diff_checker(tmp_path):
    err = "Using GPU: x\nTotal time: 12 ms\nTotal time: 12345678 ns\n"
    assert bmk.parse_time_ns(err) == 12345678
    assert bmk.parse_time_ns("nothing") is None
    a, b = tmp_path / "a", tmp_path / "b"
    a.write_text("0 0\n1 INF\n2 7\n")
    b.write_text("0 0\n1 INF\n2 7\n")
    assert bmk.diff_check(str(a), str(b)) == {"check": "pass"}
    b.write_text("0 0\n1 5\n2 7\n")
    r = bmk.diff_check(str(a), str(b))
    assert r["check"] == "fail" and r["line"] == 2 and r["output"] == "1 INF" and r["reference"] == "1 5"
    cmd = bmk.command("sssp", "in.gr", "o.txt", delta=32, source=3)
    assert cmd[1:] == ["-d", "32", "-s", "3", "-o", "o.txt", "in.gr"] and cmd[0].endswith(os.path.join("bin", "sssp"))
    assert bmk.command("bfs", "in.gr", "o.txt", delta=32)[1:] == ["-o", "o.txt", "in.gr"]
    with pytest.raises(ValueError):
        bmk.run("cc", "in.gr")
    json.dumps(r)
