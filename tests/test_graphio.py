"""Graph files and result files (no GPU).

The loaders follow the reference's reading rules (graphio.cpp:69-473), the
writers its formats (graphio.cpp:365-439), the converters mtx2gr.cpp /
gr2mtx.cpp, and the result checks regressions/checkDist.py and
checkPageRank.py.  Formats are pinned by files built here byte by byte
(struct) from the layouts the reference documents, by round trips, and by
the known answers of the reference's committed fixtures (small.mtx:
BFS [0,1,1,1,2,2,2], SSSP [0,1,4,1,2,2,2]) reached through every format.
"""
import gzip
import os
import shutil
import struct

import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import convert, resultio

from conftest import GOLDEN

SMALL = os.path.join(GOLDEN, "small.mtx")


def edges_sorted(el):
    s, d, w = el.arrays()
    cols = [s, d] if w is None else [s, d, w]
    return sorted(zip(*[c.tolist() for c in cols]))


def gr_bytes(n, ends, dsts, values=None):
    """A Lonestar binary CSR v1 file as the reference lays it out."""
    b = struct.pack("<4Q", 1, 4 if values is not None else 0, n, len(dsts))
    b += struct.pack("<%dQ" % n, *ends)
    b += struct.pack("<%dI" % len(dsts), *dsts)
    if len(dsts) % 2:
        b += struct.pack("<I", 0)
    if values is not None:
        b += struct.pack("<%dI" % len(values), *values)
    return b


# ---- .edge -----------------------------------------------------------------
def test_edge_loader_rules(tmp_path):
    p = tmp_path / "g.edge"
    p.write_text("# SNAP header\n0 1\n\n1 2\n# comment\n2 0\n7 7\n  \n3 1\n")
    el = gas.Edges.load_edge(str(p))
    s, d, w = el.arrays()
    assert w is None
    assert list(zip(s.tolist(), d.tolist())) == [(0, 1), (1, 2), (2, 0), (3, 1)]
    # the dropped self loop 7 -> 7 still counts toward the largest id
    # (graphio.cpp:133-141 without edge values)
    assert el.n_vertices == 8
    el2 = gas.Edges.load(str(p))
    assert edges_sorted(el2) == edges_sorted(el)


def test_edge_loader_errors(tmp_path):
    p = tmp_path / "bad.edge"
    p.write_text("0 1\nx y\n")
    with pytest.raises(gas.GasError, match="line 2"):
        gas.Edges.load_edge(str(p))
    with pytest.raises(gas.GasError, match="cannot open"):
        gas.Edges.load_edge(str(tmp_path / "missing.edge"))


# ---- .gr ---------------------------------------------------------------------
def test_gr_reader_on_handmade_file(tmp_path):
    # 4 vertices; out-edges 0->{1,2}, 1->{}, 2->{3}; 3 edges (odd: padded)
    p = tmp_path / "g.gr"
    p.write_bytes(gr_bytes(4, [2, 2, 3, 3], [1, 2, 3], [7, 8, 9]))
    el = gas.Edges.load_gr(str(p), True)
    s, d, w = el.arrays()
    assert el.n_vertices == 4
    assert s.tolist() == [0, 0, 2] and d.tolist() == [1, 2, 3] and w.tolist() == [7, 8, 9]
    s, d, w = gas.Edges.load_gr(str(p), False).arrays()
    assert w is None and d.tolist() == [1, 2, 3]


def test_gr_writer_layout(tmp_path):
    # list order within a source is kept; sources sorted
    el = gas.Edges.from_arrays(5, [3, 0, 3, 1, 0], [1, 4, 0, 2, 2], [5, 6, 7, 8, 9])
    p = tmp_path / "w.gr"
    el.write_gr(str(p))
    assert p.read_bytes() == gr_bytes(5, [2, 3, 3, 5, 5], [4, 2, 2, 1, 0],
                                      [6, 9, 8, 5, 7])
    el_nv = gas.Edges.from_arrays(3, [2, 0], [0, 1])
    el_nv.write_gr(str(p))
    assert p.read_bytes() == gr_bytes(3, [1, 1, 2], [1, 0])


def test_gr_reader_errors(tmp_path):
    p = tmp_path / "nv.gr"
    p.write_bytes(gr_bytes(2, [1, 1], [1]))
    # a file without values: an error only when values are wanted
    # (graphio.cpp:300-304)
    with pytest.raises(gas.GasError, match="does not have edge values"):
        gas.Edges.load_gr(str(p), True)
    assert gas.Edges.load_gr(str(p), -1).arrays()[2] is None
    full = gr_bytes(3, [1, 2, 2], [1, 2], [1, 1])
    p.write_bytes(full[:-6])
    with pytest.raises(gas.GasError, match="error reading"):
        gas.Edges.load_gr(str(p), True)
    p.write_bytes(gr_bytes(3, [2, 1, 2], [1, 2]))  # offsets go backwards
    with pytest.raises(gas.GasError, match="bad offset"):
        gas.Edges.load_gr(str(p))
    p.write_bytes(gr_bytes(2, [1, 2], [1, 5]))  # destination out of range
    with pytest.raises(gas.GasError, match="out of range"):
        gas.Edges.load_gr(str(p))
    b = bytearray(gr_bytes(2, [1, 1], [1], [3]))
    b[8:16] = struct.pack("<Q", 8)  # 8-byte edge values
    p.write_bytes(bytes(b))
    with pytest.raises(gas.GasError, match="8 bytes wide"):
        gas.Edges.load_gr(str(p), True)


# ---- .mtx writer, dispatch, gzip --------------------------------------------
def test_mtx_writer_text(tmp_path):
    p = tmp_path / "w.mtx"
    gas.Edges.from_arrays(4, [0, 3], [2, 1], [10, 20]).write_mtx(str(p))
    # one '%' in the banner, as the reference's writer prints it
    assert p.read_text() == ("%MatrixMarket matrix coordinate Integer general\n"
                             "4 4 2\n1 3 10\n4 2 20\n")
    gas.Edges.from_arrays(3, [2], [0]).write_mtx(str(p))
    assert p.read_text() == "%MatrixMarket matrix coordinate pattern general\n3 3 1\n3 1\n"


def test_dispatch_and_gzip(tmp_path):
    ref = gas.Edges.load_mtx(SMALL, True)
    gz = tmp_path / "small.mtx.gz"
    with open(SMALL, "rb") as f, gzip.open(gz, "wb") as g:
        shutil.copyfileobj(f, g)
    for path in (SMALL, str(gz)):
        el = gas.Edges.load(path, True)
        assert el.n_vertices == 7 and edges_sorted(el) == edges_sorted(ref)
    gr = tmp_path / "small.gr"
    ref.write_gr(str(gr))
    grz = tmp_path / "small.gr.gz"
    with open(gr, "rb") as f, gzip.open(grz, "wb") as g:
        shutil.copyfileobj(f, g)
    assert edges_sorted(gas.Edges.load(str(grz), True)) == edges_sorted(ref)
    with pytest.raises(gas.GasError, match="unrecognized filetype"):
        gas.Edges.load(str(tmp_path / "small.txt"))


# ---- converters, reaching the fixtures' known answers -----------------------
def test_converters_keep_known_answers(tmp_path):
    gr = str(tmp_path / "small.gr")
    back = str(tmp_path / "back.mtx")
    convert.main([SMALL, gr])
    convert.main([gr, back])
    ref = gas.Edges.load_mtx(SMALL, True)
    for path in (gr, back):
        el = gas.Edges.load(path, True)
        assert el.n_vertices == 7 and edges_sorted(el) == edges_sorted(ref)
        s, d, w = el.arrays()
        assert oracle.bfs(7, s, d, 0)[0].tolist() == [0, 1, 1, 1, 2, 2, 2]
        assert oracle.sssp(7, s, d, w, 0)[0].tolist() == [0, 1, 4, 1, 2, 2, 2]


def test_converter_symmetric_mtx(tmp_path):
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 2\n2 1 1.5\n3 1 -2.7\n")
    gr = str(tmp_path / "sym.gr")
    convert.convert(str(p), gr)
    # both directions, values truncated toward zero (float -> int)
    assert edges_sorted(gas.Edges.load_gr(gr, True)) == [
        (0, 1, 1), (0, 2, -2), (1, 0, 1), (2, 0, -2)]


# ---- result files and checks -------------------------------------------------
def test_result_file_format(tmp_path):
    assert resultio.format_results(np.array([0, -1, 2], np.int32)) == "0 0\n1 -1\n2 2\n"
    assert resultio.format_results(np.array([0.15, 1.0], np.float32)) == \
        "0 0.150000\n1 1.000000\n"
    p = str(tmp_path / "r.txt")
    resultio.write_results(p, np.array([3, resultio.SSSP_UNREACHED]))
    ids, vals = resultio.read_results(p)
    assert ids.tolist() == [0, 1] and vals.tolist() == [3, resultio.SSSP_UNREACHED]


def test_check_dist(tmp_path):
    t, g = str(tmp_path / "t"), str(tmp_path / "g")

    def run(test_lines, gold_lines):
        open(t, "w").write("".join(f"{a} {b}\n" for a, b in test_lines))
        open(g, "w").write("".join(f"{a} {b}\n" for a, b in gold_lines))
        return resultio.check_dist(t, g)

    assert run([(0, 0), (1, 2)], [(0, 0), (1, 2)])[0]
    # gold's unreached 10000000 equals the engine's 2147383647
    assert run([(0, 0), (1, resultio.SSSP_UNREACHED)], [(0, 0), (1, 10000000)])[0]
    assert not run([(0, 0), (1, 10000000)], [(0, 0), (1, resultio.SSSP_UNREACHED)])[0]
    ok, msg = run([(0, 0), (1, 3)], [(0, 0), (1, 2)])
    assert not ok and msg == "1 gold=2 ref=3"
    ok, msg = run([(0, 0), (2, 1)], [(0, 0), (1, 1), (2, 1)])
    assert not ok and msg == "vertex 1 is in gold but not in test"
    # test vertices the gold lacks are skipped
    assert run([(0, 0), (1, 5), (2, 1)], [(0, 0), (2, 1)])[0]


def test_check_pagerank_bins(tmp_path):
    t, g = str(tmp_path / "t"), str(tmp_path / "g")
    n = 1000
    gold = np.full(n, 1.0)
    resultio.write_results(g, gold)

    def run(test):
        resultio.write_results(t, test)
        return resultio.check_pagerank(t, g)[0]

    assert run(gold + 5e-5)  # everything below 1e-4
    x = gold.copy()
    x[:10] += 5e-4           # 1.0% in [1e-4, 1e-3): allowed 1%
    assert run(x)
    x[10] += 5e-4            # 1.1%
    assert not run(x)
    x = gold.copy()
    x[:5] += 5e-3            # 0.5% in [1e-3, 1e-2)
    assert run(x)
    x[5] += 5e-3
    assert not run(x)
    x = gold.copy()
    x[0] += 20.0             # any difference >= 10 fails
    assert not run(x)
    assert resultio.main(["pagerank", g, g]) == 0
    resultio.write_results(t, gold + 0.5)
    assert resultio.main(["pagerank", t, g]) == 1


# ---- against the reference's own graphio.cpp (oracle/_ref) ------------------
needs_refio = pytest.mark.skipif(not oracle.refio_available(),
                                 reason="oracle/_ref/librefgraphio.so not built "
                                        "(reference sources absent)")

TEXT_FILES = {
    "sym.mtx": "%%MatrixMarket matrix coordinate real symmetric\n% c\n4 4 4\n"
               "2 1 1.5\n3 1 -2.7\n3 3 9\n4 2 7.9\n",
    "skew.mtx": "%%MatrixMarket matrix coordinate integer skew-symmetric\n4 4 3\n"
                "2 1 3\n4 3 -5\n\n4 1 8\n",
    "loop.mtx": "%%MatrixMarket matrix coordinate integer general\n9 9 3\n"
                "1 2 4\n9 9 1\n2 3 6\n",
    "g.edge": "# header\n0 1\n1 2\n\n2 0\n7 7\n3 1\n",
}


@needs_refio
@pytest.mark.parametrize("want_values", [False, True])
def test_loaders_match_reference(tmp_path, want_values):
    paths = [SMALL, os.path.join(GOLDEN, "trivialGraph.mtx"),
             os.path.join(GOLDEN, "chesapeake.mtx")]
    for name, text in TEXT_FILES.items():
        (tmp_path / name).write_text(text)
        paths.append(str(tmp_path / name))
    gz = tmp_path / "small.mtx.gz"
    with open(SMALL, "rb") as f, gzip.open(gz, "wb") as g:
        shutil.copyfileobj(f, g)
    paths.append(str(gz))
    # a .gr with values written by gas_write_gr, read back by the reference
    gr = str(tmp_path / "small.gr")
    gas.Edges.load_mtx(SMALL, True).write_gr(gr)
    paths.append(gr)
    for path in paths:
        nv, s, d, w = oracle.refio_load(path, want_values)
        el = gas.Edges.load(path, want_values)
        ms, md, mw = el.arrays()
        assert el.n_vertices == nv, path
        assert ms.tolist() == s.tolist() and md.tolist() == d.tolist(), path
        if w is None:
            assert mw is None, path
        else:
            assert mw is not None and mw.tolist() == w.tolist(), path


@needs_refio
def test_writers_match_reference(tmp_path):
    el = gas.Edges.from_arrays(6, [3, 0, 3, 1, 0, 5], [1, 4, 0, 2, 2, 0],
                               [5, 6, -7, 8, 9, 1])
    s, d, w = el.arrays()
    # MatrixMarket text, byte for byte
    mine, theirs = tmp_path / "m.mtx", tmp_path / "r.mtx"
    el.write_mtx(str(mine))
    oracle.refio_write_mtx(str(theirs), 6, s, d, w)
    assert mine.read_bytes() == theirs.read_bytes()
    # binary CSR from the same CSR arrays (edges of a source in list order)
    order = np.argsort(s, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(s, minlength=6))])
    mine, theirs = tmp_path / "m.gr", tmp_path / "r.gr"
    el.write_gr(str(mine))
    oracle.refio_write_gr(str(theirs), 6, off, d[order], w[order])
    assert mine.read_bytes() == theirs.read_bytes()
