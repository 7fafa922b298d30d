"""Pins the CPU oracle (oracle/gasref.c, the C restatement of refgas.h).

Each restated function is checked against something it does not share code
with:
  * the reference's own engine (refgas.h + util.h compiled from its sources
    into oracle/_ref) on the same inputs;
  * hand-derived known answers on the reference's committed fixtures
    (test-graphs/small.mtx, trivialGraph.mtx, chesapeake.mtx);
  * closed forms (grid BFS depth = r + c);
  * independent textbook algorithms (queue BFS, Dijkstra, minimum reaching
    id, FP64 power iteration) and brute force on tiny graphs;
  * planted mistakes in a copy of the oracle's inputs must be caught.
"""
import os

import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import workloads

from conftest import GOLDEN

needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def random_graph(rng, n, m, wmax=100):
    s = rng.integers(0, n, m).astype(np.int32)
    d = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(1, wmax, m).astype(np.int32)
    return s, d, w


# ---- known answers on the reference's fixtures ----------------------------
def test_fixture_small():
    el = gas.Edges.load_mtx(os.path.join(GOLDEN, "small.mtx"), True)
    s, d, w = el.arrays()
    assert (el.n_vertices, el.n_edges) == (7, 11)
    assert oracle.bfs(7, s, d, 0)[0].tolist() == [0, 1, 1, 1, 2, 2, 2]
    assert oracle.sssp(7, s, d, w, 0)[0].tolist() == [0, 1, 4, 1, 2, 2, 2]
    assert oracle.cc(7, s, d)[0].tolist() == [0] * 7


def test_fixture_trivial():
    el = gas.Edges.load_mtx(os.path.join(GOLDEN, "trivialGraph.mtx"), True)
    s, d, w = el.arrays()
    assert (el.n_vertices, el.n_edges) == (5, 5)
    assert oracle.bfs(5, s, d, 0)[0].tolist() == [0, 1, 2, 3, 4]
    assert oracle.sssp(5, s, d, w, 0)[0].tolist() == [0, 1, 2, 3, 4]
    assert oracle.cc(5, s, d)[0].tolist() == [0] * 5


def test_fixture_chesapeake():
    el = gas.Edges.load_mtx(os.path.join(GOLDEN, "chesapeake.mtx"), False)
    s, d, _ = el.arrays()
    assert (el.n_vertices, el.n_edges) == (39, 340)
    src, deg = el.max_outdegree_vertex()
    assert src == 38
    depth = oracle.bfs(39, s, d, src)[0]
    want = [1] * 33 + [2] * 5 + [0]
    assert depth.tolist() == want
    assert oracle.cc(39, s, d)[0].tolist() == [0] * 39


def test_grid_closed_form():
    el = workloads.grid64()
    s, d, _ = el.arrays()
    depth, st = oracle.bfs(4096, s, d, 0)
    r, c = np.divmod(np.arange(4096), 64)
    assert np.array_equal(depth, r + c)
    assert st.supersteps == 127 + 1  # levels 0..126 plus the empty last scatter
    # SSSP with unit weights equals BFS depth
    el2 = gas.Edges.grid(64, 64)
    s2, d2, w2 = el2.arrays()
    assert np.array_equal(oracle.sssp(4096, s2, d2, w2, 0)[0], r + c)


# ---- against the reference's own engine -----------------------------------
@needs_ref
def test_csr_matches_reference_layout():
    rng = np.random.default_rng(1)
    s, d, _ = random_graph(rng, 500, 3000)
    off, dst, eid = oracle.edge_list_to_csr(500, s, d)
    roff, rdst, reid = oracle.ref_edge_list_to_csr(500, s, d)
    assert np.array_equal(off, roff)
    # within a vertex the reference's sort order is unspecified: compare sets
    for v in range(500):
        a, b = slice(off[v], off[v + 1]), slice(roff[v], roff[v + 1])
        assert sorted(dst[a]) == sorted(rdst[b])
        assert sorted(eid[a]) == sorted(reid[b])
    # the oracle's order is the stable one: edge indices ascend within a key
    for v in range(500):
        assert np.all(np.diff(eid[off[v]:off[v + 1]]) > 0)
    # CSC is CSR of the reversed edges (util.h:114-121)
    coff, csrc, ceid = oracle.edge_list_to_csc(500, s, d)
    roff2, rsrc, reid2 = oracle.edge_list_to_csr(500, d, s)
    assert np.array_equal(coff, roff2) and np.array_equal(csrc, rsrc)
    assert np.array_equal(ceid, reid2)
    assert np.array_equal(coff, np.concatenate([[0], np.cumsum(np.bincount(d, minlength=500))]))


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_engine_matches_reference_random(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    m = int(rng.integers(0, 5 * n))
    s, d, w = random_graph(rng, n, m)
    src = int(rng.integers(0, n))
    assert np.array_equal(oracle.bfs(n, s, d, src)[0], oracle.ref_bfs(n, s, d, src)[0])
    assert np.array_equal(oracle.sssp(n, s, d, w, src)[0],
                          oracle.ref_sssp(n, s, d, w, src)[0])
    assert np.array_equal(oracle.cc(n, s, d)[0], oracle.ref_cc(n, s, d)[0])
    for iters in (0, 7):
        mine = oracle.pagerank(n, s, d, iters)[0]
        theirs = oracle.ref_pagerank(n, s, d, iters)[0]
        # same arithmetic, possibly a different order within a vertex's
        # in-edges (unstable sort): within the reference's 1e-6 (pagerank.cu:156)
        assert np.max(np.abs(mine - theirs), initial=0) <= 1e-6 * max(1.0, float(np.max(theirs, initial=1)))


@needs_ref
def test_engine_matches_reference_road():
    el = workloads.road(256, 256, seed=9)
    s, d, w = el.arrays()
    n = el.n_vertices
    assert np.array_equal(oracle.bfs(n, s, d, 0)[0], oracle.ref_bfs(n, s, d, 0)[0])
    assert np.array_equal(oracle.sssp(n, s, d, w, 0)[0], oracle.ref_sssp(n, s, d, w, 0)[0])


@needs_ref
def test_engine_matches_reference_rmat_cc():
    el = workloads.rmat(12, symmetrize=True)
    s, d, _ = el.arrays()
    n = el.n_vertices
    assert np.array_equal(oracle.cc(n, s, d)[0], oracle.ref_cc(n, s, d)[0])


# ---- against independent algorithms ----------------------------------------
@pytest.mark.parametrize("seed", range(5))
def test_engine_matches_textbook(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 600))
    m = int(rng.integers(0, 4 * n))
    s, d, w = random_graph(rng, n, m)
    src = int(rng.integers(0, n))
    assert np.array_equal(oracle.bfs(n, s, d, src)[0], oracle.check_bfs(n, s, d, src))
    assert np.array_equal(oracle.sssp(n, s, d, w, src)[0],
                          oracle.check_dijkstra(n, s, d, w, src))
    assert np.array_equal(oracle.cc(n, s, d)[0], oracle.check_min_reaching(n, s, d))
    r32 = oracle.pagerank(n, s, d, 10)[0]
    r64 = oracle.check_pagerank_fp64(n, s, d, 10)
    assert np.allclose(r32, r64, rtol=1e-5, atol=1e-6)


def brute_force_sssp(n, s, d, w, src):
    # Bellman-Ford to a fixed point in exact integers
    inf = 2 ** 62
    dist = [inf] * n
    dist[src] = 0
    for _ in range(n):
        for a, b, c in zip(s.tolist(), d.tolist(), w.tolist()):
            if dist[a] + c < dist[b]:
                dist[b] = dist[a] + c
    return [x if x < inf else oracle.SSSP_UNREACHED for x in dist]


def test_brute_force_tiny_graphs():
    # every directed graph on 3 vertices with weights in {1, 2}
    pairs = [(a, b) for a in range(3) for b in range(3) if a != b]
    for mask in range(1 << len(pairs)):
        es = [p for i, p in enumerate(pairs) if mask >> i & 1]
        s = np.array([p[0] for p in es], np.int32)
        d = np.array([p[1] for p in es], np.int32)
        w = np.array([1 + (i % 2) for i in range(len(es))], np.int32)
        for src in range(3):
            assert oracle.sssp(3, s, d, w, src)[0].tolist() == brute_force_sssp(3, s, d, w, src)


def test_pagerank_closed_form_cycle():
    # directed n-cycle: every vertex has outdeg 1; r_k = 0.15 * sum_{j<=k} 0.85^j
    n = 9
    s = np.arange(n, dtype=np.int32)
    d = (s + 1) % n
    for k in (1, 5, 20):
        want = 0.15 * sum(0.85 ** j for j in range(k + 1))
        assert np.allclose(oracle.check_pagerank_fp64(n, s, d, k), want, rtol=1e-14)
        assert np.allclose(oracle.pagerank(n, s, d, k)[0], want, rtol=1e-6)


def test_ppr_closed_form_star():
    # s -> leaves, leaves -> s: two-step recurrence
    n = 5
    s = np.array([0, 0, 0, 0, 1, 2, 3, 4], np.int32)
    d = np.array([1, 2, 3, 4, 0, 0, 0, 0], np.int32)
    r = np.zeros(n)
    r[0] = 0.15
    for _ in range(6):
        nxt = np.zeros(n)
        nxt[1:] = r[0] / 4
        nxt[0] = r[1:].sum()
        r = 0.85 * nxt
        r[0] += 0.15
    assert np.allclose(oracle.check_ppr_fp64(n, s, d, 0, 6), r, rtol=1e-14)


# ---- planted mistakes must be caught ---------------------------------------
def test_checks_catch_planted_mistakes():
    el = workloads.road(64, 64, seed=2)
    s, d, w = el.arrays()
    n = el.n_vertices
    truth = oracle.check_dijkstra(n, s, d, w, 0)
    assert np.array_equal(oracle.sssp(n, s, d, w, 0)[0], truth)
    # an off-by-one in the edge term (w + 1) would be caught by Dijkstra
    assert not np.array_equal(oracle.sssp(n, s, d, w + 1, 0)[0], truth)
    # a wrong seed (0 instead of gatherZero) would be caught on unreached ids
    el2 = gas.Edges.from_arrays(4, [0, 1], [1, 2], [3, 4])
    s3, d3, w3 = el2.arrays()
    assert oracle.sssp(4, s3, d3, w3, 0)[0].tolist() == [0, 3, 7, oracle.SSSP_UNREACHED]
    # a transposed operand (edges reversed) is caught on a directed graph
    rng = np.random.default_rng(3)
    s2, d2, _ = random_graph(rng, 200, 600)
    assert not np.array_equal(oracle.bfs(200, s2, d2, 0)[0],
                              oracle.check_bfs(200, d2, s2, 0))


def _power_iteration_np(n, s, d, iters, damping=0.85, outdeg_of=None, sweeps_less=0):
    """A numpy power iteration of pagerank.cu:50, 62 with knobs for planted
    mistakes: damping, whose degree divides a contribution, sweeps run."""
    outdeg = np.bincount(s if outdeg_of is None else outdeg_of, minlength=n)
    r = np.full(n, 0.15)
    for _ in range(iters - sweeps_less):
        contrib = r[s] / np.maximum(outdeg[s], 1)
        r = 0.15 + damping * np.bincount(d, weights=contrib, minlength=n)
    return r


def test_pagerank_checks_catch_planted_mistakes():
    # the FP64 checker and the 1e-5 rel-L1 rule of the GPU tests: a numpy
    # power iteration matches the checker to rounding, and each planted
    # mistake in it (damping 0.85 -> 0.851, one sweep short, the
    # destination's degree, reversed edges) lands outside the rule
    rng = np.random.default_rng(5)
    n = 300
    s, d, _ = random_graph(rng, n, 2400)
    r64 = oracle.check_pagerank_fp64(n, s, d, 20)

    def rel(x):
        return float(np.abs(x - r64).sum() / np.abs(r64).sum())

    assert rel(_power_iteration_np(n, s, d, 20)) <= 1e-13
    assert rel(_power_iteration_np(n, s, d, 20, damping=0.851)) > 1e-5
    assert rel(_power_iteration_np(n, s, d, 20, sweeps_less=1)) > 1e-5
    assert rel(_power_iteration_np(n, s, d, 20, outdeg_of=d)) > 1e-5
    assert rel(_power_iteration_np(n, d, s, 20)) > 1e-5
    # the FP32 restatement of the engine passes the same rule
    assert rel(oracle.pagerank(n, s, d, 20)[0].astype(np.float64)) <= 1e-5


def test_certificates_catch_mistakes():
    # the one-pass BFS / SSSP certificates: exact answers pass; a depth or
    # distance moved by one at any reached vertex, an unreached vertex
    # marked reached, a reached one marked unreached and a wrong source fail
    for el in (workloads.road(32, 32, seed=3),
               workloads.rmat(10, symmetrize=True, weights=True)):
        s, d, w = el.arrays()
        n = el.n_vertices
        src = int(el.max_outdegree_vertex()[0])
        depth = oracle.check_bfs(n, s, d, src)
        dist = oracle.check_dijkstra(n, s, d, w, src)
        assert oracle.cert_bfs(n, s, d, src, depth)
        assert oracle.cert_sssp(n, s, d, w, src, dist)
        assert not oracle.cert_bfs(n, s, d, (src + 1) % n, depth)
        assert not oracle.cert_sssp(n, s, d, w, (src + 1) % n, dist)
        rng = np.random.default_rng(n)
        reached = np.nonzero((depth > 0))[0]
        for v in rng.choice(reached, min(40, len(reached)), replace=False):
            for delta in (-1, 1):
                bad = depth.copy()
                bad[v] += delta
                assert not oracle.cert_bfs(n, s, d, src, bad), (v, delta)
                bad = dist.copy()
                bad[v] += delta
                assert not oracle.cert_sssp(n, s, d, w, src, bad), (v, delta)
            bad = depth.copy()
            bad[v] = -1
            assert not oracle.cert_bfs(n, s, d, src, bad)
            bad = dist.copy()
            bad[v] = oracle.SSSP_UNREACHED
            assert not oracle.cert_sssp(n, s, d, w, src, bad)
        unreached = np.nonzero(depth < 0)[0]
        if len(unreached):
            bad = depth.copy()
            bad[unreached[0]] = 1
            assert not oracle.cert_bfs(n, s, d, src, bad)
            bad = dist.copy()
            bad[unreached[0]] = 1
            assert not oracle.cert_sssp(n, s, d, w, src, bad)
    # a graph with unreached vertices, the certificates against the textbook
    rng = np.random.default_rng(9)
    s, d, w = random_graph(rng, 500, 700)
    for src in (0, 7, 499):
        assert oracle.cert_bfs(500, s, d, src, oracle.check_bfs(500, s, d, src))
        assert oracle.cert_sssp(500, s, d, w, src, oracle.check_dijkstra(500, s, d, w, src))


@needs_ref
def test_reference_tolerance_pagerank_supersteps():
    # the restated engine's tolerance-driven run takes the reference's
    # supersteps (compiled refgas.h run(), pagerank.cu:85-101)
    for sc in (8, 10, 12):
        el = workloads.rmat(sc)
        s, d, _ = el.arrays()
        n = el.n_vertices
        mine, st = oracle.pagerank(n, s, d, 0)
        ref, steps = oracle.ref_pagerank_run(n, s, d)
        assert st.supersteps == steps
        assert np.array_equal(ref, oracle.ref_pagerank(n, s, d, 0)[0])


@needs_ref
def test_pagerank_tol_fp64_matches_reference():
    # the FP64 evaluation of the tolerance-driven run takes the reference
    # engine's supersteps and agrees with its FP32 ranks to FP32 rounding;
    # on a directed n-cycle sweep k moves every rank by 0.15 * 0.85^k, so the
    # run stops after the first sweep whose step is below tol = 0.01:
    # 0.15 * 0.85^k < 0.01 first at k = 17, i.e. 17 supersteps, leaving
    # r = 0.15 * sum_{j<=17} 0.85^j
    rng = np.random.default_rng(23)
    for trial in range(4):
        n = int(rng.integers(2, 800))
        s, d, _ = random_graph(rng, n, int(rng.integers(n, 5 * n)))
        ref, steps = oracle.ref_pagerank_run(n, s, d)
        r64, st64 = oracle.check_pagerank_tol_fp64(n, s, d)
        assert st64 == steps
        assert np.abs(ref - r64).sum() / r64.sum() <= 1e-6
    n = 9
    s = np.arange(n, dtype=np.int32)
    r64, st64 = oracle.check_pagerank_tol_fp64(n, s, (s + 1) % n)
    assert st64 == 17
    assert np.allclose(r64, 0.15 * sum(0.85 ** j for j in range(18)), rtol=1e-14)
    assert oracle.ref_pagerank_run(n, s, (s + 1) % n)[1] == 17


def test_ref_prepared_engines():
    # one reference engine bound once, many queries: each equals the
    # textbook answer (BFS, Dijkstra) or the FP64 personalised PageRank
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    el = workloads.road(48, 48, seed=3)
    s, d, w = el.arrays()
    n = el.n_vertices
    for kind in (0, 1, 2):
        eng = oracle.RefPrepared(kind, n, s, d, w)
        for src in (0, 77, n - 1):
            out, us = eng.query(src, 10)
            assert us >= 0
            if kind == 0:
                assert np.array_equal(out, oracle.check_bfs(n, s, d, src))
            elif kind == 1:
                assert np.array_equal(out, oracle.check_dijkstra(n, s, d, w, src))
            else:
                ref = oracle.check_ppr_fp64(n, s, d, src, 10)
                assert np.abs(out - ref).sum() / np.abs(ref).sum() <= 1e-6
