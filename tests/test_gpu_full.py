"""Parity at the full configuration sizes: R-MAT-21 PageRank (20 sweeps) and
symmetrised CC against the C oracle's FP64 power iteration and minimum
reaching id; R-MAT-22 64-source batches, every lane through the oracle's
certificates that pin BFS depths and shortest-path distances exactly
(size-independent properties); a served stream over road1024 + R-MAT-22,
every result through the certificates or the FP64 personalised PageRank and
against the same query answered alone."""
import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import serve, workloads

pytestmark = pytest.mark.gpu


def rel_l1(x, ref):
    return float(np.abs(x.astype(np.float64) - ref).sum() / np.abs(ref).sum())


@pytest.fixture(scope="module")
def rmat21():
    el = workloads.rmat(21)
    return el, gas.Graph(el)


@pytest.fixture(scope="module")
def rmat22s():
    el = workloads.rmat(22, symmetrize=True, weights=True)
    return el, gas.Graph(el)


def test_rmat21_pagerank_full(gpu, rmat21):
    el, g = rmat21
    s, d, _ = el.arrays()
    n = el.n_vertices
    g.autotune(gas.PROGRAM_PAGERANK)
    rank, res, st = g.pagerank(20)
    ref64 = oracle.check_pagerank_fp64(n, s, d, 20)
    assert rel_l1(rank, ref64) <= 1e-5
    assert np.max(np.abs(rank - ref64) / ref64) <= 1e-5
    r19 = oracle.check_pagerank_fp64(n, s, d, 19)
    assert abs(res[-1] - np.abs(ref64 - r19).sum()) <= 1e-3 * max(1.0, res[-1])


def test_rmat21s_cc_full(gpu):
    el = workloads.rmat(21, symmetrize=True)
    s, d, _ = el.arrays()
    g = gas.Graph(el)
    g.autotune(gas.PROGRAM_CC)
    lab, _ = g.cc()
    assert np.array_equal(lab, oracle.check_min_reaching(el.n_vertices, s, d))


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_rmat22s_batch_full(gpu, rmat22s, kind):
    el, g = rmat22s
    s, d, w = el.arrays()
    n = el.n_vertices
    deg = np.bincount(s, minlength=n)
    rng = np.random.default_rng(workloads.SEED_SOURCES)
    src = rng.choice(np.nonzero(deg > 0)[0], 64, replace=False).astype(np.int32)
    b = gas.Batch(g, kind)
    b.add(np.arange(64), src)
    b.run()
    res = b.results(range(64))
    # every ninth lane against the plain engine, every lane by certificate
    buf = gas.DeviceBuffer(4 * n)
    for lane in range(0, 64, 9):
        if kind == gas.BATCH_BFS:
            g.bfs_device(int(src[lane]), buf.ptr)
        else:
            g.sssp_device(int(src[lane]), buf.ptr)
        assert np.array_equal(res[lane], buf.to_numpy()), lane
    import concurrent.futures as cf

    def cert(lane):
        if kind == gas.BATCH_BFS:
            return oracle.cert_bfs(n, s, d, int(src[lane]), res[lane])
        return oracle.cert_sssp(n, s, d, w, int(src[lane]), res[lane])

    with cf.ThreadPoolExecutor(8) as ex:
        ok = list(ex.map(cert, range(64)))
    assert all(ok), [lane for lane in range(64) if not ok[lane]]


def test_served_full_size(gpu, rmat22s):
    # a served stream over both full-size graphs: every result checked by
    # the oracle (BFS / SSSP certificates that fix the answer exactly, PPR
    # within 1e-5 rel L1 of the FP64 power iteration) and against the same
    # query answered alone by the plain engine
    import concurrent.futures as cf
    el22, g22 = rmat22s
    el_road = workloads.road()
    road = gas.Graph(el_road)
    graphs = [road, g22]
    arrays = [el_road.arrays(), el22.arrays()]
    nvs = [el_road.n_vertices, el22.n_vertices]
    srv = gas.Server(graphs, memory_budget=32 << 30)
    st = serve.make_stream(2000, 48, graphs, seed=77)
    keep = list(range(len(st["kinds"])))
    lat, results, stats = srv.replay(st["kinds"], st["graph_ids"], st["sources"],
                                     st["arrivals"], keep=keep)
    assert np.all(lat > 0)
    queries = list(zip(st["kinds"], st["graph_ids"], st["sources"]))
    for i, (k, gi, src) in enumerate(queries):
        gg = graphs[gi]
        if k == gas.BATCH_BFS:
            assert np.array_equal(results[i], gg.bfs(int(src))[0]), i
        elif k == gas.BATCH_SSSP:
            assert np.array_equal(results[i], gg.sssp(int(src))[0]), i
        else:
            plain, _ = gg.ppr(int(src), 10)
            assert rel_l1(results[i], plain.astype(np.float64)) <= 1e-5, i

    def check(i):
        k, gi, src = queries[i]
        s, d, w = arrays[gi]
        n = nvs[gi]
        if k == gas.BATCH_BFS:
            return oracle.cert_bfs(n, s, d, int(src), results[i])
        if k == gas.BATCH_SSSP:
            return oracle.cert_sssp(n, s, d, w, int(src), results[i])
        ref = oracle.check_ppr_fp64(n, s, d, int(src), 10)
        return rel_l1(results[i], ref) <= 1e-5

    with cf.ThreadPoolExecutor(8) as ex:
        ok = list(ex.map(check, range(len(queries))))
    bad = [(i, queries[i]) for i, good in enumerate(ok) if not good]
    assert not bad, bad
    kinds = {int(k) for k, _, _ in queries}
    assert kinds == {gas.BATCH_BFS, gas.BATCH_SSSP, gas.BATCH_PPR}
