"""Parity at the full configuration sizes: R-MAT-21 PageRank (20 sweeps) and
symmetrised CC against the C oracle's FP64 power iteration and minimum
reaching id; R-MAT-22 64-source batches through certificates that pin BFS
depths and shortest-path distances exactly (size-independent properties);
a served stream over road1024 + R-MAT-22 against the plain engine."""
import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import serve, workloads

pytestmark = pytest.mark.gpu


def rel_l1(x, ref):
    return float(np.abs(x.astype(np.float64) - ref).sum() / np.abs(ref).sum())


def bfs_certificate(n, s, d, src, depth):
    """True iff depth is the BFS depth from src: depth[src] = 0; along every
    edge from a reached vertex the head is reached with depth <= tail + 1;
    every reached vertex but src has an in-edge from depth - 1."""
    if depth[src] != 0:
        return False
    ds, dt = depth[s], depth[d]
    reached = ds >= 0
    if np.any(dt[reached] < 0) or np.any(dt[reached] > ds[reached] + 1):
        return False
    tight = reached & (dt == ds + 1)
    has_parent = np.zeros(n, bool)
    has_parent[d[tight]] = True
    need = depth >= 0
    need[src] = False
    return bool(np.all(has_parent[need]))


def sssp_certificate(n, s, d, w, src, dist):
    """True iff dist is the shortest-path distance from src (positive
    weights): dist[src] = 0; no edge from a reached vertex can be relaxed;
    every reached vertex but src has a tight in-edge."""
    inf = oracle.SSSP_UNREACHED
    if dist[src] != 0:
        return False
    ds = dist[s].astype(np.int64)
    dt = dist[d].astype(np.int64)
    reached = ds < inf
    if np.any(dt[reached] > ds[reached] + w[reached]):
        return False
    tight = reached & (dt == ds + w)
    has_parent = np.zeros(n, bool)
    has_parent[d[tight]] = True
    need = dist < inf
    need[src] = False
    return bool(np.all(has_parent[need]))


def test_certificates_catch_mistakes():
    # the certificates themselves: exact answers pass, perturbed ones fail
    el = workloads.road(32, 32, seed=3)
    s, d, w = el.arrays()
    n = el.n_vertices
    depth = oracle.check_bfs(n, s, d, 5)
    dist = oracle.check_dijkstra(n, s, d, w, 5)
    assert bfs_certificate(n, s, d, 5, depth)
    assert sssp_certificate(n, s, d, w, 5, dist)
    bad = depth.copy()
    bad[np.argmax(depth)] += 1
    assert not bfs_certificate(n, s, d, 5, bad)
    bad = dist.copy()
    bad[np.argmax(dist)] -= 1
    assert not sssp_certificate(n, s, d, w, 5, bad)


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
    # every lane against the plain engine, two lanes by certificate
    buf = gas.DeviceBuffer(4 * n)
    for lane in range(0, 64, 9):
        if kind == gas.BATCH_BFS:
            g.bfs_device(int(src[lane]), buf.ptr)
        else:
            g.sssp_device(int(src[lane]), buf.ptr)
        assert np.array_equal(res[lane], buf.to_numpy()), lane
    for lane in (0, 63):
        if kind == gas.BATCH_BFS:
            assert bfs_certificate(n, s, d, int(src[lane]), res[lane]), lane
        else:
            assert sssp_certificate(n, s, d, w, int(src[lane]), res[lane]), lane


def test_served_full_size(gpu, rmat22s):
    el, g22 = rmat22s
    road = gas.Graph(workloads.road())
    graphs = [road, g22]
    srv = gas.Server(graphs, memory_budget=32 << 30)
    st = serve.make_stream(2000, 48, graphs, seed=77)
    keep = list(range(len(st["kinds"])))
    lat, results, stats = srv.replay(st["kinds"], st["graph_ids"], st["sources"],
                                     st["arrivals"], keep=keep)
    assert np.all(lat > 0)
    for i, (k, gi, src) in enumerate(zip(st["kinds"], st["graph_ids"], st["sources"])):
        gg = graphs[gi]
        if k == gas.BATCH_BFS:
            assert np.array_equal(results[i], gg.bfs(int(src))[0]), i
        elif k == gas.BATCH_SSSP:
            assert np.array_equal(results[i], gg.sssp(int(src))[0]), i
        else:
            plain, _ = gg.ppr(int(src), 10)
            assert rel_l1(results[i], plain.astype(np.float64)) <= 1e-5, i
