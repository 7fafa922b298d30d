"""GPU parity of the batched multi-source engine: every lane bit-exact
against the single-source CPU oracle (refgas.h semantics), alone and mixed
with lanes that join a running batch."""
import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import workloads

pytestmark = pytest.mark.gpu


def _sources(rng, el, n):
    s, d, _ = el.arrays()
    deg = np.bincount(s, minlength=el.n_vertices)
    cand = np.nonzero(deg > 0)[0]
    return rng.choice(cand, size=n, replace=False).astype(np.int32)


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_batch_road_64_lanes(gpu, kind):
    el = workloads.road(256, 256, seed=21)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    b = gas.Batch(g, kind)
    src = _sources(np.random.default_rng(1), el, 64)
    b.add(np.arange(64), src)
    st = b.run()
    assert st.supersteps > 0
    for lane in range(0, 64, 7):
        got = b.result(lane)
        if kind == gas.BATCH_BFS:
            want = oracle.check_bfs(n, s, d, int(src[lane]))
        else:
            want = oracle.check_dijkstra(n, s, d, w, int(src[lane]))
        assert np.array_equal(got, want), lane


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_batch_rmat_hubs(gpu, kind):
    el = workloads.rmat(14, symmetrize=True, weights=True)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    b = gas.Batch(g, kind)
    src = _sources(np.random.default_rng(2), el, 64)
    src[0] = 0  # the largest hub
    b.add(np.arange(64), src)
    b.run()
    many = b.results(range(64))  # one extraction pass for all lanes
    for lane in range(64):
        if kind == gas.BATCH_BFS:
            want = oracle.bfs(n, s, d, int(src[lane]))[0]
        else:
            want = oracle.sssp(n, s, d, w, int(src[lane]))[0]
        assert np.array_equal(many[lane], want), lane
    for lane in (0, 17, 63):
        assert np.array_equal(b.result(lane), many[lane])
    # lanes of one, two and four sectors of the distance rows
    for lanes in ([7], [16, 31], [1, 20, 33, 50]):
        part = b.results(lanes)
        for lane in lanes:
            assert np.array_equal(part[lane], many[lane]), lane


@pytest.mark.parametrize("divisor", [0, 1, 4, 64])
def test_batch_bfs_push_pull_mix(gpu, divisor):
    # 0: push only; larger divisors pull once the frontier's out-edges reach
    # E / divisor -- the direction must not change any depth
    el = workloads.rmat(13, symmetrize=True)
    s, d, _ = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    b = gas.Batch(g, gas.BATCH_BFS)
    b.set_pull_divisor(divisor)
    src = _sources(np.random.default_rng(9), el, 64)
    b.add(np.arange(64), src)
    b.run()
    many = b.results(range(0, 64, 3))
    for lane, got in many.items():
        assert np.array_equal(got, oracle.bfs(n, s, d, int(src[lane]))[0]), lane


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_continuous_batching_alone_equals_mixed(gpu, kind):
    el = workloads.road(128, 128, seed=5)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    rng = np.random.default_rng(3)
    src = _sources(rng, el, 40)
    # alone: one lane per batch
    alone = {}
    b1 = gas.Batch(g, kind)
    for i in range(3):
        b1.add([5], [src[i]])
        b1.run()
        alone[i] = b1.result(5)
        b1.release(1 << 5)
    # mixed: queries join while others run, lanes are reused after release
    b = gas.Batch(g, kind)
    b.add(np.arange(8), src[8:16])
    b.step(3)
    b.add([8, 9], src[0:2])
    b.step(2)
    b.add([10], [src[2]])
    done = {}
    while b.alive():
        b.step(1)
    for lane, qi in ((8, 0), (9, 1), (10, 2)):
        done[qi] = b.result(lane)
    for qi in range(3):
        assert np.array_equal(done[qi], alone[qi])
        want = (oracle.check_bfs(n, s, d, int(src[qi])) if kind == gas.BATCH_BFS
                else oracle.check_dijkstra(n, s, d, w, int(src[qi])))
        assert np.array_equal(done[qi], want)
    # reuse a released lane for a new query while the rest stay resident
    b.release(1 << 8)
    b.add([8], [src[20]])
    b.run()
    want = (oracle.check_bfs(n, s, d, int(src[20])) if kind == gas.BATCH_BFS
            else oracle.check_dijkstra(n, s, d, w, int(src[20])))
    assert np.array_equal(b.result(8), want)


@pytest.mark.parametrize("item_edges", [32, 64])
def test_batch_sssp_octets(gpu, item_edges):
    # a low-degree graph (road grid plus one vertex of degree 150): SSSP items
    # per (vertex, octet of lanes with a change, chunk of out-edges); lanes
    # spread over the octets, few per octet, and the hub's items (chunks x
    # octets above the hub threshold) written by the hub pass
    base = workloads.road(64, 64, seed=23)
    s0, d0, w0 = base.arrays()
    n = base.n_vertices + 1
    hub = n - 1
    rng = np.random.default_rng(11)
    nb = rng.choice(hub, size=150, replace=False).astype(np.int32)
    wt = rng.integers(1, 65, size=150).astype(np.int32)
    s = np.concatenate([s0, np.full(150, hub, np.int32), nb])
    d = np.concatenate([d0, nb, np.full(150, hub, np.int32)])
    w = np.concatenate([w0, wt, wt])
    el = gas.Edges.from_arrays(n, s, d, w)
    g = gas.Graph(el)
    b = gas.Batch(g, gas.BATCH_SSSP)
    b.set_item_edges(item_edges)
    b.set_hub_items(16)
    lanes = [0, 3, 12, 27, 40, 41, 55, 63]
    src = [hub, int(nb[0]), 5, 900, hub, 2047, int(nb[1]), 4000]
    b.add(lanes, src)
    b.run()
    got = b.results(lanes)
    for lane, sv in zip(lanes, src):
        assert np.array_equal(got[lane], oracle.check_dijkstra(n, s, d, w, sv)), lane
    # a second wave of queries into the same engine, other octets
    b.release(sum(1 << x for x in lanes))
    lanes2 = [8, 9, 30, 47]
    src2 = [hub, 17, int(nb[2]), 3333]
    b.add(lanes2, src2)
    b.run()
    for lane, sv in zip(lanes2, src2):
        assert np.array_equal(b.result(lane), oracle.check_dijkstra(n, s, d, w, sv)), lane


def test_batch_sssp_octet_many_hubs(gpu):
    # a low-degree graph (average out-degree <= 8, so octet items) with
    # 1,500 vertices of degree ~80: with every octet changed such a vertex
    # has 3 chunks x 8 octets = 24 items, above a hub threshold of 16, so a
    # round can list far more hubs than one per 16 x 32 edges
    rng = np.random.default_rng(77)
    n, nh, k = 40000, 1500, 80
    hubs = rng.choice(n, size=nh, replace=False)
    a = np.repeat(hubs, k)
    b = rng.integers(0, n, size=nh * k)
    keep = a != b
    a, b = a[keep], b[keep]
    w = rng.integers(1, 50, size=len(a))
    s = np.concatenate([a, b]).astype(np.int32)
    d = np.concatenate([b, a]).astype(np.int32)
    w = np.concatenate([w, w]).astype(np.int32)
    assert len(s) <= 8 * n
    g = gas.Graph(gas.Edges.from_arrays(n, s, d, w))
    b_ = gas.Batch(g, gas.BATCH_SSSP)
    b_.set_item_edges(32)
    b_.set_hub_items(16)
    src = rng.choice(hubs, size=64, replace=False).astype(np.int32)
    b_.add(np.arange(64), src)
    b_.run()
    got = b_.results(range(0, 64, 9))
    for lane, res in got.items():
        assert np.array_equal(res, oracle.check_dijkstra(n, s, d, w, int(src[lane]))), lane


def _random_graph(rng, n, avg_deg, hub_deg):
    # symmetric weighted graph: random pairs plus one vertex of high degree
    # (n need not be a multiple of 64: the last slab and depth tile are
    # partial)
    m = n * avg_deg // 2
    a = rng.integers(0, n, size=m)
    b = rng.integers(0, n, size=m)
    hub = int(rng.integers(0, n))
    nb = rng.choice(n, size=min(hub_deg, n - 1), replace=False)
    a = np.concatenate([a, np.full(len(nb), hub)])
    b = np.concatenate([b, nb])
    keep = a != b
    a, b = a[keep], b[keep]
    w = rng.integers(1, 100, size=len(a))
    s = np.concatenate([a, b]).astype(np.int32)
    d = np.concatenate([b, a]).astype(np.int32)
    return s, d, np.concatenate([w, w]).astype(np.int32)


@pytest.mark.parametrize("case", range(16))
def test_batch_randomized(gpu, case):
    # seeded random graphs of low and high average degree (octet SSSP items
    # or vertex-wide ones), random lane subsets, item sizes, hub thresholds
    # and distance steps; every lane against the oracle
    rng = np.random.default_rng(100 + case)
    n = int(rng.choice([999, 1000, 4099, 6000]))
    avg = int(rng.choice([3, 6, 24]))
    s, d, w = _random_graph(rng, n, avg, int(rng.choice([40, 300])))
    el = gas.Edges.from_arrays(n, s, d, w)
    g = gas.Graph(el)
    for kind in (gas.BATCH_BFS, gas.BATCH_SSSP):
        b = gas.Batch(g, kind)
        edges = int(rng.choice([8, 16] if kind == gas.BATCH_BFS else [32, 64]))
        b.set_item_edges(edges)
        b.set_hub_items(int(rng.choice([16, 32, 1 << 20])))
        if kind == gas.BATCH_SSSP:
            b.set_sssp_delta(int(rng.choice([0, 16, 64])))
        k = int(rng.integers(1, 65))
        lanes = np.sort(rng.choice(64, size=k, replace=False)).astype(np.int32)
        src = rng.integers(0, n, size=k).astype(np.int32)
        b.add(lanes, src)
        b.run()
        got = b.results(lanes)
        for lane, sv in zip(lanes, src):
            want = (oracle.check_bfs(n, s, d, int(sv)) if kind == gas.BATCH_BFS
                    else oracle.check_dijkstra(n, s, d, w, int(sv)))
            assert np.array_equal(got[int(lane)], want), (case, kind, int(lane))


def test_batch_errors(gpu):
    el = workloads.road(32, 32, seed=1)
    g = gas.Graph(el)
    b = gas.Batch(g, gas.BATCH_BFS)
    b.add([0], [3])
    with pytest.raises(gas.GasError):
        b.add([0], [4])        # lane in use
    with pytest.raises(gas.GasError):
        b.add([64], [4])       # lane out of range
    with pytest.raises(gas.GasError):
        b.add([1], [10 ** 6])  # source out of range
    nw = gas.Graph(gas.Edges.from_arrays(3, [0, 1], [1, 2]))
    with pytest.raises(gas.GasError):
        gas.Batch(nw, gas.BATCH_SSSP)  # no weights


def rel_l1(x, ref):
    return float(np.abs(x.astype(np.float64) - ref).sum() / np.abs(ref).sum())


def test_ppr_plain_and_batched_vs_fp64(gpu):
    el = workloads.rmat(14, symmetrize=True, weights=True)
    s, d, _ = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    src = _sources(np.random.default_rng(9), el, 64)
    src[1] = 0  # hub source
    # plain engine, one query
    r, _ = g.ppr(int(src[0]), 10)
    ref = oracle.check_ppr_fp64(n, s, d, int(src[0]), 10)
    assert rel_l1(r, ref) <= 1e-5
    # batched, 64 lanes
    b = gas.Batch(g, gas.BATCH_PPR, ppr_iterations=10)
    b.add(np.arange(64), src)
    st = b.run()
    assert st.supersteps == 10
    for lane in (0, 1, 17, 63):
        ref = oracle.check_ppr_fp64(n, s, d, int(src[lane]), 10)
        assert rel_l1(b.result(lane), ref) <= 1e-5, lane
    # several lanes in one pass (lanes in one, two and four 64-byte sectors
    # of the rank rows) equal the single-lane extraction bit for bit
    for lanes in ([5], [0, 15], [3, 16, 40, 63], list(range(64))):
        many = b.results(lanes)
        for lane in lanes:
            assert np.array_equal(many[lane], b.result(lane)), lane


def test_ppr_alone_equals_mixed_bitwise(gpu):
    el = workloads.road(128, 128, seed=8)
    g = gas.Graph(el)
    src = _sources(np.random.default_rng(4), el, 20)
    b1 = gas.Batch(g, gas.BATCH_PPR, ppr_iterations=10)
    b1.add([3], [src[0]])
    b1.run()
    alone = b1.result(3)
    b = gas.Batch(g, gas.BATCH_PPR, ppr_iterations=10)
    b.add(np.arange(10), src[5:15])
    b.step(4)
    b.add([11], [src[0]])
    while True:
        b.step(1)
        if not (b.alive() >> 11) & 1:
            break
    assert np.array_equal(b.result(11), alone)


@pytest.mark.parametrize("delta", [1, 16, 255])
def test_batch_sssp_distance_bound(gpu, delta):
    # a lane relaxes only below (rounds since it began) * delta; lanes that
    # join a running batch start their own bound; distances stay exact
    for el in (workloads.road(128, 128, seed=5),
               workloads.rmat(12, symmetrize=True, weights=True)):
        s, d, w = el.arrays()
        n = el.n_vertices
        g = gas.Graph(el)
        src = _sources(np.random.default_rng(delta), el, 20)
        b = gas.Batch(g, gas.BATCH_SSSP)
        b.set_sssp_delta(delta)
        b.add(np.arange(16), src[:16])
        b.step(4)
        b.add([16, 17, 18, 19], src[16:20])
        b.run()
        res = b.results(range(20))
        for lane in range(20):
            want = oracle.check_dijkstra(n, s, d, w, int(src[lane]))
            assert np.array_equal(res[lane], want), lane


def test_batch_bfs_deep_levels(gpu):
    # a 2 x 700 grid: depths up to ~700, past what one byte per lane holds
    el = gas.Edges.grid(2, 700)
    s, d, _ = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    b = gas.Batch(g, gas.BATCH_BFS)
    src = np.array([0, 699, 700, 350], np.int32)
    b.add(np.arange(4), src)
    b.step(3)
    b.add([4], [1399])  # joins three rounds late
    b.run()
    res = b.results(range(5))
    for lane, v in enumerate([0, 699, 700, 350, 1399]):
        want = oracle.check_bfs(n, s, d, v)
        assert want.max() > 254 or v == 350
        assert np.array_equal(res[lane], want), lane


def test_batch_bfs_level_limit(gpu):
    # depths are 16-bit per lane: a lane still running past 65534 levels is
    # an error, a finished lane held for its result is not
    n = 70002
    s = np.concatenate([np.arange(69999), [70000]]).astype(np.int32)
    d = np.concatenate([np.arange(1, 70000), [70001]]).astype(np.int32)
    g = gas.Graph(gas.Edges.from_arrays(n, s, d))
    b = gas.Batch(g, gas.BATCH_BFS)
    b.add([0], [70000])          # done after two levels
    b.step(66000)                # held, not released, far past the limit
    assert b.alive() == 0
    r = b.results([0])[0]
    assert r[70000] == 0 and r[70001] == 1 and np.all(r[:70000] == -1)
    b.add([1], [0])              # a path of 69999 levels
    with pytest.raises(gas.GasError) as e:
        b.run()
    assert e.value.code == 5     # GAS_ERR_STATE


def test_batch_and_server_edge_cases(gpu):
    # sources without out-edges, unreachable vertices, a one-vertex graph
    # with no edges, and a server stream of one query per kind
    src = np.array([1, 2, 0, 3], np.int32)
    dst = np.array([2, 3, 4, 4], np.int32)
    w = np.array([7, 9, 2, 5], np.int32)
    el = gas.Edges.from_arrays(6, src, dst, w)
    g = gas.Graph(el)
    s, d, ww = el.arrays()
    n = 6
    for kind in (gas.BATCH_BFS, gas.BATCH_SSSP):
        b = gas.Batch(g, kind)
        sources = [4, 5, 1, 0]  # 4 and 5 have no out-edges
        b.add(np.arange(4), sources)
        b.run()
        res = b.results(range(4))
        for lane, v in enumerate(sources):
            want = (oracle.check_bfs(n, s, d, v) if kind == gas.BATCH_BFS
                    else oracle.check_dijkstra(n, s, d, ww, v))
            assert np.array_equal(res[lane], want), (kind, lane)
    b = gas.Batch(g, gas.BATCH_PPR)
    b.add([0, 1], [5, 1])
    b.run()
    for lane, v in ((0, 5), (1, 1)):
        ref = oracle.check_ppr_fp64(n, s, d, v, 10)
        got = b.result(lane).astype(np.float64)
        assert np.abs(got - ref).sum() <= 1e-5 * np.abs(ref).sum(), lane
    one = gas.Graph(gas.Edges.from_arrays(1, [], [], []))
    b = gas.Batch(one, gas.BATCH_BFS)
    b.add([7], [0])
    b.run()
    assert b.result(7).tolist() == [0]
    srv = gas.Server([g], memory_budget=1 << 30)
    lat, results, st = srv.replay([0, 1, 2], [0, 0, 0], [4, 1, 5],
                                  np.array([0.0, 0.0, 0.0]), keep=[0, 1, 2])
    assert np.all(lat > 0)
    assert np.array_equal(results[0], oracle.check_bfs(n, s, d, 4))
    assert np.array_equal(results[1], oracle.check_dijkstra(n, s, d, ww, 1))
    ref = oracle.check_ppr_fp64(n, s, d, 5, 10)
    assert np.abs(results[2] - ref).sum() <= 1e-5 * np.abs(ref).sum()


def test_batch_autotune(gpu):
    el = workloads.road(128, 128, seed=8)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    assert g.batch_param(gas.BATCH_SSSP) == 0
    delta = g.autotune_batch(gas.BATCH_SSSP)
    assert delta in (0, 16, 32, 64)
    assert g.batch_param(gas.BATCH_SSSP) == delta
    pull = g.autotune_batch(gas.BATCH_BFS)
    assert pull in (0, 2, 8, 32)
    for kind in (gas.BATCH_BFS, gas.BATCH_SSSP):
        assert g.batch_hub_items(kind) in (16, 32, 64, 128)
    assert g.batch_item_edges(gas.BATCH_BFS) in (8, 16)
    assert g.batch_item_edges(gas.BATCH_SSSP) in (32, 64)
    # batches created now start with the tuned settings; results stay exact
    src = _sources(np.random.default_rng(4), el, 10)
    for kind in (gas.BATCH_BFS, gas.BATCH_SSSP):
        b = gas.Batch(g, kind)
        b.add(np.arange(10), src)
        b.run()
        res = b.results(range(10))
        for lane in range(10):
            want = (oracle.check_bfs(n, s, d, int(src[lane])) if kind == gas.BATCH_BFS
                    else oracle.check_dijkstra(n, s, d, w, int(src[lane])))
            assert np.array_equal(res[lane], want), (kind, lane)
    nw = gas.Graph(gas.Edges.from_arrays(3, [0, 1], [1, 2]))
    with pytest.raises(gas.GasError):
        nw.autotune_batch(gas.BATCH_SSSP)  # no weights


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_batch_hub_threshold(gpu, kind):
    # every hub threshold gives the same, exact per-lane results (R-MAT-14:
    # list vertices of hundreds of push items)
    el = workloads.rmat(14, symmetrize=True, weights=True)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    assert g.batch_hub_items(kind) == 32
    src = _sources(np.random.default_rng(6), el, 16)
    ref = [oracle.check_bfs(n, s, d, int(x)) if kind == gas.BATCH_BFS
           else oracle.check_dijkstra(n, s, d, w, int(x)) for x in src]
    for items in (16, 32, 64, 128, 1 << 20):
        b = gas.Batch(g, kind)
        b.set_hub_items(items)
        b.add(np.arange(16), src)
        b.run()
        res = b.results(range(16))
        for lane in range(16):
            assert np.array_equal(res[lane], ref[lane]), (items, lane)
    b = gas.Batch(g, kind)
    with pytest.raises(gas.GasError):
        b.set_hub_items(15)


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP, gas.BATCH_PPR])
def test_batch_degree_order(gpu, kind):
    # on a skewed graph of 2**20+ edges the batch engines run on the
    # degree-ordered copy; sources and results stay in the caller's ids
    el = workloads.rmat(16, symmetrize=True, weights=True)
    s, d, w = el.arrays()
    n = el.n_vertices
    assert el.n_edges >= 1 << 20
    src = _sources(np.random.default_rng(4), el, 64)
    src[0] = 0
    out = []
    for mode in (1, 0):
        g = gas.Graph(el)
        g.set_vertex_order(mode)
        b = gas.Batch(g, kind, ppr_iterations=10)
        b.add(np.arange(64), src)
        b.run()
        out.append([b.result(lane) for lane in range(64)])
    for lane in range(64):
        if kind == gas.BATCH_PPR:
            assert rel_l1(out[0][lane], out[1][lane]) <= 1e-5, lane
        else:
            assert np.array_equal(out[0][lane], out[1][lane]), lane
    if kind == gas.BATCH_PPR:
        # the plain engine's PPR on the ordered copy, host and device results
        g = gas.Graph(el)
        ref = oracle.check_ppr_fp64(n, s, d, int(src[5]), 10)
        r, _ = g.ppr(int(src[5]), 10)
        assert rel_l1(r, ref) <= 1e-5
        buf = gas.DeviceBuffer(4 * n)
        g.ppr_device(int(src[5]), buf.ptr, 10)
        assert rel_l1(buf.to_numpy().view(np.float32), ref) <= 1e-5
    for lane in (0, 33):
        if kind == gas.BATCH_BFS:
            want = oracle.bfs(n, s, d, int(src[lane]))[0]
        elif kind == gas.BATCH_SSSP:
            want = oracle.sssp(n, s, d, w, int(src[lane]))[0]
        else:
            ref = oracle.check_ppr_fp64(n, s, d, int(src[lane]), 10)
            assert rel_l1(out[0][lane], ref) <= 1e-5
            continue
        assert np.array_equal(out[0][lane], want), lane


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
@pytest.mark.parametrize("graph", ["road", "rmat"])
def test_batch_round_graph_replay(gpu, kind, graph):
    # rounds issued as replays of a captured graph of 1, 2 or 4 rounds give
    # the results of direct launches, exact per lane: one capture is replayed
    # many times, lanes join and leave between replays (the lanes in use and
    # the round live in device memory), and pull / distance-bound settings
    # recapture it
    if graph == "road":
        el = workloads.road(96, 96, seed=31)
    else:
        el = workloads.rmat(13, symmetrize=True, weights=True)
    s, d, w = el.arrays()
    n = el.n_vertices
    g = gas.Graph(el)
    src = _sources(np.random.default_rng(12), el, 48)

    def want(x):
        return (oracle.check_bfs(n, s, d, int(x)) if kind == gas.BATCH_BFS
                else oracle.check_dijkstra(n, s, d, w, int(x)))

    ref = {i: want(x) for i, x in enumerate(src)}
    for k in (1, 2, 4):
        b = gas.Batch(g, kind)
        b.set_graph_rounds(k)
        if kind == gas.BATCH_BFS:
            b.set_pull_divisor(8 if k == 2 else 0)
        else:
            b.set_sssp_delta(16 if k == 2 else 0)
        # a whole batch through runToCompletion's replays
        b.add(np.arange(32), src[:32])
        b.run()
        res = b.results(range(32))
        for lane in range(32):
            assert np.array_equal(res[lane], ref[lane]), (k, lane)
        b.release((1 << 32) - 1)
        # continuous: lanes join between replays, a running lane is released
        b.add(np.arange(8), src[32:40])
        b.step(k)
        b.add([8, 9], src[40:42])
        b.release(1 << 3)  # released while running: its result is dropped
        b.step(2 * k)
        b.add([10], [src[42]])
        while b.alive():
            b.step(k)
        res = b.results([0, 1, 2, 4, 5, 6, 7, 8, 9, 10])
        for lane, qi in ((0, 32), (1, 33), (2, 34), (4, 36), (5, 37), (6, 38),
                         (7, 39), (8, 40), (9, 41), (10, 42)):
            assert np.array_equal(res[lane], ref[qi]), (k, lane)
    with pytest.raises(gas.GasError):
        gas.Batch(g, kind).set_graph_rounds(9)


@pytest.mark.parametrize("kind", [gas.BATCH_BFS, gas.BATCH_SSSP])
def test_batch_item_edges(gpu, kind):
    # push items cut shorter (BFS 8 instead of 16 out-edges, SSSP 32 instead
    # of 64) give the same exact per-lane results, with hubs written by
    # workgroups and by list threads, on R-MAT-14 and a road graph; sizes
    # outside the kind's range are refused
    small, large = (8, 16) if kind == gas.BATCH_BFS else (32, 64)
    for el in (workloads.rmat(14, symmetrize=True, weights=True),
               workloads.road(96, 96, seed=17)):
        s, d, w = el.arrays()
        n = el.n_vertices
        g = gas.Graph(el)
        assert g.batch_item_edges(kind) == large
        src = _sources(np.random.default_rng(21), el, 24)
        ref = [oracle.check_bfs(n, s, d, int(x)) if kind == gas.BATCH_BFS
               else oracle.check_dijkstra(n, s, d, w, int(x)) for x in src]
        for edges, hub in ((small, 16), (small, 1 << 20), (large, 16)):
            b = gas.Batch(g, kind)
            b.set_item_edges(edges)
            b.set_hub_items(hub)
            b.add(np.arange(24), src)
            b.run()
            res = b.results(range(24))
            for lane in range(24):
                assert np.array_equal(res[lane], ref[lane]), (edges, hub, lane)
    b = gas.Batch(g, kind)
    for bad in (small // 2, large * 2, small + 1):
        with pytest.raises(gas.GasError):
            b.set_item_edges(bad)
