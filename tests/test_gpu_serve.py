"""GPU tests of the continuous-batching server: every served result equals
the query run alone (bit-exact for BFS/SSSP; for PPR identical to a batch
lane alone, or to the plain engine when the idle server answered it alone)
and the CPU oracle; admission waits when memory is short and refuses queries that can
never fit."""
import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import serve, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gang", [(16, 1), (0, 0), (1, 0)])
def test_served_results_equal_alone(gpu, gang):
    # gang (16, 1): the default; (0, 0): always join a running batch;
    # (1, 0): any waiting query holds until the running batch drains
    road = gas.Graph(workloads.road(96, 96, seed=2))
    rmat = gas.Graph(workloads.rmat(12, symmetrize=True, weights=True))
    graphs = [road, rmat]
    edges = [workloads.road(96, 96, seed=2), workloads.rmat(12, symmetrize=True, weights=True)]
    stream = serve.make_stream(rate_qps=30000, n=150, graphs=graphs, seed=11)
    srv = gas.Server(graphs, memory_budget=1 << 34)
    srv.set_gang(*gang)
    keep = list(range(0, 150, 3))
    out, lat, results = serve.serve(srv, stream, keep=keep)
    assert out["queries"] == 150 and np.all(lat > 0)
    assert out["max_concurrent"] > 1  # queries overlapped
    for q in keep:
        k, gi, s = stream["kinds"][q], stream["graph_ids"][q], int(stream["sources"][q])
        src, dst, w = edges[gi].arrays()
        n = graphs[gi].n_vertices
        if k == gas.BATCH_BFS:
            assert np.array_equal(results[q], oracle.check_bfs(n, src, dst, s)), q
        elif k == gas.BATCH_SSSP:
            assert np.array_equal(results[q], oracle.check_dijkstra(n, src, dst, w, s)), q
        else:
            # a batch lane is bit-identical alone and mixed; a query the
            # idle server answered alone comes from the plain engine
            b = gas.Batch(graphs[gi], gas.BATCH_PPR)
            b.add([0], [s])
            b.run()
            plain, _ = graphs[gi].ppr(s, 10)
            assert (np.array_equal(results[q], b.result(0))
                    or np.array_equal(results[q], plain)), q
            ref = oracle.check_ppr_fp64(n, src, dst, s, 10)
            assert np.abs(results[q] - ref).sum() <= 1e-5 * np.abs(ref).sum()


def test_admission_by_memory(gpu):
    # the budget bounds the device memory the server holds: batch engine
    # state (charged when built), lane result buffers (charged when first
    # used) and the solo buffers; admission waits for room, evicts idle
    # engines, and refuses a query that can never fit
    el = workloads.road(64, 64, seed=3)
    road = gas.Graph(el)
    V = road.n_vertices
    s, d, w = el.arrays()
    srv = gas.Server([road], memory_budget=1 << 34)
    srv.set_solo(0, 0)  # every query through the batch engines
    serve.serve(srv, serve.make_stream(100000, 3, [road], mix=(0, 1, 0), seed=1))
    fp = [srv.engine_bytes(k, 0) for k in (gas.BATCH_BFS, gas.BATCH_SSSP, gas.BATCH_PPR)]
    assert min(fp) > 64 * V  # 64 lanes of state each
    solo = 2 * 4 * V         # the two solo threads' result buffers
    # room for the SSSP engine and two result buffers: two queries at once
    budget = solo + fp[gas.BATCH_SSSP] + 2 * 4 * V
    srv.set_budget(budget)
    stream = serve.make_stream(rate_qps=100000, n=12, graphs=[road], mix=(0, 1, 0), seed=4)
    out, lat, results = serve.serve(srv, stream, keep=list(range(12)))
    assert out["max_concurrent"] <= 2
    assert out["admissions_deferred"] > 0
    assert 0 < out["device_bytes"] <= budget and srv.budget()[1] <= budget
    for q in range(12):
        assert np.array_equal(results[q], oracle.check_dijkstra(V, s, d, w, int(stream["sources"][q])))
    # room for one batch engine at a time: BFS and SSSP queries take turns,
    # each engine evicted while idle to build the other
    budget = solo + max(fp[0], fp[1]) + 4 * 4 * V
    srv.set_budget(budget)
    mixed = serve.make_stream(rate_qps=100000, n=16, graphs=[road], mix=(0.5, 0.5, 0), seed=6)
    out, lat, results = serve.serve(srv, mixed, keep=list(range(16)))
    assert out["evictions"] > 0 and out["device_bytes"] <= budget
    for q in range(16):
        src = int(mixed["sources"][q])
        want = (oracle.check_bfs(V, s, d, src) if mixed["kinds"][q] == gas.BATCH_BFS
                else oracle.check_dijkstra(V, s, d, w, src))
        assert np.array_equal(results[q], want), q
    # an SSSP query needs its engine and one result buffer beside the solo
    # buffers: one byte less can never fit
    srv.set_budget(solo + fp[gas.BATCH_SSSP] + 4 * V - 1)
    with pytest.raises(gas.GasError) as ex:
        serve.serve(srv, stream)
    assert ex.value.code == 6  # GAS_ERR_FULL


def test_online_submit_poll_wait(gpu):
    # queries handed to a running server one at a time from two host
    # threads: every result equals the oracle's (PPR within 1e-5 of FP64),
    # each ticket reports done with a positive latency, and the server
    # refuses replays while online and submissions after the stop
    import threading
    import time
    els = [workloads.road(96, 96, seed=2), workloads.rmat(12, symmetrize=True, weights=True)]
    graphs = [gas.Graph(e) for e in els]
    srv = gas.Server(graphs, memory_budget=0)  # 80% of the free memory
    assert srv.budget()[0] > 1 << 30
    srv.start()
    rng = np.random.default_rng(12)
    qs = []
    for i in range(80):
        gi = int(rng.integers(0, 2))
        k = int(rng.choice([gas.BATCH_BFS, gas.BATCH_SSSP, gas.BATCH_PPR], p=[0.5, 0.3, 0.2]))
        src = int(rng.integers(0, graphs[gi].n_vertices))
        qs.append((k, gi, src))
    outs = [np.zeros(graphs[gi].n_vertices, np.float32 if k == gas.BATCH_PPR else np.int32)
            for k, gi, _ in qs]
    tickets = [None] * len(qs)

    def submitter(part):
        for i in part:
            k, gi, src = qs[i]
            tickets[i] = srv.submit(k, gi, src, outs[i])
            if i % 7 == 0:
                time.sleep(0.002)

    ths = [threading.Thread(target=submitter, args=(range(j, len(qs), 2),)) for j in (0, 1)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    with pytest.raises(gas.GasError):
        srv.replay([0], [0], [0], [0.0])
    lats = [srv.wait(t, 120.0) for t in tickets]
    assert all(x > 0 for x in lats)
    assert all(srv.poll(t)[0] == 1 for t in tickets)
    st = srv.stop()
    assert st.elapsed_s > 0 and 0 <= st.solo <= len(qs)
    with pytest.raises(gas.GasError):
        srv.submit(0, 0, 0)
    for i, (k, gi, src) in enumerate(qs):
        s, d, w = els[gi].arrays()
        n = graphs[gi].n_vertices
        if k == gas.BATCH_BFS:
            assert np.array_equal(outs[i], oracle.check_bfs(n, s, d, src)), i
        elif k == gas.BATCH_SSSP:
            assert np.array_equal(outs[i], oracle.check_dijkstra(n, s, d, w, src)), i
        else:
            ref = oracle.check_ppr_fp64(n, s, d, src, 10)
            assert np.abs(outs[i] - ref).sum() <= 1e-5 * np.abs(ref).sum(), i
    # a second session on the same server, every query through the batch
    # engines: a burst shares batches, and each result equals the first
    # session's for the same query
    srv.set_solo(0, 0)
    srv.start()
    again = [np.zeros_like(o) for o in outs[:40]]
    tk = [srv.submit(k, gi, src, again[i]) for i, (k, gi, src) in enumerate(qs[:40])]
    for t in tk:
        srv.wait(t, 120.0)
    st = srv.stop()
    assert st.max_concurrent > 1 and st.solo == 0
    for i, (k, gi, src) in enumerate(qs[:40]):
        if k == gas.BATCH_PPR:
            assert np.abs(again[i] - outs[i]).sum() <= 1e-5 * np.abs(outs[i]).sum(), i
        else:
            assert np.array_equal(again[i], outs[i]), i


def test_idle_server_answers_alone(gpu):
    # at a low rate queries find the server idle and are answered by the
    # plain engine at once (a batch of one), PPR included
    el = workloads.road(64, 64, seed=8)
    road = gas.Graph(el)
    srv = gas.Server([road], memory_budget=1 << 30)
    stream = serve.make_stream(rate_qps=200, n=24, graphs=[road], seed=13)
    out, lat, results = serve.serve(srv, stream, keep=list(range(24)))
    assert 0 < out["solo"] <= 24
    s, d, w = el.arrays()
    n = road.n_vertices
    for q in range(24):
        k, src = stream["kinds"][q], int(stream["sources"][q])
        if k == gas.BATCH_BFS:
            assert np.array_equal(results[q], oracle.check_bfs(n, s, d, src)), q
        elif k == gas.BATCH_SSSP:
            assert np.array_equal(results[q], oracle.check_dijkstra(n, s, d, w, src)), q
        else:
            ref = oracle.check_ppr_fp64(n, s, d, src, 10)
            assert np.abs(results[q] - ref).sum() <= 1e-5 * np.abs(ref).sum(), q
