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
    road = gas.Graph(workloads.road(64, 64, seed=3))
    el = workloads.road(64, 64, seed=3)
    V = road.n_vertices
    budget = 2 * gas.query_bytes(gas.BATCH_SSSP, V)  # two SSSP queries at once
    srv = gas.Server([road], memory_budget=budget)
    srv.set_solo(0, 0)  # every query through the batch engine
    stream = serve.make_stream(rate_qps=100000, n=12, graphs=[road], mix=(0, 1, 0), seed=4)
    out, lat, results = serve.serve(srv, stream, keep=list(range(12)))
    assert out["max_concurrent"] <= 2
    assert out["admissions_deferred"] > 0
    s, d, w = el.arrays()
    for q in range(12):
        assert np.array_equal(results[q], oracle.check_dijkstra(V, s, d, w, int(stream["sources"][q])))
    tiny = gas.Server([road], memory_budget=gas.query_bytes(gas.BATCH_SSSP, V) - 1)
    with pytest.raises(gas.GasError) as ex:
        serve.serve(tiny, stream)
    assert ex.value.code == 6  # GAS_ERR_FULL


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
