"""Host-side pieces of the served measurement (no GPU): the Poisson query
stream, the FIFO model of the plain engine, and the latency summary."""
import numpy as np

import vertexapi2_amd as gas
from vertexapi2_amd import serve


class _StubGraph:
    """Just the CSR offsets make_stream reads."""

    def __init__(self, degrees):
        self._off = np.concatenate([[0], np.cumsum(degrees)]).astype(np.int32)

    def host_csr(self):
        return self._off, None, None


def test_fifo_latencies_closed_form():
    lat, end = serve.fifo_latencies(np.array([0.0, 1.0, 1.5, 10.0]),
                                    np.array([2.0, 1.0, 1.0, 0.5]))
    # finishes at 2, 3, 4 (queued) and 10.5 (idle server)
    assert np.allclose(lat, [2000.0, 2000.0, 2500.0, 500.0])
    assert end == 10.5


def test_stream_rate_mix_and_sources():
    g0 = _StubGraph([0, 3, 0, 1, 2])          # vertices 1, 3, 4 have out-edges
    g1 = _StubGraph([5] * 10)
    st = serve.make_stream(rate_qps=2000, n=20000, graphs=[g0, g1],
                           mix=(0.5, 0.3, 0.2), seed=3)
    gaps = np.diff(st["arrivals"])
    assert st["arrivals"][0] == 0.0 and np.all(gaps >= 0)
    assert abs(gaps.mean() * 2000 - 1.0) < 0.03       # Poisson rate
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.05  # exponential gaps
    frac = np.bincount(st["kinds"], minlength=3) / 20000
    assert np.allclose(frac, [0.5, 0.3, 0.2], atol=0.02)
    assert abs(np.mean(st["graph_ids"]) - 0.5) < 0.02
    s0 = st["sources"][st["graph_ids"] == 0]
    assert set(np.unique(s0)) == {1, 3, 4}          # out-degree >= 1 only
    # reproducible from the seed
    again = serve.make_stream(rate_qps=2000, n=20000, graphs=[g0, g1],
                              mix=(0.5, 0.3, 0.2), seed=3)
    assert all(np.array_equal(st[k], again[k])
               for k in ("kinds", "graph_ids", "sources", "arrivals"))


def test_summarize_percentiles():
    lat = np.arange(1, 101, dtype=np.float64)
    out = serve.summarize(lat, np.zeros(100), elapsed_s=2.0)
    assert out["queries"] == 100 and out["qps"] == 50.0
    assert out["p50_ms"] == 50.5 and out["p99_ms"] == 99.01
    assert serve.KIND_NAMES[gas.BATCH_PPR] == "ppr"
