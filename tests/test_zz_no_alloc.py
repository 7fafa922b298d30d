"""The allocation-free PageRank entry point (gas_pagerank_device_noalloc)
against the existing one (gas_pagerank_device) through the C ABI: two
calls in a row, the second on a smaller graph, each identical to the
allocating run (same kernels, same tiers) and within 1e-5 relative L1 of
the FP64 power iteration; a graph beyond the scratch is refused before any
work."""
import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import workloads

pytestmark = pytest.mark.gpu


def rel_l1(x, ref):
    return float(np.abs(x.astype(np.float64) - ref).sum() / np.abs(ref).sum())


def _both(el, iterations):
    g = gas.Graph(el)
    n = el.n_vertices
    a, b = gas.DeviceBuffer(4 * n), gas.DeviceBuffer(4 * n)
    ra, rb = gas.DeviceBuffer(8 * iterations), gas.DeviceBuffer(8 * iterations)
    g.pagerank_device(a.ptr, iterations, ra.ptr)
    g.pagerank_device_noalloc(b.ptr, iterations, rb.ptr)
    return (a.to_numpy(np.float32), b.to_numpy(np.float32),
            ra.to_numpy(np.float64), rb.to_numpy(np.float64))


def test_noalloc_pagerank_matches_allocating_run(gpu):
    # R-MAT-18 (hub tiles, degree-ordered copy), then a smaller road graph
    for el, iters in ((workloads.rmat(18), 20), (workloads.road(200, 200, seed=4), 7)):
        s, d, _ = el.arrays()
        n = el.n_vertices
        want, got, rwant, rgot = _both(el, iters)
        assert np.array_equal(got, want)
        assert np.array_equal(rgot, rwant)
        assert rel_l1(got, oracle.check_pagerank_fp64(n, s, d, iters)) <= 1e-5


def test_noalloc_pagerank_refuses_large_graphs(gpu):
    g = gas.Graph(workloads.road(1400, 1400, seed=2))  # 1.96M vertices
    buf = gas.DeviceBuffer(4 * g.n_vertices)
    with pytest.raises(gas.GasError) as e:
        g.pagerank_device_noalloc(buf.ptr, 3)
    assert e.value.code == 1 and "scratch" in str(e.value)
