"""A user program with an edge-writing scatter, compiled against the
engine's template header (tests/template/edge_trace.hip), on the GPU:
distances, per-edge scatter counts and recorded differences equal the
reference engine's run() (refgas.h:170-222; oracle/_ref) and the Python
restatement of it, through run() and the step-by-step API; with
scatterActivate(false) edges stay untouched (the reference GPU engine's
behaviour for every program, gpugas.h:666-668)."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import vertexapi2_amd as gas
from vertexapi2_amd import workloads

from conftest import ROOT
from test_template import edge_trace_py, small_graph

LIB = os.path.join(ROOT, "tests", "template", "_lib", "libedge_trace.so")
_lib = None


def et_lib():
    global _lib
    if _lib is None:
        assert os.path.exists(LIB), f"{LIB} missing: run tests/template/build.py"
        gas.lib()
        _lib = C.CDLL(LIB)
    return _lib


def et_run(n, s, d, w, source, mode):
    s, d, w = (np.ascontiguousarray(a, np.int32) for a in (s, d, w))
    dist = np.zeros(n, np.int32)
    hits = np.zeros(len(s), np.int32)
    last = np.zeros(len(s), np.int32)
    steps = C.c_int()
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
    rc = et_lib().et_run(C.c_int(n), C.c_int(len(s)), p(s), p(d), p(w), C.c_int(source),
                         C.c_int(mode), p(dist), p(hits), p(last), C.byref(steps))
    assert rc == 0, gas.lib().gas_last_error().decode()
    return dist, hits, last, steps.value


def hub_graph():
    # vertex 5: 300 out-edges (wave-cooperative scatter) and 2500 in-edges
    # (hub gather); a chain through every vertex for long frontiers
    n = 3000
    pairs = {(5, i) for i in range(100, 400)} | {(i, 5) for i in range(400, 2900)}
    pairs |= {(i, i + 1) for i in range(n - 1)}
    rng = np.random.default_rng(9)
    while len(pairs) < 9000:
        u, v = (int(x) for x in rng.integers(0, n, 2))
        if u != v:
            pairs.add((u, v))
    order = rng.permutation(len(pairs))
    arr = np.array(sorted(pairs), np.int32)[order]
    return n, arr[:, 0].copy(), arr[:, 1].copy(), rng.integers(1, 30, len(arr)).astype(np.int32)


def cases():
    yield "random80", small_graph(4, n=80, m=500), 3
    el = workloads.road(24, 24, seed=5)
    s, d, w = el.arrays()
    yield "road24", (el.n_vertices, s, d, w), 0
    yield "hub3000", hub_graph(), 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["random80", "road24", "hub3000"])
def test_edge_scatter_matches_reference(gpu, name):
    (n, s, d, w), src = next((g, src) for nm, g, src in cases() if nm == name)
    ref = edge_trace_py(n, s, d, w, src)
    if oracle.ref_available():
        rr = oracle.ref_edge_trace(n, s, d, w, src)
        for a, b in zip(ref[:3], rr[:3]):
            assert a.tolist() == b.tolist()
    assert ref[0].tolist() == oracle.check_dijkstra(n, s, d, w, src).tolist()
    for mode in (0, 1):
        dist, hits, last, steps = et_run(n, s, d, w, src, mode)
        assert np.array_equal(dist, ref[0]), (name, mode)
        assert np.array_equal(hits, ref[1]), (name, mode)
        assert np.array_equal(last, ref[2]), (name, mode)
        if mode == 1:
            assert steps == ref[3]
    dist, hits, last, _ = et_run(n, s, d, w, src, 2)
    assert np.array_equal(dist, ref[0])
    assert not hits.any() and not last.any()
