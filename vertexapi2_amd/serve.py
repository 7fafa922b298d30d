"""Served query streams: Poisson arrivals, the continuous-batching server,
and the plain engine answering the same stream one query at a time.

    stream = make_stream(rate_qps=2000, n=1000, graphs=[road, rmat], seed=5)
    served = serve(server, stream)            # continuous batching
    lat, end = plain_replay(graphs, stream)    # the plain engine, one at a time
"""
from __future__ import annotations

import time

import numpy as np

from . import BATCH_BFS, BATCH_PPR, BATCH_SSSP

KIND_NAMES = {BATCH_BFS: "bfs", BATCH_SSSP: "sssp", BATCH_PPR: "ppr"}


def make_stream(rate_qps: float, n: int, graphs, mix=(0.5, 0.3, 0.2), seed: int = 1,
                source_degree_min: int = 1):
    """n queries with exponential inter-arrival times (Poisson process of
    `rate_qps`), kinds drawn from `mix` (BFS, SSSP, PPR), a graph drawn
    uniformly, and a source drawn uniformly among vertices with out-degree
    >= source_degree_min."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_qps, n))
    arrivals -= arrivals[0]
    kinds = rng.choice([BATCH_BFS, BATCH_SSSP, BATCH_PPR], size=n, p=list(mix)).astype(np.int32)
    gids = rng.integers(0, len(graphs), n).astype(np.int32)
    cands = []
    for g in graphs:
        off = g.host_csr()[0]
        cands.append(np.nonzero(np.diff(off) >= source_degree_min)[0].astype(np.int32))
    sources = np.array([cands[gi][rng.integers(0, len(cands[gi]))] for gi in gids],
                       dtype=np.int32)
    return {"kinds": kinds, "graph_ids": gids, "sources": sources, "arrivals": arrivals,
            "rate_qps": rate_qps}


def summarize(latency_ms, arrivals, elapsed_s=None):
    lat = np.asarray(latency_ms)
    n = len(lat)
    if elapsed_s is None:
        elapsed_s = float(np.max(arrivals + lat / 1e3))
    return {"queries": n, "qps": round(n / elapsed_s, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "elapsed_s": round(elapsed_s, 4)}


def serve(server, stream, keep=None):
    lat, results, st = server.replay(stream["kinds"], stream["graph_ids"], stream["sources"],
                                     stream["arrivals"], keep=keep)
    out = summarize(lat, stream["arrivals"], st.elapsed_s)
    out.update(st.as_dict())
    return out, lat, results


def plain_service_times(graphs, stream, ppr_iterations: int = 10):
    """Seconds each query takes on the plain engine answering one query at a
    time with its result left in HBM (gas_bfs_device / gas_sssp_device /
    gas_ppr_device: device-side initial state, run, device copy of the
    result), measured back to back on the host clock."""
    from . import DeviceBuffer
    bufs = [DeviceBuffer(4 * g.n_vertices) for g in graphs]
    svc = np.zeros(len(stream["kinds"]))
    for i, (k, gi, s) in enumerate(zip(stream["kinds"], stream["graph_ids"],
                                       stream["sources"])):
        g, buf = graphs[gi], bufs[gi]
        t = time.perf_counter()
        if k == BATCH_BFS:
            g.bfs_device(int(s), buf.ptr)
        elif k == BATCH_SSSP:
            g.sssp_device(int(s), buf.ptr)
        else:
            g.ppr_device(int(s), buf.ptr, ppr_iterations)
        svc[i] = time.perf_counter() - t
    return svc


def plain_replay(graphs, stream, ppr_iterations: int = 10):
    """The plain engine serving the stream itself: queries answered one at a
    time in arrival order, each started no earlier than its arrival on the
    host clock (gas_bfs_device / gas_sssp_device / gas_ppr_device, result
    left in HBM).  Returns (latency_ms per query, seconds from the first
    arrival to the last completion)."""
    from . import DeviceBuffer
    bufs = [DeviceBuffer(4 * g.n_vertices) for g in graphs]
    arrivals = np.asarray(stream["arrivals"], dtype=np.float64)
    n = len(arrivals)
    finish = np.zeros(n)
    t0 = time.perf_counter()
    for i, (k, gi, s) in enumerate(zip(stream["kinds"], stream["graph_ids"],
                                       stream["sources"])):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        g, buf = graphs[gi], bufs[gi]
        if k == BATCH_BFS:
            g.bfs_device(int(s), buf.ptr)
        elif k == BATCH_SSSP:
            g.sssp_device(int(s), buf.ptr)
        else:
            g.ppr_device(int(s), buf.ptr, ppr_iterations)
        finish[i] = time.perf_counter() - t0
    return (finish - arrivals) * 1e3, float(finish[-1]) if n else 0.0


def fifo_latencies(arrivals, service_s):
    """Latencies of a single FIFO server with the given service times."""
    finish = np.zeros(len(arrivals))
    t = 0.0
    for i, (a, s) in enumerate(zip(arrivals, service_s)):
        t = max(t, a) + s
        finish[i] = t
    return (finish - arrivals) * 1e3, float(finish[-1]) if len(finish) else 0.0
