"""Benchmark of the GAS hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--extra/--no-extra]

Workload (the headline configuration): BFS and SSSP from vertex 0 on the
1024x1024 road-like graph (3% links dropped, 1% long links, weights
U[1,64]).  One step = one BFS run plus one SSSP run of the plain engine to
completion on the graph already resident in HBM, from frontier
initialisation to results back on the host (the reference's "Took" window,
sssp.cu:77-81, bfs.cu:100-111).

value = traversed edges per second in GTEPS: (E_r(BFS) + E_r(SSSP)) / time,
with E_r the summed out-degree of the vertices each run reaches.

roofline: the dominant kernel of the step (by device time) with its
algorithmic bytes per launch over its average launch duration, both measured
here with HIP events around every launch of that kernel.
cpu_baseline: the reference's own CPU engine (refgas.h, built from its
sources into oracle/_ref) on the same workload, one host core.

Extra per-app lines (grid BFS, R-MAT-21 PageRank and CC) are reported under
"apps" unless --no-extra.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import vertexapi2_amd as gas  # noqa: E402
from vertexapi2_amd import workloads  # noqa: E402

METRIC = "GTEPS + % HBM roofline per app; batched-served QPS & p99 vs plain 1-source"
HBM_PEAK_GBS = 8000.0


def reached_edges(depth_or_dist, csr_off, unreached):
    deg = np.diff(csr_off).astype(np.int64)
    return int(deg[depth_or_dist != unreached].sum())


def gather_bytes(st) -> float:
    """Algorithmic bytes of the gather kernel for a run (int32 programs):
    per gathered edge 4 (source id) + 4 (edge weight) + 4 (source state);
    per active vertex 8 (two offsets) + 4 (result)."""
    return 12.0 * st.gathered_edges + 12.0 * st.active_total


def scatter_bytes(st) -> float:
    """Per scattered edge 4 (destination id) + 4 (frontier word update);
    per active vertex 8 (offsets) + 4 (predicated degree)."""
    return 8.0 * st.scattered_edges + 12.0 * st.active_total


def run_plain(g, csr_off, steps, warmup):
    for _ in range(warmup):
        g.bfs(0)
        g.sssp(0)
    times, sts = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        depth, sb = g.bfs(0)
        dist, ss = g.sssp(0)
        times.append(time.perf_counter() - ts)
        sts.append((sb, ss))
    total = time.perf_counter() - t0
    er = reached_edges(depth, csr_off, -1) + reached_edges(dist, csr_off, gas.SSSP_UNREACHED)
    return total, times, er, sts[-1], depth, dist


def kernel_roofline(g):
    """Per-launch event timing of the gather and scatter kernels over one
    step; returns the dominant one's figures."""
    gas.set_kernel_timing(True)
    try:
        _, sb = g.bfs(0)
        _, ss = g.sssp(0)
    finally:
        gas.set_kernel_timing(False)
    cands = []
    g_launch = ss.gather_launches
    if g_launch:
        cands.append(("k_gather<SSSP>", ss.gather_ms, g_launch, gather_bytes(ss)))
    s_launch = sb.scatter_launches + ss.scatter_launches
    if s_launch:
        cands.append(("k_scatter<BFS,SSSP>", sb.scatter_ms + ss.scatter_ms, s_launch,
                      scatter_bytes(sb) + scatter_bytes(ss)))
    name, ms, n, byts = max(cands, key=lambda c: c[1])
    avg_s = ms / 1e3 / n
    achieved = byts / n / avg_s / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None, "avg_launch_us": round(avg_s * 1e6, 3), "launches": n,
            "algorithmic_bytes_per_launch": round(byts / n, 1)}


def cpu_baseline(el, budget_s=12.0):
    import oracle
    if not oracle.ref_available():
        return None
    s, d, w = el.arrays()
    n = el.n_vertices
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, s + 1, 1)
    off = np.cumsum(off)
    reps, us_total, er_total = 0, 0, 0
    t0 = time.perf_counter()
    while True:
        depth, us1 = oracle.ref_bfs(n, s, d, 0)
        dist, us2 = oracle.ref_sssp(n, s, d, w, 0)
        us_total += us1 + us2
        er_total += reached_edges(depth, off, -1) + reached_edges(dist, off, gas.SSSP_UNREACHED)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 20:
            break
    gteps = er_total / (us_total * 1e-6) / 1e9
    return {"value": round(gteps, 5), "unit": "GTEPS", "cores": 1, "kind": "reference",
            "sample": f"{reps} x (BFS + SSSP from vertex 0) on the full 1024x1024 road graph "
                      f"with refgas.h (GASEngineRef) single-threaded; {us_total / reps / 1e3:.1f} "
                      f"ms per step in the reference's Took window"}


def app_lines(steps):
    out = {}
    # grid BFS
    el = workloads.grid64()
    g = gas.Graph(el)
    off = g.host_csr()[0]
    g.bfs(0)
    ts = []
    for _ in range(max(3, steps)):
        t = time.perf_counter()
        depth, st = g.bfs(0)
        ts.append(time.perf_counter() - t)
    er = reached_edges(depth, off, -1)
    t = float(np.median(ts))
    out["grid64_bfs"] = {"gteps": round(er / t / 1e9, 4), "ms": round(t * 1e3, 3),
                         "device_ms": round(st.ms, 3), "supersteps": st.supersteps,
                         "roofline_frac": round((20 * 4096 + 8 * er) / t / 1e12 / 8, 6)}
    del g, el
    # R-MAT-21 PageRank, 20 sweeps
    el = workloads.rmat(21)
    g = gas.Graph(el)
    V, E = g.n_vertices, g.n_edges
    g.pagerank(20)
    ts = []
    for _ in range(max(3, steps)):
        t = time.perf_counter()
        r, res, st = g.pagerank(20)
        ts.append(time.perf_counter() - t)
    t = float(np.median(ts))
    byts = (8.0 * E + 16.0 * V) * 20
    out["rmat21_pagerank20"] = {"gteps": round(20 * E / t / 1e9, 4), "ms": round(t * 1e3, 3),
                                "device_ms": round(st.ms, 3), "V": V, "E": E,
                                "roofline_frac": round(byts / t / 1e9 / HBM_PEAK_GBS, 5),
                                "device_roofline_frac": round(byts / (st.ms * 1e-3) / 1e9 /
                                                              HBM_PEAK_GBS, 5)}
    del g, el
    el = workloads.rmat(21, symmetrize=True)
    g = gas.Graph(el)
    g.cc()
    ts = []
    for _ in range(max(3, steps)):
        t = time.perf_counter()
        lab, st = g.cc()
        ts.append(time.perf_counter() - t)
    t = float(np.median(ts))
    out["rmat21s_cc"] = {"gteps": round(st.gathered_edges / t / 1e9, 4),
                         "ms": round(t * 1e3, 3), "device_ms": round(st.ms, 3),
                         "supersteps": st.supersteps, "E": g.n_edges}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--extra", dest="extra", action="store_true", default=True)
    ap.add_argument("--no-extra", dest="extra", action="store_false")
    ap.add_argument("--no-cpu", dest="cpu", action="store_false", default=True)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the engine on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if gas.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    el = workloads.road()
    g = gas.Graph(el)
    csr_off = g.host_csr()[0]
    total, times, er, (sb, ss), depth, dist = run_plain(g, csr_off, args.steps, args.warmup)
    ms_per_step = total / args.steps * 1e3
    gteps = er * args.steps / total / 1e9
    roof = kernel_roofline(g)
    line = {
        "metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator, graph resident in HBM)",
        "config": {"workload": "road1024: BFS + SSSP from vertex 0, 1024x1024 road-like grid, "
                               f"V={g.n_vertices}, E={g.n_edges}, weights U[1,64]",
                   "edges_traversed_per_step": er},
        "step_detail": {"bfs_device_ms": round(sb.ms, 4), "sssp_device_ms": round(ss.ms, 4),
                        "bfs_supersteps": sb.supersteps, "sssp_supersteps": ss.supersteps,
                        "median_step_ms": round(float(np.median(times)) * 1e3, 4)},
        "roofline": roof,
    }
    del g
    if args.cpu:
        cb = cpu_baseline(el)
        line["cpu_baseline"] = cb
    if args.extra:
        line["apps"] = app_lines(min(args.steps, 5))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
