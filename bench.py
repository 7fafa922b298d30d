"""Benchmark of the GAS hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extra] [--no-cpu]

Headline workload: BFS and SSSP from vertex 0 on the 1024x1024 road-like
graph (3% of links dropped, 1% long links, int32 weights U[1,64]).  One step
= one BFS run plus one SSSP run of the plain engine to completion on the
graph already resident in HBM: initial vertex state written on the device,
supersteps until the frontier is empty, result copied into a device buffer
(gas_bfs_device / gas_sssp_device; the reference's "Took" window is the run
alone, sssp.cu:77-81, bfs.cu:100-111).  Degree tiers are autotuned once per
program before the warmup.

value = traversed edges per second in GTEPS: (E_r(BFS) + E_r(SSSP)) / time,
with E_r the summed out-degree of the vertices the run reaches, and time the
host wall clock around the K timed steps (each step ends synchronised).

roofline: after the timed steps, the same K steps run again with every
engine kernel launch bracketed by HIP events on the engine's stream (the
event packets roughly double a road step, so that pass is not the value);
the dominant kernel (largest summed time) is reported with its algorithmic
bytes per launch (model below) over its average launch duration.  traffic: HBM bytes per launch of that kernel
from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept under profiles/
(pmc_<workload>.json), or null when none matches.
cpu_baseline: the reference's own CPU engine (refgas.h built from its
sources into oracle/_ref) on the same workload, one host core.

Further lines under "apps" (unless --no-extra): grid BFS (64x64), R-MAT-21
PageRank (20 sweeps) and CC (symmetrised), each with its dominant-kernel
roofline; "batched": 64-source BFS and SSSP batches on symmetrised weighted
R-MAT-22 against the plain engine answering the same sources one at a time;
"served": Poisson query streams (50/30/20 BFS/SSSP/PPR over road1024 and
R-MAT-22) replayed through the continuous-batching server, with QPS, p50
and p99 against a FIFO of the plain engine's measured service times.
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
from vertexapi2_amd import serve, workloads  # noqa: E402

METRIC = "GTEPS + % HBM roofline per app; batched-served QPS & p99 vs plain 1-source"
HBM_PEAK_GBS = 8000.0

# Bytes of one vertex's state, one edge's value and one gather result.
VD = {"bfs": 4, "sssp": 4, "cc": 4, "pagerank": 8, "ppr": 12}
ED = {"sssp": 4}
GR = 4
# Bytes a gather reads per in-edge for the source: the state, or for the
# source-only programs (PageRank, PPR) the per-vertex contribution
# rank / outdeg (4 bytes, refreshed by apply).
SRC = {"bfs": 4, "sssp": 4, "cc": 4, "pagerank": 4, "ppr": 4}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def reached_edges(values, csr_off, unreached):
    deg = np.diff(csr_off).astype(np.int64)
    return int(deg[values != unreached].sum())


# ---------------------------------------------------------------------------
# Algorithmic bytes per kernel (each datum counted once, lower bounds).
#   k_gather_vp   per gathered vertex: 8 (CSC offset pair) + 4 (result);
#                 per in-edge: 4 (source id) + edge value + source read (SRC)
#   k_apply       per active vertex: 4 (slot id) + 4 (gather result, when
#                 there is a gather) + state; per scattered out-edge: 4
#                 (destination id) + 1 (activation byte)
#   k_hgather     per hub in-edge: 4 + edge value + source state; per hub:
#                 8 + 4
#   k_compact1    V (activation bytes); k_compact2: V + 4 per listed id
#   k_fused       per active vertex: 8 (offsets) + state read + state write;
#                 per in-edge as k_gather_vp; per scattered edge 5
#   k_small       (runs that never left the one-workgroup mode) per active
#                 vertex: 4 (list id) + 8 (CSR offsets) + state read + write,
#                 with a gather + 8 (CSC offsets); per in-edge as
#                 k_gather_vp; per scattered edge 5
def kernel_models(app, st, n_vertices, have_gather, hub_split=None):
    """{kernel: total algorithmic bytes over the run}.  hub_split =
    (hub_vertices, hub_edges) per sweep for all-active runs, splitting the
    gather between k_gather_vp and k_hgather; without it every gathered
    edge is booked to k_gather_vp (graphs whose in-degrees stay under the
    hub tier)."""
    vd, ed, sv = VD[app], ED.get(app, 0), SRC[app]
    A, Eg, Es = st.active_total, st.gathered_edges, st.scattered_edges
    out = {}
    ks = st.kernels()
    if set(ks) <= {"k_small"}:
        out["k_small"] = (A * (12 + 2 * vd + (8 if have_gather else 0)) +
                          Eg * (4 + ed + sv) + Es * 5)
        return out
    if have_gather and "k_fused" in ks and "k_gather_vp" not in ks:
        # every superstep ran as one launch: gather + apply + scatter
        out["k_fused"] = A * (8 + 2 * vd) + Eg * (4 + ed + sv) + Es * 5
        return out
    if have_gather:
        if hub_split is not None:
            sweeps = st.supersteps
            hv, he = hub_split
            out["k_gather_vp"] = ((A - sweeps * hv) * (8 + GR) +
                                  (Eg - sweeps * he) * (4 + ed + sv))
            out["k_hgather"] = sweeps * (hv * (8 + GR) + he * (4 + ed + sv))
        else:
            out["k_gather_vp"] = A * (8 + GR) + Eg * (4 + ed + sv)
    out["k_apply"] = A * (4 + (GR if have_gather else 0) + vd) + Es * 5
    launches = st.kernels().get("k_compact1", (0, 0))[1]
    out["k_compact1"] = launches * n_vertices
    out["k_compact2"] = launches * n_vertices + 4 * A
    return out


def roofline_table(stats_list, models_list):
    """Merges per-run kernel times and byte models; returns (dominant entry,
    table)."""
    tot = {}
    for st, models in zip(stats_list, models_list):
        for name, (ms, n) in st.kernels().items():
            t = tot.setdefault(name, [0.0, 0, 0.0, True])
            t[0] += ms
            t[1] += n
            if name in models:
                t[2] += models[name]
            else:
                t[3] = False
    table = {}
    for name, (ms, n, byts, known) in tot.items():
        avg_s = ms / 1e3 / n
        e = {"launches": n, "avg_launch_us": round(avg_s * 1e6, 3),
             "total_ms": round(ms, 4)}
        if known and byts > 0:
            gbs = byts / n / avg_s / 1e9
            e.update({"algorithmic_bytes_per_launch": round(byts / n, 1),
                      "achieved_gbs": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 5)})
        table[name] = e
    dom = max(table, key=lambda k: table[k]["total_ms"])
    return dom, table


def roofline_entry(dom, table, workload):
    e = table[dom]
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"pmc_{workload}.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            rec = json.load(f)
        k = rec.get("kernels", {}).get(dom)
        if k is not None:
            traffic = k.get("hbm_bytes_per_launch")
    achieved = e.get("achieved_gbs")
    return {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": e.get("frac"), "traffic": traffic,
            "avg_launch_us": e["avg_launch_us"], "launches": e["launches"],
            "algorithmic_bytes_per_launch": e.get("algorithmic_bytes_per_launch")}


# ---------------------------------------------------------------------------
def headline(g, steps, warmup):
    tiers = {"bfs": g.autotune(gas.PROGRAM_BFS), "sssp": g.autotune(gas.PROGRAM_SSSP)}
    V = g.n_vertices
    dbuf, sbuf = gas.DeviceBuffer(4 * V), gas.DeviceBuffer(4 * V)
    for _ in range(warmup):
        g.bfs_device(0, dbuf.ptr)
        g.sssp_device(0, sbuf.ptr)
    # timed region
    t0 = time.perf_counter()
    for _ in range(steps):
        sb0 = g.bfs_device(0, dbuf.ptr)
        ss0 = g.sssp_device(0, sbuf.ptr)
    total = time.perf_counter() - t0
    # the same steps again with every engine launch bracketed by HIP events
    # (the event packets slow the step down, so this pass is not the value)
    gas.set_kernel_timing(True)
    sts = []
    try:
        t1 = time.perf_counter()
        for _ in range(steps):
            sb = g.bfs_device(0, dbuf.ptr)
            ss = g.sssp_device(0, sbuf.ptr)
            sts.append((sb, ss))
        total_timed = time.perf_counter() - t1
    finally:
        gas.set_kernel_timing(False)
    depth, dist = dbuf.to_numpy(), sbuf.to_numpy()
    csr_off = g.host_csr()[0]
    er = (reached_edges(depth, csr_off, -1) +
          reached_edges(dist, csr_off, gas.SSSP_UNREACHED))
    stats, models = [], []
    for sb, ss in sts:
        stats += [sb, ss]
        models += [kernel_models("bfs", sb, V, False), kernel_models("sssp", ss, V, True)]
    dom, table = roofline_table(stats, models)
    sb, ss = sb0, ss0
    return {"total_s": total, "total_timed_s": total_timed, "er": er, "tiers": tiers,
            "dom": dom, "table": table, "bfs": sb, "sssp": ss}


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


def timed_runs(fn, reps):
    ts, last = [], None
    for _ in range(reps):
        t = time.perf_counter()
        last = fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), last


def app_lines(steps):
    out = {}
    reps = max(3, steps)
    # grid BFS, 64x64
    g = gas.Graph(workloads.grid64())
    V = g.n_vertices
    buf = gas.DeviceBuffer(4 * V)
    g.autotune(gas.PROGRAM_BFS)
    g.bfs_device(0, buf.ptr)
    t, _ = timed_runs(lambda: g.bfs_device(0, buf.ptr), reps)
    gas.set_kernel_timing(True)
    st = g.bfs_device(0, buf.ptr)
    gas.set_kernel_timing(False)
    er = reached_edges(buf.to_numpy(), g.host_csr()[0], -1)
    dom, table = roofline_table([st], [kernel_models("bfs", st, V, False)])
    out["grid64_bfs"] = {"gteps": round(er / t / 1e9, 5), "ms": round(t * 1e3, 4),
                         "device_ms": round(st.ms, 4), "supersteps": st.supersteps,
                         "roofline": roofline_entry(dom, table, "grid64_bfs")}
    del g, buf
    log("grid64 done")
    # R-MAT-21 PageRank, 20 sweeps
    g = gas.Graph(workloads.rmat(21))
    V, E = g.n_vertices, g.n_edges
    tiers = g.autotune(gas.PROGRAM_PAGERANK)
    _, res, _ = g.pagerank(20)  # host-result run: the per-sweep residuals
    rbuf = gas.DeviceBuffer(4 * V)
    g.pagerank_device(rbuf.ptr, 20)
    t, st0 = timed_runs(lambda: g.pagerank_device(rbuf.ptr, 20), reps)
    gas.set_kernel_timing(True)
    st = g.pagerank_device(rbuf.ptr, 20)
    gas.set_kernel_timing(False)
    coff = g.host_csc()[0]
    indeg = np.diff(coff)
    hub = indeg > tiers[1]
    split = (int(hub.sum()), int(indeg[hub].sum()))
    dom, table = roofline_table([st], [kernel_models("pagerank", st, V, True, split)])
    out["rmat21_pagerank20"] = {
        "gteps": round(20 * E / t / 1e9, 4), "ms": round(t * 1e3, 3),
        "device_ms": round(st0.ms, 3), "V": V, "E": E, "tiers": tiers,
        "final_residual_l1": float(res[-1]),
        "roofline": roofline_entry(dom, table, "rmat21_pagerank20"), "kernels": table}
    del g
    log("pagerank done")
    # R-MAT-21 symmetrised, CC
    g = gas.Graph(workloads.rmat(21, symmetrize=True))
    V = g.n_vertices
    tiers = g.autotune(gas.PROGRAM_CC)
    lab, _ = g.cc()
    lbuf = gas.DeviceBuffer(4 * V)
    g.cc_device(lbuf.ptr)
    t, st0 = timed_runs(lambda: g.cc_device(lbuf.ptr), reps)
    assert np.array_equal(lbuf.to_numpy(), lab)
    out["rmat21s_cc"] = {"gteps": round(st0.gathered_edges / t / 1e9, 4),
                         "ms": round(t * 1e3, 3), "device_ms": round(st0.ms, 3),
                         "supersteps": st0.supersteps, "E": g.n_edges, "tiers": tiers,
                         "components": int(len(np.unique(lab)))}
    del g
    log("cc done")
    return out


def batched_lines(g22, n_plain=8):
    """64-source batches on R-MAT-22 (symmetrised, weighted) against the
    plain engine on the same sources one at a time."""
    V = g22.n_vertices
    csr_off = g22.host_csr()[0]
    deg = np.diff(csr_off)
    rng = np.random.default_rng(workloads.SEED_SOURCES)
    srcs = rng.choice(np.nonzero(deg > 0)[0], 64, replace=False).astype(np.int32)
    out = {}
    buf = gas.DeviceBuffer(4 * V)
    for kind, name, unreached in ((gas.BATCH_BFS, "bfs", -1),
                                  (gas.BATCH_SSSP, "sssp", gas.SSSP_UNREACHED)):
        b = gas.Batch(g22, kind)
        lanes = np.arange(64, dtype=np.int32)
        b.add(lanes, srcs)  # warm (allocations, code objects)
        b.run()
        b.release((1 << 64) - 1)
        ts = []
        for _ in range(3):
            b.add(lanes, srcs)
            t = time.perf_counter()
            st = b.run()
            ts.append(time.perf_counter() - t)
            if _ < 2:
                b.release((1 << 64) - 1)
        t_b = float(np.median(ts))
        er = 0
        for lane, res in b.results(range(64)).items():
            er += reached_edges(res, csr_off, unreached)
        b.release((1 << 64) - 1)
        del b
        fn = g22.bfs_device if kind == gas.BATCH_BFS else g22.sssp_device
        fn(int(srcs[0]), buf.ptr)
        tp = []
        for s in srcs[:n_plain]:
            t = time.perf_counter()
            fn(int(s), buf.ptr)
            tp.append(time.perf_counter() - t)
        t_p = float(np.mean(tp))
        out[f"rmat22s_{name}64"] = {
            "batch_ms": round(t_b * 1e3, 3), "batch_qps": round(64 / t_b, 1),
            "batch_gteps": round(er / t_b / 1e9, 4), "rounds": st.supersteps,
            "plain_ms_per_query": round(t_p * 1e3, 3), "plain_qps": round(1 / t_p, 1),
            "qps_vs_plain": round(t_p * 64 / t_b, 2)}
    return out


def served_lines(graphs, rates=(100, 1000, 5000, 20000), mix=(0.5, 0.3, 0.2)):
    srv = gas.Server(graphs, memory_budget=64 << 30)
    serve.serve(srv, serve.make_stream(1000, 64, graphs, mix=mix, seed=1))  # warm engines
    ref = serve.make_stream(1000, 120, graphs, mix=mix, seed=7)
    svc = serve.plain_service_times(graphs, ref)
    out = {"graphs": ["road1024", "rmat22s"],
           "mix": "/".join(str(round(100 * m)) for m in mix) + " bfs/sssp/ppr",
           "ppr_iterations": 10, "memory_budget_gib": 64,
           "plain_mean_service_ms": round(float(svc.mean() * 1e3), 3),
           "plain_saturated_qps": round(float(1 / svc.mean()), 1), "rates": {}}
    rng = np.random.default_rng(11)
    for rate in rates:
        n = int(min(3000, max(200, rate)))
        st = serve.make_stream(rate, n, graphs, mix=mix, seed=rate)
        res, lat, _ = serve.serve(srv, st)
        # plain engine, FIFO: service time of each query drawn from the
        # measured service times of the same kind and graph
        s_plain = np.zeros(n)
        for k in (0, 1, 2):
            for gi in range(len(graphs)):
                m = (st["kinds"] == k) & (st["graph_ids"] == gi)
                pool = svc[(ref["kinds"] == k) & (ref["graph_ids"] == gi)]
                if not len(pool):
                    pool = svc
                s_plain[m] = rng.choice(pool, int(m.sum()))
        lat_p, end_p = serve.fifo_latencies(st["arrivals"], s_plain)
        out["rates"][str(rate)] = {
            "served": {k: res[k] for k in ("qps", "p50_ms", "p99_ms", "max_concurrent",
                                           "rounds", "solo")},
            "plain_fifo": {"qps": round(n / end_p, 2),
                           "p50_ms": round(float(np.percentile(lat_p, 50)), 3),
                           "p99_ms": round(float(np.percentile(lat_p, 99)), 3)},
            "qps_vs_plain": round(res["qps"] / (n / end_p), 2)}
        log(f"served rate {rate} ({out['mix']}) done")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--extra", dest="extra", action="store_true", default=True)
    ap.add_argument("--no-extra", dest="extra", action="store_false")
    ap.add_argument("--no-cpu", dest="cpu", action="store_false", default=True)
    ap.add_argument("--lines", default="headline,apps,batched,served",
                    help="comma-separated subset of the extra lines to run "
                         "(the headline always runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one device copy of 256 MiB, whose FETCH_SIZE / "
                         "WRITE_SIZE readings scale the counters of the engine's kernels")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs on a single MI355X here "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if gas.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    el = workloads.road()
    g = gas.Graph(el)
    h = headline(g, args.steps, args.warmup)
    log("headline done")
    gteps = h["er"] * args.steps / h["total_s"] / 1e9
    line = {
        "metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(h["total_s"] / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator, graph resident in HBM)",
        "config": {"workload": "road1024: BFS + SSSP from vertex 0, 1024x1024 road-like grid, "
                               f"V={g.n_vertices}, E={g.n_edges}, weights U[1,64]",
                   "edges_traversed_per_step": h["er"], "tiers": h["tiers"]},
        "step_detail": {
            "bfs_device_ms": round(h["bfs"].ms, 4), "sssp_device_ms": round(h["sssp"].ms, 4),
            "bfs_supersteps": h["bfs"].supersteps, "sssp_supersteps": h["sssp"].supersteps,
            "ms_per_step_with_kernel_events": round(h["total_timed_s"] / args.steps * 1e3, 4)},
        "roofline": roofline_entry(h["dom"], h["table"], "road1024"),
        "kernels": h["table"],
    }
    del g
    if args.cpu:
        line["cpu_baseline"] = cpu_baseline(el)
        log("cpu baseline done")
    del el
    want = set(args.lines.split(","))
    if args.extra:
        if "apps" in want:
            line["apps"] = app_lines(min(args.steps, 5))
        if want & {"batched", "served"}:
            road = gas.Graph(workloads.road())
            g22 = gas.Graph(workloads.rmat(22, symmetrize=True, weights=True))
            log("rmat22 built")
            # batch autotuner: BFS pull divisor and SSSP distance step per graph
            tuned = {}
            for name, g in (("road1024", road), ("rmat22s", g22)):
                tuned[name] = {"bfs_pull_divisor": g.autotune_batch(gas.BATCH_BFS),
                               "sssp_delta": g.autotune_batch(gas.BATCH_SSSP)}
            line["batch_tuning"] = tuned
            log(f"batch tuning {tuned}")
            if "batched" in want:
                line["batched"] = batched_lines(g22)
                log("batched done")
            if "served" in want:
                line["served"] = served_lines([road, g22])
                # BFS / SSSP only (the 50/30 part of the mix), at saturation
                line["served_bfs_sssp"] = served_lines([road, g22], rates=(20000,),
                                                       mix=(0.625, 0.375, 0.0))
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        # one device-to-device copy of 256 MiB (the runtime's copy kernel)
        a, b = gas.DeviceBuffer(256 << 20), gas.DeviceBuffer(256 << 20)
        gas.device_copy(b, a)
        log("copy calibration: 268435456 bytes read, 268435456 written")


if __name__ == "__main__":
    main()
