"""Benchmark of the GAS hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extra] [--no-cpu]

Headline workload: BFS and SSSP from vertex 0 on the 1024x1024 road-like
graph (3% of links dropped, 1% long links, int32 weights U[1,64]).  One step
= one BFS run plus one SSSP run of the plain engine to completion on the
graph already resident in HBM: initial vertex state written on the device,
supersteps until the frontier is empty, result copied into a device buffer
(gas_bfs_device / gas_sssp_device).  Degree tiers are autotuned once per
program before the warmup.

value = traversed edges per second in GTEPS: (E_r(BFS) + E_r(SSSP)) / time,
with E_r the summed out-degree of the vertices the run reaches, and time the
host wall clock around the K timed steps (each step ends synchronised).
results_on_host: the same K steps through gas_bfs / gas_sssp, the
reference's "Took" window (frontier initialisation through results in host
memory, sssp.cu:77-81, bfs.cu:100-111): host-side initial state, upload,
run, download.

roofline: after the timed steps, the same K steps run again with every
engine kernel launch bracketed by HIP events on the engine's stream; the
dominant kernel (largest summed time) is reported with its algorithmic bytes
per launch (model below, unit counts from the engine's device statistics)
over its average event-bracketed duration (frac).  Event brackets include
dispatch latency, so the entry also gives the same bytes over the kernel's
average duration in the rocprofv3 kernel-trace summary of this workload
(profiles/kstats_<workload>.json: frac_rocprof).  traffic: HBM bytes per
launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary
(profiles/pmc_<workload>.json), or null.
roofline_8d: the whole step against the compulsory bytes of the scope
table: BFS 20 V_r + 8 E_r, SSSP 20 V_r + 12 E_r (V_r reached vertices, each
datum once, re-relaxations not credited), PageRank (8 E + 16 V) per sweep,
CC sum over supersteps of 16 A + 8 G + 4 S + 8 R.
cpu_baseline: the reference's own CPU engine (refgas.h built from its
sources into oracle/_ref) on the same workload, one host core, with the CPU
model.  The batched and served lines add an all-core figure: one prepared
reference engine per core of the GPU's CPU share (16 threads), all running
at once, each answering queries of one (kind, graph) for a bounded time.

Further lines under "apps" (unless --no-extra): grid BFS (64x64), R-MAT-21
PageRank (20 sweeps, with its error against an FP64 power iteration
computed here with numpy) and CC (symmetrised), each with its dominant
kernel's roofline and roofline_8d; "batched": 64-source BFS and SSSP batches
on symmetrised weighted R-MAT-22 against the plain engine answering the same
sources one at a time; "served": Poisson query streams (50/30/20 BFS/SSSP/PPR
over road1024 and R-MAT-22) at 100 ... 20,000 queries/s replayed through the
continuous-batching server and through the plain engine answering the same
stream itself one query at a time, with QPS, p50 and p99.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import vertexapi2_amd as gas  # noqa: E402
from vertexapi2_amd import serve, workloads  # noqa: E402

METRIC = "GTEPS + % HBM roofline per app; batched-served QPS & p99 vs plain 1-source"
HBM_PEAK_GBS = 8000.0
SERVED_RATES = (100, 200, 500, 1000, 2000, 5000, 10000, 20000)

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


def reached(values, csr_off, unreached):
    """(V_r, E_r): reached vertices and their summed out-degree."""
    deg = np.diff(csr_off).astype(np.int64)
    m = values != unreached
    return int(m.sum()), int(deg[m].sum())


def bytes_8d(app, vr, er):
    """Compulsory bytes of one BFS / SSSP run (scope-table model)."""
    return 20 * vr + (12 if app == "sssp" else 8) * er


def frac_entry(byts, seconds):
    gbs = byts / seconds / 1e9
    return {"bytes": int(byts), "seconds": round(seconds, 6), "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_share():
    """Worker processes for all-core figures: the CPU share of one GPU on
    the box (16 threads), fewer if fewer are available."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


# ---------------------------------------------------------------------------
# Algorithmic bytes per kernel (each datum counted once, lower bounds).
#   k_gather_vp   per gathered vertex: 8 (CSC offset pair) + 4 (result);
#                 per in-edge: 4 (source id) + edge value + source read (SRC)
#   k_hgather     the same for the hub vertices and their in-edges (the
#                 engine counts both: hub_vertices_total, hub_edges_total)
#   k_apply       per active vertex: 4 (slot id) + 4 (gather result, when
#                 there is a gather) + state; per scattered out-edge: 4
#                 (destination id) + 1 (activation byte)
#   k_compact1    V (activation bytes); k_compact2: V + 4 per listed id
#   k_fused       per active vertex: 8 (offsets) + state read + state write;
#                 per in-edge as k_gather_vp; per scattered edge 5
#   k_push        (min-programs pushed by their changed vertices) per
#                 frontier vertex: 1 (activation byte) + state read + state
#                 write (its own entry of the other buffer); per pushed
#                 edge: 4 (destination id) + edge value + destination state
#                 read + 1 (mark)
#   k_small       (runs that never left the one-workgroup mode) per active
#                 vertex: 4 (list id) + 8 (CSR offsets) + state read + write,
#                 with a gather + 8 (CSC offsets); per in-edge as
#                 k_gather_vp; per scattered edge 5
def kernel_models(app, st, n_vertices, have_gather):
    """{kernel: total algorithmic bytes over the run}."""
    vd, ed, sv = VD[app], ED.get(app, 0), SRC[app]
    A, Eg, Es = st.active_total, st.gathered_edges, st.scattered_edges
    Hv, He = st.hub_vertices_total, st.hub_edges_total
    out = {}
    ks = st.kernels()
    if set(ks) <= {"k_small"}:
        out["k_small"] = (A * (12 + 2 * vd + (8 if have_gather else 0)) +
                          Eg * (4 + ed + sv) + Es * 5)
        return out
    if have_gather and "k_push" in ks:
        # every superstep pushed: the frontier is the changed vertices
        # (returned_total counts them, scattered_edges the pushes)
        out["k_push"] = st.returned_total * (1 + 2 * vd) + Es * (4 + ed + vd + 1)
        return out
    if have_gather and "k_fused" in ks and "k_gather_vp" not in ks:
        # every superstep ran as one launch: gather + apply + scatter
        out["k_fused"] = A * (8 + 2 * vd) + Eg * (4 + ed + sv) + Es * 5
        return out
    if have_gather:
        out["k_gather_vp"] = (A - Hv) * (8 + GR) + (Eg - He) * (4 + ed + sv)
        out["k_hgather"] = Hv * (8 + GR) + He * (4 + ed + sv)
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


# Timed regions, cut out of rocprofv3 kernel traces by k_trace_mark
# dispatches of grid 64 * (2 * id - 1) (start) and 64 * (2 * id) (end).
REGIONS = {"road1024": 1, "grid64_bfs": 2, "rmat21_pagerank20": 3, "rmat21s_cc": 4,
           "rmat22s_batched_bfs": 5, "rmat22s_batched_sssp": 6,
           "road1024_batched_bfs": 7, "road1024_batched_sssp": 8, "copy_calibration": 9}
REGIONS.update({f"served_{r}": 10 + i for i, r in enumerate(SERVED_RATES)})


def mark(region, end=False):
    gas.trace_mark(2 * REGIONS[region] - (0 if end else 1))


def _profile(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def roofline_entry(dom, table, workload):
    e = table[dom]
    ks = _profile(f"kstats_{workload}.json")
    # the engine books k_small_lds (the LDS-resident form) in k_small's slot
    kname = dom
    if ks is not None and dom not in ks.get("kernels", {}) and f"{dom}_lds" in ks["kernels"]:
        kname = f"{dom}_lds"
    traffic = None
    rec = _profile(f"pmc_{workload}.json")
    if rec is not None:
        k = rec.get("kernels", {}).get(kname)
        if k is not None:
            traffic = k.get("hbm_bytes_per_launch")
    achieved = e.get("achieved_gbs")
    out = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": e.get("frac"), "traffic": traffic,
           "avg_launch_us": e["avg_launch_us"], "launches": e["launches"],
           "algorithmic_bytes_per_launch": e.get("algorithmic_bytes_per_launch")}
    b = e.get("algorithmic_bytes_per_launch")
    if ks is not None and b and kname in ks.get("kernels", {}):
        us = ks["kernels"][kname]["avg_us"]
        gbs = b / (us * 1e-6) / 1e9
        out.update({"rocprof_kernel": kname, "rocprof_avg_us": us,
                    "achieved_rocprof": round(gbs, 2),
                    "frac_rocprof": round(gbs / HBM_PEAK_GBS, 5)})
    return out


def fp64_pagerank(n, s, d, iterations, source=-1):
    """FP64 power iteration of rank = tele + 0.85 sum rank/outdeg over
    in-edges (pagerank.cu:50, 62), tele = 0.15 everywhere (source < 0) or at
    the source only (personalised), written here with numpy."""
    s = np.asarray(s, np.int64)
    d = np.asarray(d, np.int64)
    outdeg = np.bincount(s, minlength=n).astype(np.float64)
    tele = np.full(n, 0.15) if source < 0 else np.zeros(n)
    if source >= 0:
        tele[source] = 0.15
    r = tele.copy()
    inv = np.zeros(n)
    nz = outdeg > 0
    inv[nz] = 1.0 / outdeg[nz]
    for _ in range(iterations):
        r = tele + 0.85 * np.bincount(d, weights=(r * inv)[s], minlength=n)
    return r


def fp_errors(x, ref):
    x = np.asarray(x, np.float64)
    diff = np.abs(x - ref)
    nz = ref != 0
    return {"rel_l1": float(diff.sum() / np.abs(ref).sum()),
            "max_rel": float((diff[nz] / np.abs(ref[nz])).max(initial=0.0)),
            "max_abs": float(diff.max(initial=0.0))}


# ---------------------------------------------------------------------------
def headline(g, steps, warmup):
    tiers = {"bfs": g.autotune(gas.PROGRAM_BFS), "sssp": g.autotune(gas.PROGRAM_SSSP)}
    V = g.n_vertices
    dbuf, sbuf = gas.DeviceBuffer(4 * V), gas.DeviceBuffer(4 * V)
    for _ in range(warmup):
        g.bfs_device(0, dbuf.ptr)
        g.sssp_device(0, sbuf.ptr)
    # timed region
    mark("road1024")
    t0 = time.perf_counter()
    for _ in range(steps):
        sb0 = g.bfs_device(0, dbuf.ptr)
        ss0 = g.sssp_device(0, sbuf.ptr)
    total = time.perf_counter() - t0
    mark("road1024", end=True)
    # results on the host (the reference's Took window)
    g.bfs(0)
    g.sssp(0)
    t1 = time.perf_counter()
    for _ in range(steps):
        g.bfs(0)
        g.sssp(0)
    total_host = time.perf_counter() - t1
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
    vb, eb = reached(depth, csr_off, -1)
    vs, es = reached(dist, csr_off, gas.SSSP_UNREACHED)
    stats, models = [], []
    for sb, ss in sts:
        stats += [sb, ss]
        models += [kernel_models("bfs", sb, V, False), kernel_models("sssp", ss, V, True)]
    dom, table = roofline_table(stats, models)
    return {"total_s": total, "total_host_s": total_host, "total_timed_s": total_timed,
            "er": eb + es, "bytes_8d": bytes_8d("bfs", vb, eb) + bytes_8d("sssp", vs, es),
            "tiers": tiers, "dom": dom, "table": table, "bfs": sb0, "sssp": ss0}


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
        er_total += reached(depth, off, -1)[1] + reached(dist, off, gas.SSSP_UNREACHED)[1]
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 20:
            break
    gteps = er_total / (us_total * 1e-6) / 1e9
    return {"value": round(gteps, 5), "unit": "GTEPS", "cores": 1, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": f"{reps} x (BFS + SSSP from vertex 0) on the full 1024x1024 road graph "
                      f"with refgas.h (GASEngineRef) single-threaded; {us_total / reps / 1e3:.1f} "
                      f"ms per step in the reference's Took window"}


def cpu_all_cores(graphs, budget_s=15.0, seed=5):
    """Per (graph, kind) mean service time of the reference engine with every
    core of the CPU share busy at once: one worker process per core, worker
    i answering queries of combination i mod 6 (BFS, SSSP, PPR on each of the
    two graphs) from uniform sources with out-edges, for budget_s seconds
    after binding its engine to the graph.  graphs: [(name, Edges)]."""
    import oracle
    if not oracle.ref_available():
        return None
    nw = cpu_share()
    tmp = tempfile.mkdtemp(prefix="gas_cpu_")
    combos, files = [], []
    try:
        for gi, (name, el) in enumerate(graphs):
            s, d, w = el.arrays()
            pre = os.path.join(tmp, name)
            np.save(pre + "_s.npy", s)
            np.save(pre + "_d.npy", d)
            np.save(pre + "_w.npy", w)
            files += [pre + "_s.npy", pre + "_d.npy", pre + "_w.npy"]
            deg = np.bincount(s, minlength=el.n_vertices)
            cand = np.nonzero(deg > 0)[0]
            for kind in (0, 1, 2):
                combos.append((name, kind, pre, el.n_vertices, cand))
        rng = np.random.default_rng(seed)
        procs = []
        env = dict(os.environ, PYTHONPATH=ROOT, OMP_NUM_THREADS="1")
        for i in range(nw):
            name, kind, pre, n, cand = combos[i % len(combos)]
            spec = {"prefix": pre, "n_vertices": int(n), "kind": kind, "budget_s": budget_s,
                    "iterations": 10,
                    "sources": rng.choice(cand, 256).astype(int).tolist()}
            path = os.path.join(tmp, f"spec{i}.json")
            with open(path, "w") as f:
                json.dump(spec, f)
            files.append(path)
            procs.append((i % len(combos), subprocess.Popen(
                [sys.executable, "-m", "oracle.cpu_worker", path], cwd=ROOT, env=env,
                stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        agg = {}
        for ci, p in procs:
            out, err = p.communicate(timeout=900)
            if p.returncode != 0:
                raise RuntimeError(f"cpu worker failed: {err[-2000:]}")
            r = json.loads(out.strip().splitlines()[-1])
            a = agg.setdefault(ci, {"queries": 0, "took_us": 0, "instances": 0})
            a["queries"] += r["queries"]
            a["took_us"] += r["took_us"]
            a["instances"] += 1
    finally:
        for f in files:
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(tmp)
    out = {}
    for ci, a in agg.items():
        name, kind = combos[ci][0], combos[ci][1]
        out[(name, kind)] = {"mean_ms": a["took_us"] / a["queries"] / 1e3,
                             "queries": a["queries"], "instances": a["instances"]}
    return {"workers": nw, "cpu_model": cpu_model(), "combos": out, "budget_s": budget_s}


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
    mark("grid64_bfs")
    t, _ = timed_runs(lambda: g.bfs_device(0, buf.ptr), reps)
    mark("grid64_bfs", end=True)
    gas.set_kernel_timing(True)
    st = g.bfs_device(0, buf.ptr)
    gas.set_kernel_timing(False)
    vr, er = reached(buf.to_numpy(), g.host_csr()[0], -1)
    dom, table = roofline_table([st], [kernel_models("bfs", st, V, False)])
    out["grid64_bfs"] = {"gteps": round(er / t / 1e9, 5), "ms": round(t * 1e3, 4),
                         "device_ms": round(st.ms, 4), "supersteps": st.supersteps,
                         "roofline": roofline_entry(dom, table, "grid64_bfs"),
                         "roofline_8d": frac_entry(bytes_8d("bfs", vr, er), t)}
    del g, buf
    log("grid64 done")
    # R-MAT-21 PageRank, 20 sweeps
    el = workloads.rmat(21)
    g = gas.Graph(el)
    V, E = g.n_vertices, g.n_edges
    tiers = g.autotune(gas.PROGRAM_PAGERANK)
    _, res, _ = g.pagerank(20)  # host-result run: the per-sweep residuals
    rbuf = gas.DeviceBuffer(4 * V)
    g.pagerank_device(rbuf.ptr, 20)
    mark("rmat21_pagerank20")
    t, st0 = timed_runs(lambda: g.pagerank_device(rbuf.ptr, 20), reps)
    mark("rmat21_pagerank20", end=True)
    gas.set_kernel_timing(True)
    st = g.pagerank_device(rbuf.ptr, 20)
    gas.set_kernel_timing(False)
    rank = rbuf.to_numpy(np.float32)
    dom, table = roofline_table([st], [kernel_models("pagerank", st, V, True)])
    s, d, _ = el.arrays()
    err = fp_errors(rank, fp64_pagerank(V, s, d, 20))
    out["rmat21_pagerank20"] = {
        "gteps": round(20 * E / t / 1e9, 4), "ms": round(t * 1e3, 3),
        "device_ms": round(st0.ms, 3), "V": V, "E": E, "tiers": tiers,
        "final_residual_l1": float(res[-1]), "l1_vs_fp64": err,
        "roofline": roofline_entry(dom, table, "rmat21_pagerank20"),
        "roofline_8d": frac_entry(20 * (8 * E + 16 * V), t), "kernels": table}
    del g, el, s, d
    log("pagerank done")
    # R-MAT-21 symmetrised, CC
    g = gas.Graph(workloads.rmat(21, symmetrize=True))
    V = g.n_vertices
    tiers = g.autotune(gas.PROGRAM_CC)
    lab, _ = g.cc()
    lbuf = gas.DeviceBuffer(4 * V)
    g.cc_device(lbuf.ptr)
    mark("rmat21s_cc")
    t, st0 = timed_runs(lambda: g.cc_device(lbuf.ptr), reps)
    mark("rmat21s_cc", end=True)
    assert np.array_equal(lbuf.to_numpy(), lab)
    gas.set_kernel_timing(True)
    st = g.cc_device(lbuf.ptr)
    gas.set_kernel_timing(False)
    dom, table = roofline_table([st], [kernel_models("cc", st, V, True)])
    b8 = (16 * st0.active_total + 8 * st0.gathered_edges + 4 * st0.scattered_edges +
          8 * st0.returned_total)
    out["rmat21s_cc"] = {"gteps": round(st0.gathered_edges / t / 1e9, 4),
                         "ms": round(t * 1e3, 3), "device_ms": round(st0.ms, 3),
                         "supersteps": st0.supersteps, "E": g.n_edges, "tiers": tiers,
                         "components": int(len(np.unique(lab))),
                         "roofline": roofline_entry(dom, table, "rmat21s_cc"),
                         "roofline_8d": frac_entry(b8, t), "kernels": table}
    del g
    log("cc done")
    return out


def batched_lines(g22, n_plain=8, name22="rmat22s"):
    """64-source batches on a weighted graph (R-MAT-22 symmetrised, or the
    road grid) against the plain engine on the same sources one at a time."""
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
        mark(f"{name22}_batched_{name}")
        for _ in range(3):
            b.add(lanes, srcs)
            t = time.perf_counter()
            st = b.run()
            ts.append(time.perf_counter() - t)
            if _ < 2:
                b.release((1 << 64) - 1)
        mark(f"{name22}_batched_{name}", end=True)
        t_b = float(np.median(ts))
        er = 0
        any_lane = np.zeros(V, bool)
        for lane, res in b.results(range(64)).items():
            m = res != unreached
            any_lane |= m
            er += int(deg[m].astype(np.int64).sum())
        # compulsory bytes of the batch, each datum once: per vertex some
        # lane reaches, its offsets and its lane row (BFS: 16-bit depths of
        # 64 lanes written + visited / frontier masks; SSSP: 64 int32
        # distances read and written); per out-edge of those, its
        # destination id (+ weight) and the destination's masks (BFS) or
        # row (SSSP) once
        va = int(any_lane.sum())
        ea = int(deg[any_lane].astype(np.int64).sum())
        if kind == gas.BATCH_BFS:
            b8 = va * (8 + 16 + 2 * 64) + ea * (4 + 16)
        else:
            b8 = va * (8 + 2 * 4 * 64) + ea * (8 + 4 * 64)
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
        line = {
            "batch_ms": round(t_b * 1e3, 3), "batch_qps": round(64 / t_b, 1),
            "batch_gteps": round(er / t_b / 1e9, 4), "rounds": st.supersteps,
            "plain_ms_per_query": round(t_p * 1e3, 3), "plain_qps": round(1 / t_p, 1),
            "qps_vs_plain": round(t_p * 64 / t_b, 2),
            "roofline_8d": frac_entry(b8, t_b)}
        ks = _profile(f"kstats_{name22}_batched_{name}.json")
        pm = _profile(f"pmc_{name22}_batched_{name}.json")
        if ks is not None and pm is not None:
            kern = ks.get("dominant")
            k = ks["kernels"].get(kern) if kern else None
            tr = pm.get("kernels", {}).get(kern) if kern else None
            if k and tr:
                gbs = tr["hbm_bytes_per_launch"] / (k["avg_us"] * 1e-6) / 1e9
                line["roofline"] = {
                    "kernel": kern, "bound": "hbm", "basis": "measured HBM traffic "
                    "(rocprofv3 FETCH_SIZE + WRITE_SIZE) per launch over the kernel's "
                    "rocprofv3 average duration", "achieved": round(gbs, 2),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 5),
                    "traffic": tr["hbm_bytes_per_launch"], "avg_launch_us": k["avg_us"]}
        out[f"{name22}_{name}64"] = line
    return out


def served_lines(graphs, edge_lists=None, rates=SERVED_RATES, mix=(0.5, 0.3, 0.2),
                 err_rate=1000):
    srv = gas.Server(graphs, memory_budget=64 << 30)
    serve.serve(srv, serve.make_stream(1000, 64, graphs, mix=mix, seed=1))  # warm engines
    out = {"graphs": ["road1024", "rmat22s"],
           "mix": "/".join(str(round(100 * m)) for m in mix) + " bfs/sssp/ppr",
           "ppr_iterations": 10, "memory_budget_gib": 64,
           "plain": "the plain engine serving the same stream one query at a time "
                    "in arrival order (gas_*_device, results in HBM)", "rates": {}}
    for rate in rates:
        n = int(min(2000, max(200, rate)))
        st = serve.make_stream(rate, n, graphs, mix=mix, seed=rate)
        keep = None
        if edge_lists is not None and rate == err_rate and mix[2] > 0:
            ppr = np.nonzero(st["kinds"] == gas.BATCH_PPR)[0]
            keep = [int(q) for gi in range(len(graphs))
                    for q in ppr[st["graph_ids"][ppr] == gi][:2 if gi == 0 else 1]]
        region = f"served_{rate}" if mix == (0.5, 0.3, 0.2) else None
        if region:
            mark(region)
        res, lat, results = serve.serve(srv, st, keep=keep)
        if region:
            mark(region, end=True)
        lat_p, end_p = serve.plain_replay(graphs, st)
        plain_qps = n / end_p
        entry = {
            "queries": n,
            "served": {k: res[k] for k in ("qps", "p50_ms", "p99_ms", "max_concurrent",
                                           "rounds", "solo")},
            "plain": {"qps": round(plain_qps, 2),
                      "p50_ms": round(float(np.percentile(lat_p, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat_p, 99)), 3)},
            "qps_vs_plain": round(res["qps"] / plain_qps, 2)}
        if keep:
            errs = []
            for q in keep:
                s, d, _ = edge_lists[st["graph_ids"][q]].arrays()
                nv = graphs[st["graph_ids"][q]].n_vertices
                e = fp_errors(results[q], fp64_pagerank(nv, s, d, 10, int(st["sources"][q])))
                e["graph"] = out["graphs"][st["graph_ids"][q]]
                errs.append(e)
            entry["ppr_l1_vs_fp64"] = errs
        out["rates"][str(rate)] = entry
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
    step_s = h["total_s"] / args.steps
    host_s = h["total_host_s"] / args.steps
    line = {
        "metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator, graph resident in HBM)",
        "config": {"workload": "road1024: BFS + SSSP from vertex 0, 1024x1024 road-like grid, "
                               f"V={g.n_vertices}, E={g.n_edges}, weights U[1,64]",
                   "edges_traversed_per_step": h["er"], "tiers": h["tiers"]},
        "step_detail": {
            "bfs_device_ms": round(h["bfs"].ms, 4), "sssp_device_ms": round(h["sssp"].ms, 4),
            "bfs_supersteps": h["bfs"].supersteps, "sssp_supersteps": h["sssp"].supersteps,
            "ms_per_step_with_kernel_events": round(h["total_timed_s"] / args.steps * 1e3, 4)},
        "results_on_host": {"ms_per_step": round(host_s * 1e3, 4),
                            "gteps": round(h["er"] / host_s / 1e9, 4),
                            "window": "gas_bfs + gas_sssp: host initial state, upload, run, "
                                      "download (sssp.cu:77-81)"},
        "roofline": roofline_entry(h["dom"], h["table"], "road1024"),
        "roofline_8d": frac_entry(h["bytes_8d"], step_s),
        "kernels": h["table"],
    }
    del g
    if args.cpu:
        line["cpu_baseline"] = cpu_baseline(el)
        log("cpu baseline done")
    want = set(args.lines.split(","))
    if args.extra:
        if "apps" in want:
            line["apps"] = app_lines(min(args.steps, 5))
        if want & {"batched", "served"}:
            road = gas.Graph(el)
            el22 = workloads.rmat(22, symmetrize=True, weights=True)
            t = time.perf_counter()
            g22 = gas.Graph(el22)
            line["rmat22s_graph_create_s"] = round(time.perf_counter() - t, 3)
            log("rmat22 built")
            # batch autotuner: BFS pull divisor and SSSP distance step per graph
            tuned = {}
            for name, gg in (("road1024", road), ("rmat22s", g22)):
                tuned[name] = {"bfs_pull_divisor": gg.autotune_batch(gas.BATCH_BFS),
                               "sssp_delta": gg.autotune_batch(gas.BATCH_SSSP),
                               "bfs_hub_items": gg.batch_hub_items(gas.BATCH_BFS),
                               "sssp_hub_items": gg.batch_hub_items(gas.BATCH_SSSP),
                               "bfs_item_edges": gg.batch_item_edges(gas.BATCH_BFS),
                               "sssp_item_edges": gg.batch_item_edges(gas.BATCH_SSSP)}
            line["batch_tuning"] = tuned
            log(f"batch tuning {tuned}")
            if "batched" in want:
                line["batched"] = batched_lines(g22)
                # high-diameter graph: 50-100 sparse rounds per batch
                line["batched"].update(batched_lines(road, name22="road1024"))
                log("batched done")
            if "served" in want:
                line["served"] = served_lines([road, g22], [el, el22])
                # BFS / SSSP only (the 50/30 part of the mix), at saturation
                line["served_bfs_sssp"] = served_lines([road, g22], rates=(20000,),
                                                       mix=(0.625, 0.375, 0.0))
            if args.cpu:
                allc = cpu_all_cores([("road1024", el), ("rmat22s", el22)])
                log("cpu all-core done")
                if allc is not None:
                    attach_cpu_all_cores(line, allc)
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        # one device-to-device copy of 256 MiB (the runtime's copy kernel)
        a, b = gas.DeviceBuffer(256 << 20), gas.DeviceBuffer(256 << 20)
        mark("copy_calibration")
        gas.device_copy(b, a)
        mark("copy_calibration", end=True)
        log("copy calibration: 268435456 bytes read, 268435456 written")


def attach_cpu_all_cores(line, allc):
    """All-core CPU figures beside the batched and served lines."""
    nw = allc["workers"]
    combos = allc["combos"]
    desc = (f"{nw} worker processes at once (the CPU share of one GPU), each with one "
            f"engine of the reference (refgas.h) bound to its graph, answering uniform "
            f"sources of one (kind, graph) for {allc['budget_s']:g} s after binding it; "
            f"all-core queries/s = {nw} / mean "
            f"service time")
    per = {f"{g}_{serve.KIND_NAMES[k]}": {"mean_ms": round(v["mean_ms"], 3),
                                          "queries": v["queries"], "instances": v["instances"]}
           for (g, k), v in combos.items()}
    base = {"kind": "reference", "cores": nw, "cpu_model": allc["cpu_model"],
            "sample": desc, "per_kind": per}
    for name, kind in (("rmat22s_bfs64", 0), ("rmat22s_sssp64", 1)):
        b = line.get("batched", {}).get(name)
        c = combos.get(("rmat22s", kind))
        if b is not None and c is not None:
            qps = nw / (c["mean_ms"] / 1e3)
            b["cpu_all_cores"] = dict(base, qps=round(qps, 3),
                                      batch_vs_cpu=round(b["batch_qps"] / qps, 1))
    for key, mix in (("served", (0.5, 0.3, 0.2)), ("served_bfs_sssp", (0.625, 0.375, 0.0))):
        s = line.get(key)
        if s is None:
            continue
        mean = 0.0
        for k, frac in enumerate(mix):
            for g in ("road1024", "rmat22s"):
                c = combos.get((g, k))
                if c is not None and frac > 0:
                    mean += 0.5 * frac * c["mean_ms"]
        if mean > 0:
            s["cpu_all_cores"] = dict(base, saturated_qps=round(nw / (mean / 1e3), 3),
                                      mean_service_ms=round(mean, 3))


if __name__ == "__main__":
    main()
