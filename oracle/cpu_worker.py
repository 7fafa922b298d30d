"""One CPU instance of the reference engine for the all-core baselines --
TEST INFRASTRUCTURE ONLY (run by bench.py's cpu_baseline leg, never by the
product).

    python -m oracle.cpu_worker SPEC.json

SPEC: {"prefix": path prefix of <prefix>_{s,d,w}.npy edge arrays,
       "n_vertices": V, "kind": 0 BFS | 1 SSSP | 2 personalised PageRank,
       "sources": [...], "budget_s": seconds, "iterations": PPR sweeps}

Binds one engine of the reference (refgas.h compiled into oracle/_ref) to
the graph once, then answers queries from `sources` in turn until the
budget is spent (at least one).  Prints one JSON line: queries answered and
the summed microseconds of the reference's Took windows (frontier
initialisation through results, sssp.cu:77-81), and the wall seconds of the
query loop.  Imports nothing that touches a GPU.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

import oracle


def main(spec_path: str) -> None:
    with open(spec_path) as f:
        spec = json.load(f)
    pre = spec["prefix"]
    s = np.load(pre + "_s.npy", mmap_mode="r")
    d = np.load(pre + "_d.npy", mmap_mode="r")
    w = np.load(pre + "_w.npy", mmap_mode="r") if spec["kind"] == 1 else None
    t = time.perf_counter()
    eng = oracle.RefPrepared(spec["kind"], spec["n_vertices"], s, d, w)
    prep_s = time.perf_counter() - t
    n, us = 0, 0
    t0 = time.perf_counter()
    for src in spec["sources"]:
        _, u = eng.query(int(src), int(spec.get("iterations", 10)))
        us += u
        n += 1
        if time.perf_counter() - t0 >= spec["budget_s"]:
            break
    wall = time.perf_counter() - t0
    print(json.dumps({"queries": n, "took_us": us, "wall_s": wall, "prepare_s": prep_s}),
          flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
