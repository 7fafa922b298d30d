"""Per-vertex result files and the checks that compare them with golds.

Result files hold one "vid value" line per vertex in id order, as the
reference's apps write them: "%d %d" for BFS depths, SSSP distances and CC
labels (bfs.cu:115-119, sssp.cu:85-89, connected_component.cu:84-88),
"%d %f" for PageRank ranks (pagerank.cu:76-81).

`check_dist` and `check_pagerank` follow the reference's regression
checkers (regressions/checkDist.py:33-48, regressions/checkPageRank.py:
40-82) against golds such as PowerGraph's outputs:

    python -m vertexapi2_amd.resultio dist test gold
    python -m vertexapi2_amd.resultio pagerank test gold

exit with 0 when the files agree and 1 otherwise, like the scripts.
"""
from __future__ import annotations

import sys

import numpy as np

SSSP_UNREACHED = 2147383647  # INT_MAX - 100000 (sssp.cu:33)
GOLD_UNREACHED = 10000000    # unreached distance in the golds (checkDist.py:43)

# histogram bins of |gold - test| and the largest fraction of vertices each
# may hold (checkPageRank.py:77-82): < 1e-4 any, [1e-4, 1e-3) 1%,
# [1e-3, 1e-2) 0.5%, ..., [1, 10) 5e-6, >= 10 none
PR_TOL_VALS = (1e-4, 1e-3, 1e-2, 1e-1, 1.0, 10.0)
PR_TOL_ALLOWED = (1.0, 1e-2, 5e-3, 5e-4, 5e-5, 5e-6, 0.0)


def format_results(values) -> str:
    """The text of a result file for per-vertex `values` (integers as %d,
    floating point as %f)."""
    a = np.asarray(values)
    if np.issubdtype(a.dtype, np.floating):
        return "".join("%d %f\n" % (i, float(x)) for i, x in enumerate(a))
    return "".join("%d %d\n" % (i, int(x)) for i, x in enumerate(a))


def write_results(path: str, values) -> None:
    with open(path, "w") as f:
        f.write(format_results(values))


def read_results(path: str):
    """(ids, values) of a "vid value" file, in file order; values are int64
    when every value parses as an integer, float64 otherwise."""
    ids, vals = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            ids.append(int(parts[0]))
            vals.append(parts[1])
    try:
        v = np.array([int(x) for x in vals], np.int64)
    except ValueError:
        v = np.array([float(x) for x in vals], np.float64)
    return np.array(ids, np.int64), v


def check_dist(test_path: str, gold_path: str):
    """(ok, message): every gold vertex must appear in the test file with
    the same value, a gold 10000000 matching a test 2147383647 (the
    reference's unreached distance).  Both files are read as sorted by
    vertex id; test vertices absent from the gold are ignored."""
    t_ids, t_val = read_results(test_path)
    g_ids, g_val = read_results(gold_path)
    j = 0
    for vid, gv in zip(g_ids.tolist(), g_val.tolist()):
        while j < len(t_ids) and t_ids[j] < vid:
            j += 1
        if j >= len(t_ids) or t_ids[j] != vid:
            return False, "vertex %d is in gold but not in test" % vid
        tv = int(t_val[j])
        if tv != gv and not (gv == GOLD_UNREACHED and tv == SSSP_UNREACHED):
            return False, "%d gold=%d ref=%d" % (vid, gv, tv)
        j += 1
    return True, "No differences found"


def check_pagerank(test_path: str, gold_path: str, tol_vals=PR_TOL_VALS,
                   tol_allowed=PR_TOL_ALLOWED):
    """(ok, message): the fraction of test vertices whose |gold - test|
    falls in each bin of `tol_vals` must not exceed `tol_allowed`; vertices
    missing from the gold count toward the total only."""
    t_ids, t_val = read_results(test_path)
    g_ids, g_val = read_results(gold_path)
    gold = dict(zip(g_ids.tolist(), g_val.astype(np.float64).tolist()))
    hist = [0] * (len(tol_vals) + 1)
    missing = 0
    for vid, val in zip(t_ids.tolist(), t_val.astype(np.float64).tolist()):
        if vid not in gold:
            missing += 1
            continue
        diff = abs(gold[vid] - val)
        pos = len(tol_vals) - 1
        while pos >= 0 and diff < tol_vals[pos]:
            pos -= 1
        hist[pos + 1] += 1
    total = float(len(t_ids))
    for idx, count in enumerate(hist):
        frac = count / total if total else 0.0
        if frac > tol_allowed[idx]:
            return False, "Percentage too high: %g %g" % (tol_allowed[idx], frac)
    note = " (%d test vertices not in gold)" % missing if missing else ""
    return True, "ranks within tolerance" + note


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 3 or argv[0] not in ("dist", "pagerank"):
        print("Usage: python -m vertexapi2_amd.resultio dist|pagerank test gold",
              file=sys.stderr)
        return 1
    fn = check_dist if argv[0] == "dist" else check_pagerank
    ok, msg = fn(argv[1], argv[2])
    print(msg)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
