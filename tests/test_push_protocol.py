"""Host emulation of the push superstep protocol of min-programs (k_push in
include/gas/kernels.hpp) against the oracle's pull engine (refgas.h:143-210
restated in oracle/gasref.c).

The emulation follows the kernel's rules step by step, with the order of
every write inside a superstep shuffled (the GPU's atomics land in any
order):
  * state is double-buffered: pushes read the snapshot A and lower B;
  * a frontier vertex first lowers B to its own value (its "self copy":
    B still holds the value of two supersteps ago there);
  * a push x to u is applied, and u marked as changed for the next
    superstep, exactly when x < A[u];
  * the first superstep runs on every vertex, the later ones on the marked
    vertices; the reported frontier size counts the vertices that push
    (those with out-edges), and supersteps enqueued past the first empty
    report change nothing.
Results and superstep counts must equal the pull's.  A planted mistake (no
self copy) must be caught by the same comparison.
"""
import numpy as np
import pytest

import oracle
from vertexapi2_amd import workloads

INF = 2147483647 - 100000  # SSSP's unreached value (sssp.cu:32-33)


def _csr(n, s, d, w):
    o = np.argsort(s, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(s, minlength=n))])
    return off, d[o], (w[o] if w is not None else None)


def push_run(n, s, d, w, init, skip_inf, rng, extra=2, self_copy=True):
    off, dst, wc = _csr(n, s, d, w)
    A, B = init.copy(), init.copy()
    fl = [np.zeros(n, bool), np.zeros(n, bool)]
    p, first, done_at, steps = 0, True, None, 0
    while True:
        if first:
            front = np.arange(n)
        else:
            front = np.nonzero(fl[p])[0]
            fl[p][:] = False
        events, active = [], 0
        for v in front:
            a = A[v]
            deg = off[v + 1] - off[v]
            if deg > 0:
                active += 1
            if not first and self_copy:
                events.append((0, v, a))
            if deg > 0 and not (skip_inf and a >= INF):
                for k in range(off[v], off[v + 1]):
                    events.append((1, dst[k], a + (wc[k] if wc is not None else 0)))
        for i in rng.permutation(len(events)):
            kind, u, x = events[i]
            if kind == 0:
                B[u] = min(B[u], x)
            elif x < A[u]:
                B[u] = min(B[u], x)
                fl[p ^ 1][u] = True
        report = n if first else active
        A, B = B, A
        p ^= 1
        first = False
        steps += 1
        if done_at is None and report == 0:
            done_at = steps - 1
        if done_at is not None and steps >= done_at + 1 + extra:
            return A, done_at
        if steps > 2 * n + 8:  # more supersteps than any pull run takes
            return A, None


def _cases():
    rng = np.random.default_rng(0)
    out = []
    el = workloads.road(32, 48, seed=3)
    out.append(("road32x48", el.n_vertices, *el.arrays()))
    for t in range(4):
        n = int(rng.integers(5, 300))
        m = int(rng.integers(0, 3 * n))
        s = rng.integers(0, n, m).astype(np.int32)
        d = rng.integers(0, n, m).astype(np.int32)
        w = rng.integers(1, 50, m).astype(np.int32)
        out.append((f"rand{t}", n, s, d, w))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_push_protocol_equals_pull(case):
    name, n, s, d, w = case
    rng = np.random.default_rng(1)
    for src in (0, n // 2):
        init = np.full(n, INF, np.int64)
        init[src] = 0
        for extra in (0, 3):
            res, steps = push_run(n, s, d, w, init, True, rng, extra=extra)
            ref, st = oracle.sssp(n, s, d, w, src)
            assert np.array_equal(res, ref), (name, src)
            assert steps == st.supersteps, (name, src)
    res, steps = push_run(n, s, d, None, np.arange(n, dtype=np.int64), False, rng)
    ref, st = oracle.cc(n, s, d)
    assert np.array_equal(res, ref), name
    assert steps == st.supersteps, name


def test_push_protocol_without_self_copy_is_caught():
    el = workloads.road(12, 16, seed=3)
    n = el.n_vertices
    s, d, w = el.arrays()
    init = np.full(n, INF, np.int64)
    init[0] = 0
    res, steps = push_run(n, s, d, w, init, True, np.random.default_rng(2), self_copy=False)
    ref, st = oracle.sssp(n, s, d, w, 0)
    assert not (np.array_equal(res, ref) and steps == st.supersteps)
