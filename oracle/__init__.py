"""CPU oracle loader -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  It binds:

  _lib/liboracle.so   the C restatement of the reference CPU engine
                      (gasref.c; refgas.h:30-225, util.h:48-121) and the
                      independent checkers;
  _ref/librefgas.so   the reference's own refgas.h/util.h/util.cpp compiled
                      from its sources (absent when they were not available
                      at build time);
  _ref/librefgraphio.so  the reference's graphio.cpp (graph file loaders and
                      writers), likewise.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "_lib", "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "librefgas.so")
REFIO_LIB = os.path.join(HERE, "_ref", "librefgraphio.so")
SSSP_UNREACHED = 2147483647 - 100000


def build() -> None:
    """make -C oracle (liboracle always; _ref when the reference is present)."""
    res = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")


class Stats(C.Structure):
    _fields_ = [("supersteps", C.c_int64), ("active_total", C.c_int64),
                ("gathered", C.c_int64), ("returned", C.c_int64),
                ("scattered", C.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_o = None
_r = None


def lib() -> C.CDLL:
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        _o = C.CDLL(ORACLE_LIB)
    return _o


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref() -> C.CDLL:
    global _r
    if _r is None:
        _r = C.CDLL(REF_LIB)
        for f in ("ref_bfs", "ref_sssp", "ref_cc", "ref_pagerank"):
            getattr(_r, f).restype = C.c_int64
    return _r


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def ref_edge_trace(n, s, d, w, source):
    """The reference engine's run() on the edge-scattering SSSP of
    tests/template/edge_trace.hip: (dist, hits, last, supersteps)."""
    s, d, w = _i32(s), _i32(d), _i32(w)
    dist = np.zeros(n, np.int32)
    hits = np.zeros(len(s), np.int32)
    last = np.zeros(len(s), np.int32)
    steps = ref().ref_edge_trace(C.c_int(n), C.c_int(len(s)), _p(s), _p(d), _p(w),
                                 C.c_int(source), _p(dist), _p(hits), _p(last))
    return dist, hits, last, steps


_io = None


def refio_available() -> bool:
    return os.path.exists(REFIO_LIB)


def _refio() -> C.CDLL:
    global _io
    if _io is None:
        _io = C.CDLL(REFIO_LIB)
    return _io


def refio_load(path: str, want_values: bool):
    """The reference's loadGraph (graphio.cpp:441-473) on a well-formed
    file: (n_vertices, srcs, dsts, values or None)."""
    r = _refio()
    nv, ne = C.c_int(), C.c_int()
    ps, pd, pw = (C.POINTER(C.c_int)() for _ in range(3))
    r.refio_load(path.encode(), C.c_int(1 if want_values else 0), C.byref(nv),
                 C.byref(ne), C.byref(ps), C.byref(pd), C.byref(pw))
    n = ne.value

    def take(p):
        if not p:
            return None
        a = np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(np.int32)
        r.refio_free(p)
        return a

    return nv.value, take(ps), take(pd), take(pw)


def refio_write_gr(path: str, n_vertices: int, offsets, dsts, values) -> None:
    """The reference's writeGraph_binaryCSR (graphio.cpp:365-414)."""
    off, d = _i32(offsets), _i32(dsts)
    w = None if values is None else _i32(values)
    _refio().refio_write_gr(path.encode(), C.c_int(n_vertices), C.c_int(len(d)),
                            _p(off), _p(d), _p(w) if w is not None else None)


def refio_write_mtx(path: str, n_vertices: int, srcs, dsts, values) -> None:
    """The reference's writeGraph_mtx (graphio.cpp:416-439)."""
    s, d = _i32(srcs), _i32(dsts)
    w = None if values is None else _i32(values)
    _refio().refio_write_mtx(path.encode(), C.c_int(n_vertices), C.c_int(len(d)),
                             _p(s), _p(d), _p(w) if w is not None else None)


def _p(a, t=C.c_int32):
    return a.ctypes.data_as(C.POINTER(t))


def bfs(nV, srcs, dsts, source):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    st = Stats()
    assert lib().gasref_bfs(nV, len(s), _p(s), _p(d), source, _p(out), C.byref(st)) == 0
    return out, st


def sssp(nV, srcs, dsts, weights, source):
    s, d, w = _i32(srcs), _i32(dsts), _i32(weights)
    out = np.zeros(nV, np.int32)
    st = Stats()
    assert lib().gasref_sssp(nV, len(s), _p(s), _p(d), _p(w), source, _p(out),
                             C.byref(st)) == 0
    return out, st


def cc(nV, srcs, dsts):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    st = Stats()
    assert lib().gasref_cc(nV, len(s), _p(s), _p(d), _p(out), C.byref(st)) == 0
    return out, st


def pagerank(nV, srcs, dsts, iterations):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.float32)
    st = Stats()
    assert lib().gasref_pagerank(nV, len(s), _p(s), _p(d), iterations,
                                 _p(out, C.c_float), C.byref(st)) == 0
    return out, st


def check_bfs(nV, srcs, dsts, source):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    assert lib().gasref_check_bfs(nV, len(s), _p(s), _p(d), source, _p(out)) == 0
    return out


def check_dijkstra(nV, srcs, dsts, weights, source):
    s, d, w = _i32(srcs), _i32(dsts), _i32(weights)
    out = np.zeros(nV, np.int32)
    assert lib().gasref_check_dijkstra(nV, len(s), _p(s), _p(d), _p(w), source,
                                       _p(out)) == 0
    return out


def check_min_reaching(nV, srcs, dsts):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    assert lib().gasref_check_min_reaching(nV, len(s), _p(s), _p(d), _p(out)) == 0
    return out


def check_pagerank_fp64(nV, srcs, dsts, iterations):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.float64)
    assert lib().gasref_check_pagerank_fp64(nV, len(s), _p(s), _p(d), iterations,
                                            _p(out, C.c_double)) == 0
    return out


def check_ppr_fp64(nV, srcs, dsts, source, iterations):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.float64)
    assert lib().gasref_check_ppr_fp64(nV, len(s), _p(s), _p(d), source, iterations,
                                       _p(out, C.c_double)) == 0
    return out


def check_pagerank_tol_fp64(nV, srcs, dsts, tol=0.01):
    """The tolerance-driven PageRank run (pagerank.cu:85-101: every vertex
    active, then refgas.h:215-222's loop -- gather over in-edges, apply
    rank = 0.15 + 0.85 sum, activate the out-neighbours of vertices whose rank
    moved by >= tol, next frontier) evaluated in FP64 with numpy: (ranks,
    supersteps).  Independent of the C restatement; where its activation
    decisions match the FP32 engines' (equal superstep counts), it is the
    exact value their FP32 arithmetic approximates."""
    s = np.asarray(srcs, np.int64)
    d = np.asarray(dsts, np.int64)
    outdeg = np.bincount(s, minlength=nV).astype(np.float64)
    inv = np.zeros(nV)
    inv[outdeg > 0] = 1.0 / outdeg[outdeg > 0]
    rank = np.full(nV, 0.15)
    active = np.arange(nV)
    steps = 0
    while len(active):
        sums = np.bincount(d, weights=(rank * inv)[s], minlength=nV)
        new = 0.15 + 0.85 * sums[active]
        ret = np.abs(new - rank[active]) >= tol
        rank[active] = new
        returned = np.zeros(nV, bool)
        returned[active[ret]] = True
        flags = np.zeros(nV, bool)
        flags[d[returned[s]]] = True
        active = np.nonzero(flags)[0]
        steps += 1
    return rank, steps


def edge_list_to_csc(nV, srcs, dsts):
    """CSC of util.h:114-121: (offsets, sources, edge index) by destination,
    equal destinations in edge-list order."""
    s, d = _i32(srcs), _i32(dsts)
    off = np.zeros(nV + 1, np.int32)
    out = np.zeros(len(s), np.int32)
    eid = np.zeros(len(s), np.int32)
    lib().gasref_edge_list_to_csc(nV, len(s), _p(s), _p(d), _p(off), _p(out), _p(eid))
    return off, out, eid


def cert_bfs(nV, srcs, dsts, source, depth) -> bool:
    """True iff depth is exactly the BFS depth from source (one pass)."""
    s, d, x = _i32(srcs), _i32(dsts), _i32(depth)
    return bool(lib().gasref_cert_bfs(nV, len(s), _p(s), _p(d), source, _p(x)))


def cert_sssp(nV, srcs, dsts, weights, source, dist) -> bool:
    """True iff dist is exactly the shortest-path distance from source
    (unreached 2147383647; positive weights), in one pass."""
    s, d, w, x = _i32(srcs), _i32(dsts), _i32(weights), _i32(dist)
    return bool(lib().gasref_cert_sssp(nV, len(s), _p(s), _p(d), _p(w), source, _p(x)))


def edge_list_to_csr(nV, srcs, dsts):
    s, d = _i32(srcs), _i32(dsts)
    off = np.zeros(nV + 1, np.int32)
    out = np.zeros(len(s), np.int32)
    eid = np.zeros(len(s), np.int32)
    lib().gasref_edge_list_to_csr(nV, len(s), _p(s), _p(d), _p(off), _p(out), _p(eid))
    return off, out, eid


# ---- the reference's own engine (oracle/_ref) ------------------------------
def ref_bfs(nV, srcs, dsts, source):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    us = ref().ref_bfs(nV, len(s), _p(s), _p(d), source, _p(out))
    return out, us


def ref_sssp(nV, srcs, dsts, weights, source):
    s, d, w = _i32(srcs), _i32(dsts), _i32(weights)
    out = np.zeros(nV, np.int32)
    us = ref().ref_sssp(nV, len(s), _p(s), _p(d), _p(w), source, _p(out))
    return out, us


def ref_cc(nV, srcs, dsts):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.int32)
    us = ref().ref_cc(nV, len(s), _p(s), _p(d), _p(out))
    return out, us


def ref_pagerank(nV, srcs, dsts, iterations):
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.float32)
    us = ref().ref_pagerank(nV, len(s), _p(s), _p(d), iterations, _p(out, C.c_float))
    return out, us


def ref_pagerank_run(nV, srcs, dsts):
    """The reference engine's tolerance-driven PageRank (run(), tol 0.01):
    (ranks, supersteps)."""
    s, d = _i32(srcs), _i32(dsts)
    out = np.zeros(nV, np.float32)
    steps = ref().ref_pagerank_run(nV, len(s), _p(s), _p(d), _p(out, C.c_float))
    return out, steps


class RefPrepared:
    """An engine of the reference (compiled refgas.h) bound to a graph once
    and reused query after query: kind 0 BFS, 1 SSSP, 2 personalised
    PageRank (`iterations` sweeps)."""

    def __init__(self, kind, nV, srcs, dsts, weights=None):
        r = ref()
        r.ref_prepare.restype = C.c_void_p
        r.ref_prepared_query.restype = C.c_int64
        r.ref_prepared_query.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        r.ref_prepared_free.argtypes = [C.c_void_p]
        s, d = _i32(srcs), _i32(dsts)
        w = _i32(weights) if weights is not None else None
        self.kind, self.nV = kind, nV
        self.h = r.ref_prepare(kind, nV, len(s), _p(s), _p(d),
                               _p(w) if w is not None else None)

    def query(self, source, iterations=10):
        """(result, microseconds of the reference's Took window)."""
        out = np.zeros(self.nV, np.float32 if self.kind == 2 else np.int32)
        us = ref().ref_prepared_query(self.h, int(source), int(iterations),
                                      out.ctypes.data_as(C.c_void_p))
        return out, us

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_prepared_free(self.h)
            self.h = None


def ref_edge_list_to_csr(nV, srcs, dsts):
    s, d = _i32(srcs), _i32(dsts)
    off = np.zeros(nV + 1, np.int32)
    out = np.zeros(len(s), np.int32)
    eid = np.zeros(len(s), np.int32)
    ref().ref_edge_list_to_csr(nV, len(s), _p(s), _p(d), _p(off), _p(out), _p(eid))
    return off, out, eid
