"""vertexapi2_amd -- Python bindings of libgas.so, the MI355X GAS engine.

The library is the product; this module only binds its C ABI (include/gas.h)
with ctypes and moves numpy arrays across it.  Loading fails loudly when
libgas.so has not been built: there is no fallback implementation.

    import vertexapi2_amd as gas
    el = gas.Edges.road(1024, 1024, seed=1)
    g = gas.Graph(el)
    depth, stats = g.bfs(0)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgas.so")

PROGRAM_BFS, PROGRAM_SSSP, PROGRAM_PAGERANK, PROGRAM_CC = 0, 1, 2, 3
SSSP_UNREACHED = 2147483647 - 100000  # sssp.cu:32-33 gatherZero


class GasError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class EdgesView(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("n_edges", C.c_int32),
                ("srcs", C.POINTER(C.c_int32)), ("dsts", C.POINTER(C.c_int32)),
                ("weights", C.POINTER(C.c_int32))]


class GraphView(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("n_edges", C.c_int32),
                ("csc_offsets", C.c_void_p), ("csc_srcs", C.c_void_p),
                ("csr_offsets", C.c_void_p), ("csr_dsts", C.c_void_p),
                ("csc_weights", C.c_void_p), ("csr_weights", C.c_void_p),
                ("device_bytes", C.c_size_t)]


class RunStats(C.Structure):
    _fields_ = [("supersteps", C.c_int32), ("reserved", C.c_int32),
                ("gathered_edges", C.c_int64), ("scattered_edges", C.c_int64),
                ("active_total", C.c_int64), ("ms", C.c_double),
                ("gather_ms", C.c_double), ("scatter_ms", C.c_double),
                ("gather_launches", C.c_int32), ("scatter_launches", C.c_int32),
                ("kernel_ms", C.c_double * 13), ("kernel_launches", C.c_int32 * 13),
                ("returned_total", C.c_int64), ("hub_vertices_total", C.c_int64),
                ("hub_edges_total", C.c_int64)]

    def as_dict(self) -> dict:
        out = {f: getattr(self, f) for f, _ in self._fields_
               if f not in ("reserved", "kernel_ms", "kernel_launches")}
        out["kernels"] = self.kernels()
        return out

    def kernels(self) -> dict:
        """{kernel name: (summed ms, launches)} of the timed launches."""
        return {name: (self.kernel_ms[i], self.kernel_launches[i])
                for i, name in enumerate(KERNEL_NAMES) if self.kernel_launches[i]}


KERNEL_NAMES = ["k_gather_vp", "k_apply", "k_scatter_vp", "k_hscan1", "k_hscan2",
                "k_hgather", "k_hfix", "k_hscatter", "k_compact1", "k_compact2",
                "k_fused", "k_small", "k_push"]


_lib = None


def lib() -> C.CDLL:
    """The loaded libgas.so (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m vertexapi2_amd.build` (hipcc, gfx950)")
        _lib = C.CDLL(LIB_PATH)
        _lib.gas_last_error.restype = C.c_char_p
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise GasError(rc, lib().gas_last_error().decode())


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a: np.ndarray, ctype=C.c_int32):
    return a.ctypes.data_as(C.POINTER(ctype))


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().gasrt_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def device_info(device: int = 0) -> dict:
    cu, lds, hbm = C.c_int(0), C.c_int(0), C.c_size_t(0)
    name = C.create_string_buffer(256)
    _check(lib().gasrt_device_info(device, C.byref(cu), C.byref(lds), C.byref(hbm),
                                   name, C.c_size_t(256)))
    return {"cu_count": cu.value, "lds_per_block": lds.value,
            "hbm_bytes": hbm.value, "name": name.value.decode()}


def set_kernel_timing(enable: bool) -> None:
    _check(lib().gas_set_kernel_timing(1 if enable else 0))


def trace_mark(tag: int) -> None:
    """Empty kernel of `tag` wave64 workgroups on the default stream, waited
    for: a cut point in profiler kernel traces (grid size 64 * tag)."""
    _check(lib().gas_trace_mark(C.c_uint32(tag)))


def set_device_build(min_edges: int) -> None:
    """Graphs of at least `min_edges` edges get their CSC / CSR laid out on
    the GPU (default 2**20; 0: every graph)."""
    _check(lib().gas_graph_set_device_build(C.c_int64(min_edges)))


def download_i32(ptr: int, n: int) -> np.ndarray:
    """n int32 values from device memory (synchronous)."""
    out = np.zeros(max(n, 0), np.int32)
    if n > 0:
        _check(lib().gasrt_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                      C.c_size_t(4 * n), None))
    return out


class DeviceBuffer:
    """Device memory from gasrt_malloc (freed with the object)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().gasrt_malloc(C.byref(p), C.c_size_t(self.nbytes)))
        self.ptr = p.value

    def to_numpy(self, dtype=np.int32) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        _check(lib().gasrt_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                      C.c_size_t(out.nbytes), None))
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gasrt_free(C.c_void_p(self.ptr))
            self.ptr = None


def device_copy(dst: DeviceBuffer, src: DeviceBuffer) -> None:
    """Synchronous device-to-device copy of min(dst, src) bytes."""
    n = min(dst.nbytes, src.nbytes)
    _check(lib().gasrt_memcpy_d2d(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(n),
                                  None))
    _check(lib().gasrt_stream_synchronize(None))


def _want(want_values) -> int:
    """Loader value flag: True / False, or -1 for whatever the file holds."""
    return -1 if want_values == -1 else (1 if want_values else 0)


class Edges:
    """Host edge list (gas_edges_t)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gas_edges_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_arrays(cls, n_vertices: int, srcs, dsts, weights=None) -> "Edges":
        s, d = _i32(srcs), _i32(dsts)
        w = None if weights is None else _i32(weights)
        h = C.c_void_p()
        _check(lib().gas_edges_from_arrays(
            C.c_int32(n_vertices), C.c_int32(len(s)), _ptr(s), _ptr(d),
            _ptr(w) if w is not None else None, C.byref(h)))
        return cls(h)

    @classmethod
    def grid(cls, rows: int, cols: int) -> "Edges":
        h = C.c_void_p()
        _check(lib().gas_gen_grid(C.c_int32(rows), C.c_int32(cols), C.byref(h)))
        return cls(h)

    @classmethod
    def road(cls, rows: int, cols: int, drop_frac: float = 0.03,
             long_frac: float = 0.01, wmin: int = 1, wmax: int = 64,
             seed: int = 1) -> "Edges":
        h = C.c_void_p()
        _check(lib().gas_gen_road(C.c_int32(rows), C.c_int32(cols),
                                  C.c_double(drop_frac), C.c_double(long_frac),
                                  C.c_int32(wmin), C.c_int32(wmax),
                                  C.c_uint64(seed), C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57,
             b: float = 0.19, c: float = 0.19, seed: int = 1,
             symmetrize: bool = False, wmin: int = 1, wmax: int = 0) -> "Edges":
        h = C.c_void_p()
        _check(lib().gas_gen_rmat(C.c_int32(scale), C.c_int32(edge_factor),
                                  C.c_double(a), C.c_double(b), C.c_double(c),
                                  C.c_uint64(seed), C.c_int(1 if symmetrize else 0),
                                  C.c_int32(wmin), C.c_int32(wmax), C.byref(h)))
        return cls(h)

    @classmethod
    def load_mtx(cls, path: str, want_values: bool = False) -> "Edges":
        h = C.c_void_p()
        _check(lib().gas_load_mtx(path.encode(), C.c_int(_want(want_values)),
                                  C.byref(h)))
        return cls(h)

    @classmethod
    def load_edge(cls, path: str) -> "Edges":
        """GraphLab SNAP edge list (0-based "src dst" lines, '#' comments)."""
        h = C.c_void_p()
        _check(lib().gas_load_edge(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def load_gr(cls, path: str, want_values: bool = False) -> "Edges":
        """Lonestar binary CSR (.gr), expanded to an edge list in CSR order."""
        h = C.c_void_p()
        _check(lib().gas_load_gr(path.encode(), C.c_int(_want(want_values)),
                                 C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path: str, want_values: bool = False) -> "Edges":
        """By extension (.edge, .mtx, .gr, optionally .gz), as loadGraph
        (graphio.cpp:441-473)."""
        h = C.c_void_p()
        _check(lib().gas_load_graph(path.encode(), C.c_int(_want(want_values)),
                                    C.byref(h)))
        return cls(h)

    def write_gr(self, path: str) -> None:
        _check(lib().gas_write_gr(path.encode(), self._h))

    def write_mtx(self, path: str) -> None:
        _check(lib().gas_write_mtx(path.encode(), self._h))

    def view(self) -> EdgesView:
        v = EdgesView()
        _check(lib().gas_edges_view_get(self._h, C.byref(v)))
        return v

    @property
    def n_vertices(self) -> int:
        return self.view().n_vertices

    @property
    def n_edges(self) -> int:
        return self.view().n_edges

    def arrays(self):
        """(srcs, dsts, weights or None) as numpy copies."""
        v = self.view()
        n = v.n_edges
        if n == 0:
            e = np.zeros(0, np.int32)
            return e, e.copy(), (None if not v.weights else e.copy())
        s = np.ctypeslib.as_array(v.srcs, shape=(n,)).copy()
        d = np.ctypeslib.as_array(v.dsts, shape=(n,)).copy()
        w = np.ctypeslib.as_array(v.weights, shape=(n,)).copy() if v.weights else None
        return s, d, w

    def max_outdegree_vertex(self):
        v, d = C.c_int32(0), C.c_int32(0)
        _check(lib().gas_edges_max_outdegree_vertex(self._h, C.byref(v), C.byref(d)))
        return v.value, d.value


class Graph:
    """A graph resident in HBM (gas_graph_t): CSC + CSR + CSC-order weights."""

    def __init__(self, edges: Edges):
        h = C.c_void_p()
        _check(lib().gas_graph_create(edges.handle, C.byref(h)))
        self._h = h
        v = self.view()
        self.n_vertices, self.n_edges = v.n_vertices, v.n_edges
        self.has_weights = bool(v.csc_weights)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gas_graph_free(self._h)
            self._h = None

    def free(self) -> None:
        self.__del__()

    @property
    def handle(self):
        return self._h

    def view(self) -> GraphView:
        v = GraphView()
        _check(lib().gas_graph_view_get(self._h, C.byref(v)))
        return v

    def host_csc(self):
        off = np.zeros(self.n_vertices + 1, np.int32)
        src = np.zeros(self.n_edges, np.int32)
        eid = np.zeros(self.n_edges, np.int32)
        _check(lib().gas_graph_host_csc(self._h, _ptr(off), _ptr(src), _ptr(eid)))
        return off, src, eid

    def host_csr(self):
        off = np.zeros(self.n_vertices + 1, np.int32)
        dst = np.zeros(self.n_edges, np.int32)
        eid = np.zeros(self.n_edges, np.int32)
        _check(lib().gas_graph_host_csr(self._h, _ptr(off), _ptr(dst), _ptr(eid)))
        return off, dst, eid

    def set_tiers(self, program: int, tiers) -> None:
        t = _i32(tiers)
        assert t.shape == (4,)
        _check(lib().gas_graph_set_tiers(self._h, C.c_int(program), _ptr(t)))

    def get_tiers(self, program: int):
        t = np.zeros(4, np.int32)
        _check(lib().gas_graph_get_tiers(self._h, C.c_int(program), _ptr(t)))
        return t.tolist()

    def set_frontier_switch(self, program: int, divisor: int) -> None:
        """List frontier while |A| * divisor < V, bitmap otherwise."""
        _check(lib().gas_graph_set_frontier_switch(self._h, C.c_int(program),
                                                   C.c_int32(divisor)))

    def autotune_batch(self, kind: int) -> int:
        """Times a 64-lane batch of `kind` (BATCH_BFS: pull divisor,
        BATCH_SSSP: distance step) per candidate and keeps the fastest for
        the batch engines created on this graph afterwards; returns it."""
        p = C.c_int32(0)
        _check(lib().gas_graph_autotune_batch(self._h, C.c_int(kind), C.byref(p)))
        return p.value

    def batch_param(self, kind: int) -> int:
        p = C.c_int32(0)
        _check(lib().gas_graph_batch_param(self._h, C.c_int(kind), C.byref(p)))
        return p.value

    def batch_hub_items(self, kind: int) -> int:
        """Hub threshold (push items) the batch engines on this graph start
        with: the tuned value, 32 until tuned."""
        p = C.c_int32(0)
        _check(lib().gas_graph_batch_hub_items(self._h, C.c_int(kind), C.byref(p)))
        return p.value

    def batch_item_edges(self, kind: int) -> int:
        """Out-edges per push item of the batch engines on this graph (the
        tuned value; BFS 16, SSSP 64 until tuned)."""
        p = C.c_int32(0)
        _check(lib().gas_graph_batch_item_edges(self._h, C.c_int(kind), C.byref(p)))
        return p.value

    def batch_graph_rounds(self, kind: int) -> int:
        """Rounds per replay of the captured round graph the batch engines
        on this graph start with (0: direct launches, the default)."""
        p = C.c_int32(0)
        _check(lib().gas_graph_batch_graph_rounds(self._h, C.c_int(kind), C.byref(p)))
        return p.value

    def set_batch_graph_rounds(self, kind: int, rounds: int) -> None:
        _check(lib().gas_graph_set_batch_graph_rounds(self._h, C.c_int(kind),
                                                      C.c_int32(rounds)))

    def set_push(self, program: int, enable: bool) -> None:
        """Push supersteps of whole SSSP / CC runs (same results and
        superstep counts as the pull)."""
        _check(lib().gas_graph_set_push(self._h, C.c_int(program), C.c_int32(1 if enable else 0)))

    def set_small_frontier(self, program: int, max_vertices: int) -> None:
        """Runs supersteps of list frontiers of at most `max_vertices` inside
        one workgroup on graphs of degree <= 64 (0: off)."""
        _check(lib().gas_graph_set_small_frontier(self._h, C.c_int(program),
                                                  C.c_int32(max_vertices)))

    def autotune(self, program: int):
        """Times candidate degree tiers for `program` on this graph and keeps
        the fastest; returns [gather t1, gather t2, scatter t1, scatter t2]."""
        t = np.zeros(4, np.int32)
        _check(lib().gas_graph_autotune(self._h, C.c_int(program), _ptr(t)))
        return t.tolist()

    def bfs(self, source: int):
        out = np.zeros(self.n_vertices, np.int32)
        st = RunStats()
        _check(lib().gas_bfs(self._h, C.c_int32(source), _ptr(out), C.byref(st)))
        return out, st

    def sssp(self, source: int):
        out = np.zeros(self.n_vertices, np.int32)
        st = RunStats()
        _check(lib().gas_sssp(self._h, C.c_int32(source), _ptr(out), C.byref(st)))
        return out, st

    def bfs_device(self, source: int, out_ptr: int):
        """BFS with the depths written to device memory at out_ptr
        (n_vertices int32); returns the run statistics."""
        st = RunStats()
        _check(lib().gas_bfs_device(self._h, C.c_int32(source), C.c_void_p(out_ptr),
                                    C.byref(st)))
        return st

    def sssp_device(self, source: int, out_ptr: int):
        """SSSP with the distances written to device memory at out_ptr."""
        st = RunStats()
        _check(lib().gas_sssp_device(self._h, C.c_int32(source), C.c_void_p(out_ptr),
                                     C.byref(st)))
        return st

    def cc_device(self, out_ptr: int):
        """Connected components with the labels written to device memory."""
        st = RunStats()
        _check(lib().gas_cc_device(self._h, C.c_void_p(out_ptr), C.byref(st)))
        return st

    def set_vertex_order(self, mode: int) -> None:
        """1 (default): PageRank and CC run on a copy renumbered by
        descending in-degree (skewed graphs of 2**20+ edges); 0: on the graph as
        given."""
        _check(lib().gas_graph_set_vertex_order(self._h, C.c_int32(mode)))

    def pagerank_device(self, out_ptr: int, iterations: int = 20, residual_ptr=None):
        """Fixed-sweep PageRank with the ranks (and, if given, the per-sweep
        FP64 residuals) written to device memory."""
        st = RunStats()
        _check(lib().gas_pagerank_device(self._h, C.c_int32(iterations),
                                         C.c_void_p(out_ptr),
                                         C.c_void_p(residual_ptr) if residual_ptr else None,
                                         C.byref(st)))
        return st

    def pagerank_device_noalloc(self, out_ptr: int, iterations: int = 20,
                                residual_ptr=None, stream=None) -> None:
        """gas_pagerank_device_noalloc: the fixed-sweep run with every buffer
        in the library's fixed device scratch (no allocation)."""
        _check(lib().gas_pagerank_device_noalloc(
            self._h, C.c_int32(iterations), C.c_void_p(out_ptr),
            C.c_void_p(residual_ptr) if residual_ptr else None,
            C.c_void_p(stream) if stream else None))

    def ppr_device(self, source: int, out_ptr: int, iterations: int = 10):
        """Personalised PageRank with the ranks written to device memory."""
        st = RunStats()
        _check(lib().gas_ppr_device(self._h, C.c_int32(source), C.c_int32(iterations),
                                    C.c_void_p(out_ptr), C.byref(st)))
        return st

    def cc(self):
        out = np.zeros(self.n_vertices, np.int32)
        st = RunStats()
        _check(lib().gas_cc(self._h, _ptr(out), C.byref(st)))
        return out, st

    def ppr(self, source: int, iterations: int = 10):
        out = np.zeros(self.n_vertices, np.float32)
        st = RunStats()
        _check(lib().gas_ppr(self._h, C.c_int32(source), C.c_int32(iterations),
                             _ptr(out, C.c_float), C.byref(st)))
        return out, st

    def pagerank(self, iterations: int = 20):
        out = np.zeros(self.n_vertices, np.float32)
        res = np.zeros(max(iterations, 1), np.float64)
        st = RunStats()
        _check(lib().gas_pagerank(self._h, C.c_int32(iterations), _ptr(out, C.c_float),
                                  _ptr(res, C.c_double), C.byref(st)))
        return out, (res[:iterations] if iterations > 0 else None), st


# VertexData record layouts of the four programs (include/gas/programs.hpp).
VERTEX_DTYPES = {
    PROGRAM_BFS: np.dtype([("depth", np.int32)]),
    PROGRAM_SSSP: np.dtype(np.int32),
    PROGRAM_PAGERANK: np.dtype([("rank", np.float32), ("numOutEdges", np.int32)]),
    PROGRAM_CC: np.dtype(np.int32),
}


class Engine:
    """The reference engine API step by step (gpugas.h:212-772).

    `vertex_data` is a numpy array of the program's VertexData records; it
    is retained and written back by get_results(), as in the reference.
    """

    def __init__(self, program: int, n_vertices: int, vertex_data: np.ndarray,
                 srcs, dsts, edge_data=None):
        assert vertex_data.dtype == VERTEX_DTYPES[program]
        assert vertex_data.flags["C_CONTIGUOUS"] and len(vertex_data) == n_vertices
        self.vertex_data = vertex_data
        s, d = _i32(srcs), _i32(dsts)
        self._ed = None if edge_data is None else _i32(edge_data)
        h = C.c_void_p()
        _check(lib().gas_engine_create(
            C.c_int(program), C.c_int32(n_vertices), vertex_data.ctypes.data_as(C.c_void_p),
            C.c_int32(len(s)),
            self._ed.ctypes.data_as(C.c_void_p) if self._ed is not None else None,
            _ptr(s), _ptr(d), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gas_engine_destroy(self._h)
            self._h = None

    def setActive(self, start: int, end: int) -> None:
        _check(lib().gas_engine_set_active(self._h, C.c_int32(start), C.c_int32(end)))

    def countActive(self) -> int:
        n = C.c_int32(0)
        _check(lib().gas_engine_count_active(self._h, C.byref(n)))
        return n.value

    def gatherApply(self, haveGather: bool = True) -> None:
        _check(lib().gas_engine_gather_apply(self._h, C.c_int(1 if haveGather else 0)))

    def scatterActivate(self, haveScatter: bool = True) -> None:
        _check(lib().gas_engine_scatter_activate(self._h, C.c_int(1 if haveScatter else 0)))

    def nextIter(self) -> int:
        n = C.c_int32(0)
        _check(lib().gas_engine_next_iter(self._h, C.byref(n)))
        return n.value

    def run(self) -> None:
        _check(lib().gas_engine_run(self._h))

    def getResults(self) -> None:
        _check(lib().gas_engine_get_results(self._h))

    def setIteration(self, it: int) -> None:
        _check(lib().gas_engine_set_iteration(self._h, C.c_int32(it)))

    def stats(self) -> RunStats:
        st = RunStats()
        _check(lib().gas_engine_stats(self._h, C.byref(st)))
        return st


BATCH_BFS, BATCH_SSSP, BATCH_PPR = 0, 1, 2


class Batch:
    """Up to 64 BFS or SSSP queries sharing edge reads (gas_batch_t)."""

    def __init__(self, graph: Graph, kind: int, ppr_iterations: int = 10):
        h = C.c_void_p()
        if kind == BATCH_PPR:
            _check(lib().gas_batch_create_ppr(graph.handle, C.c_int32(ppr_iterations),
                                              C.byref(h)))
        else:
            _check(lib().gas_batch_create(graph.handle, C.c_int(kind), C.byref(h)))
        self._h = h
        self.graph = graph  # keeps the resident graph alive
        self.kind = kind

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gas_batch_free(self._h)
            self._h = None

    def add(self, lanes, sources) -> None:
        la, so = _i32(lanes), _i32(sources)
        _check(lib().gas_batch_add(self._h, _ptr(la), _ptr(so), C.c_int32(len(la))))

    def step(self, rounds: int = 1) -> None:
        _check(lib().gas_batch_step(self._h, C.c_int32(rounds)))

    def alive(self) -> int:
        x = C.c_uint64(0)
        _check(lib().gas_batch_alive(self._h, C.byref(x)))
        return x.value

    def used(self) -> int:
        x = C.c_uint64(0)
        _check(lib().gas_batch_used(self._h, C.byref(x)))
        return x.value

    def release(self, lanes_mask: int) -> None:
        _check(lib().gas_batch_release(self._h, C.c_uint64(lanes_mask)))

    def run(self) -> RunStats:
        st = RunStats()
        _check(lib().gas_batch_run(self._h, C.byref(st)))
        return st

    def results(self, lanes) -> dict:
        """{lane: result} for several lanes, extracted in one device pass."""
        lanes = [int(s) for s in lanes]  # numpy ints would overflow 1 << s
        if any(s < 0 or s >= 64 for s in lanes):
            raise GasError(1, "lanes are 0..63")
        bufs = {s: DeviceBuffer(4 * self.graph.n_vertices) for s in lanes}
        ptrs = (C.c_void_p * 64)()
        mask = 0
        for s, b in bufs.items():
            ptrs[s] = b.ptr
            mask |= 1 << s
        if self.kind == BATCH_PPR:
            _check(lib().gas_batch_results_f32_device(self._h, C.c_uint64(mask), ptrs))
            return {s: b.to_numpy(np.float32) for s, b in bufs.items()}
        _check(lib().gas_batch_results_device(self._h, C.c_uint64(mask), ptrs))
        return {s: b.to_numpy() for s, b in bufs.items()}

    def result(self, lane: int) -> np.ndarray:
        if self.kind == BATCH_PPR:
            out = np.zeros(self.graph.n_vertices, np.float32)
            _check(lib().gas_batch_result_f32(self._h, C.c_int32(lane),
                                              _ptr(out, C.c_float)))
            return out
        out = np.zeros(self.graph.n_vertices, np.int32)
        _check(lib().gas_batch_result(self._h, C.c_int32(lane), _ptr(out)))
        return out

    def set_pull_divisor(self, divisor: int) -> None:
        _check(lib().gas_batch_set_pull_divisor(self._h, C.c_int32(divisor)))

    def set_hub_items(self, items: int) -> None:
        """List vertices with more push items than this have their items
        written by a whole workgroup (>= 16)."""
        _check(lib().gas_batch_set_hub_items(self._h, C.c_int32(items)))

    def set_item_edges(self, edges: int) -> None:
        """BFS / SSSP: out-edges per push item (BFS 8 or 16, SSSP 32 or 64)."""
        _check(lib().gas_batch_set_item_edges(self._h, C.c_int32(edges)))

    def set_graph_rounds(self, rounds: int) -> None:
        """BFS / SSSP: issue rounds as replays of a captured graph of
        `rounds` rounds (1..8); 0 launches each round directly."""
        _check(lib().gas_batch_set_graph_rounds(self._h, C.c_int32(rounds)))

    def set_sssp_delta(self, delta: int) -> None:
        """SSSP lanes relax a vertex only while its distance is below
        (rounds since the lane began) * delta (0: no bound)."""
        _check(lib().gas_batch_set_sssp_delta(self._h, C.c_int32(delta)))


class ServeStats(C.Structure):
    _fields_ = [("elapsed_s", C.c_double), ("rounds", C.c_int64),
                ("admissions_deferred", C.c_int64), ("max_concurrent", C.c_int32),
                ("solo", C.c_int32), ("evictions", C.c_int64), ("device_bytes", C.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def query_bytes(kind: int, n_vertices: int) -> int:
    f = lib().gas_server_query_bytes
    f.restype = C.c_size_t
    return int(f(C.c_int(kind), C.c_int64(n_vertices)))


class Server:
    """Continuous-batching server over resident graphs (gas_server_t)."""

    def __init__(self, graphs, memory_budget: int, ppr_iterations: int = 10):
        self.graphs = list(graphs)
        arr = (C.c_void_p * len(self.graphs))(*[g.handle.value for g in self.graphs])
        h = C.c_void_p()
        _check(lib().gas_server_create(arr, C.c_int32(len(self.graphs)),
                                       C.c_size_t(memory_budget),
                                       C.c_int32(ppr_iterations), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().gas_server_free(self._h)
            self._h = None

    def set_solo(self, max_ms: float, idle_max_ms: float) -> None:
        """Plain-engine cost caps for answering a query alone (0, 0: never)."""
        _check(lib().gas_server_set_solo(self._h, C.c_double(max_ms), C.c_double(idle_max_ms)))

    def set_gang(self, backlog: int, rounds: int) -> None:
        """While `backlog` queries wait for a BFS/SSSP engine, they join its
        batch only within `rounds` rounds of its start (0: always join)."""
        _check(lib().gas_server_set_gang(self._h, C.c_int32(backlog), C.c_int32(rounds)))

    def budget(self):
        """(budget, device bytes the server holds now)."""
        b, h = C.c_size_t(), C.c_size_t()
        _check(lib().gas_server_budget(self._h, C.byref(b), C.byref(h)))
        return b.value, h.value

    def set_budget(self, nbytes: int) -> None:
        _check(lib().gas_server_set_budget(self._h, C.c_size_t(nbytes)))

    def engine_bytes(self, kind: int, graph_id: int) -> int:
        b = C.c_size_t()
        _check(lib().gas_server_engine_bytes(self._h, C.c_int32(kind), C.c_int32(graph_id),
                                             C.byref(b)))
        return b.value

    # ---- online serving ---------------------------------------------------
    def start(self) -> None:
        """Runs the serving loop on the server's own thread (submit / poll /
        wait / stop)."""
        _check(lib().gas_server_start(self._h))

    def submit(self, kind: int, graph_id: int, source: int, out: np.ndarray = None) -> int:
        """Hands one query to the running server; returns its ticket.  `out`
        (V int32 for BFS / SSSP, V float32 for PPR, kept alive by the
        caller until the ticket is done) receives the result."""
        if out is not None:
            want = np.float32 if kind == BATCH_PPR else np.int32
            n = self.graphs[graph_id].n_vertices
            if out.dtype != want or out.size != n or not out.flags["C_CONTIGUOUS"]:
                raise ValueError("result array must be contiguous, %d x %s" % (n, want))
        t = C.c_int64()
        _check(lib().gas_server_submit(
            self._h, C.c_int32(kind), C.c_int32(graph_id), C.c_int32(source),
            C.c_void_p(out.ctypes.data) if out is not None else None, C.byref(t)))
        return t.value

    def poll(self, ticket: int):
        """(state, latency_ms): state 0 waiting / running, 1 done."""
        st, lat = C.c_int32(), C.c_double()
        _check(lib().gas_server_poll(self._h, C.c_int64(ticket), C.byref(st), C.byref(lat)))
        return st.value, lat.value

    def wait(self, ticket: int, timeout_s: float = 60.0) -> float:
        """Blocks until the ticket is done; returns its latency in ms."""
        lat = C.c_double()
        _check(lib().gas_server_wait(self._h, C.c_int64(ticket), C.c_double(timeout_s),
                                     C.byref(lat)))
        return lat.value

    def stop(self) -> "ServeStats":
        st = ServeStats()
        _check(lib().gas_server_stop(self._h, C.byref(st)))
        return st

    def replay(self, kinds, graph_ids, sources, arrivals_s, keep=None):
        """Serves the stream in real time; returns (latency_ms, results, stats)
        where results maps query index -> array for the indices in `keep`."""
        k, gi, so = _i32(kinds), _i32(graph_ids), _i32(sources)
        ar = np.ascontiguousarray(arrivals_s, dtype=np.float64)
        n = len(k)
        lat = np.zeros(n, np.float64)
        results = {}
        ptrs = None
        if keep:
            ptrs = (C.c_void_p * n)()
            for q in keep:
                dt = np.float32 if k[q] == BATCH_PPR else np.int32
                results[q] = np.zeros(self.graphs[gi[q]].n_vertices, dt)
                ptrs[q] = results[q].ctypes.data
        st = ServeStats()
        _check(lib().gas_server_replay(self._h, C.c_int32(n), _ptr(k), _ptr(gi), _ptr(so),
                                       ar.ctypes.data_as(C.POINTER(C.c_double)), ptrs,
                                       lat.ctypes.data_as(C.POINTER(C.c_double)),
                                       C.byref(st)))
        return lat, results, st
