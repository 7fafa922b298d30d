"""The C ABI: libgas.so loads on a machine without a GPU and exports every
function include/gas.h declares; calls that need no device behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gas.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    # return type, name, '(' at the start of a declaration
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*"
                       r"((?:gas|gasrt)_[A-Za-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("gas_bfs", "gas_sssp", "gas_cc", "gas_pagerank", "gas_ppr",
                 "gas_engine_create", "gas_engine_run", "gas_batch_create",
                 "gas_server_replay", "gasrt_malloc", "gas_last_error"):
        assert must in names
    assert len(names) > 80


def test_library_exports_every_declared_symbol():
    import vertexapi2_amd as gas
    lib = gas.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_calls_without_gpu():
    import vertexapi2_amd as gas
    el = gas.Edges.grid(4, 4)
    s, d, w = el.arrays()
    assert el.n_vertices == 16 and len(s) == 48
    # errors come back as status codes with a message
    lib = gas.lib()
    assert lib.gas_bfs(None, 0, None, None) == 1
    assert b"gas_bfs" in lib.gas_last_error()
    f = lib.gas_server_query_bytes
    f.restype = C.c_size_t
    assert f(C.c_int(0), C.c_int64(1 << 20)) == gas.query_bytes(0, 1 << 20) > 0
    assert np.all(np.asarray(w) == 1)


@pytest.mark.parametrize("name", ["small.mtx", "trivialGraph.mtx", "chesapeake.mtx"])
def test_mtx_loader(name):
    import vertexapi2_amd as gas
    el = gas.Edges.load_mtx(os.path.join(ROOT, "tests", "golden", name))
    assert el.n_vertices > 0 and el.n_edges > 0
