"""Graph file converters (mtx2gr.cpp, gr2mtx.cpp).

    python -m vertexapi2_amd.convert input output

Reads any format `Edges.load` reads (.edge / .mtx / .gr, optionally .gz)
with the values the file holds, and writes Lonestar binary CSR when the
output ends in ".gr", MatrixMarket when it ends in ".mtx".  As in
mtx2gr.cpp:41-51 the .gr output holds each source's edges in input order.
"""
from __future__ import annotations

import sys

from . import Edges


def convert(src: str, dst: str) -> Edges:
    el = Edges.load(src, want_values=-1)
    if dst.endswith(".gr"):
        el.write_gr(dst)
    elif dst.endswith(".mtx"):
        el.write_mtx(dst)
    else:
        raise ValueError(f"unrecognized output extension: {dst} (use .gr or .mtx)")
    return el


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("Usage: python -m vertexapi2_amd.convert input output", file=sys.stderr)
        return 1
    el = convert(argv[0], argv[1])
    v = el.view()
    print(f"Read input file with {v.n_vertices} vertices and {v.n_edges} edges")
    return 0


if __name__ == "__main__":
    sys.exit(main())
