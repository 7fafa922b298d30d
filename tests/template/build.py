"""Builds the template-API test programs (gfx950) into tests/template/_lib.

    python tests/template/build.py

Each .hip file here includes the engine's template header and links
libgas.so (found next to the package through an $ORIGIN-relative rpath), as
a program of a reference user would.  Test infrastructure only.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from vertexapi2_amd import build as gasbuild  # noqa: E402

LIBDIR = os.path.join(HERE, "_lib")
PROGRAMS = ["edge_trace"]


def lib_path(name: str) -> str:
    return os.path.join(LIBDIR, f"lib{name}.so")


def build() -> None:
    os.makedirs(LIBDIR, exist_ok=True)
    for name in PROGRAMS:
        out = lib_path(name)
        cmd = [gasbuild.hipcc(), *gasbuild.COMMON, *gasbuild.HIP_FLAGS, "-shared",
               os.path.join(HERE, name + ".hip"), "-o", out + ".tmp",
               "-L", gasbuild.LIBDIR, "-lgas",
               "-Wl,-rpath,$ORIGIN/../../../vertexapi2_amd/_lib"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build()
    print("built", ", ".join(lib_path(n) for n in PROGRAMS))
