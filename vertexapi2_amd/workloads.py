"""The synthetic workloads the engine is measured on (seeded, reproducible).

    grid64      64x64 4-neighbour grid, 4,096 vertices, 16,128 directed edges,
                unit weights; BFS from vertex 0
    road1024    1024x1024 grid, 3% of links dropped, 1% random long links,
                symmetric int32 weights U[1,64]; BFS and SSSP from vertex 0
    rmat21      R-MAT scale 21, edge factor 16, a/b/c/d .57/.19/.19/.05,
                self loops and duplicates removed (directed, PageRank), and
                its symmetrised form (CC)
    rmat22s     R-MAT scale 22, symmetrised, weights U[1,255]; 64-source
                batched BFS/SSSP
"""
from __future__ import annotations

from . import Edges

SEED_ROAD = 20240601
SEED_RMAT21 = 20240621
SEED_RMAT22 = 20240622
SEED_SOURCES = 20240664


def grid64() -> Edges:
    return Edges.grid(64, 64)


def road(rows: int = 1024, cols: int = 1024, seed: int = SEED_ROAD) -> Edges:
    return Edges.road(rows, cols, drop_frac=0.03, long_frac=0.01, wmin=1, wmax=64,
                      seed=seed)


def rmat(scale: int = 21, symmetrize: bool = False, weights: bool = False,
         seed: int | None = None) -> Edges:
    if seed is None:
        seed = SEED_RMAT21 if scale == 21 else SEED_RMAT22 if scale == 22 else 7 + scale
    return Edges.rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed, symmetrize=symmetrize,
                      wmin=1, wmax=255 if weights else 0)
