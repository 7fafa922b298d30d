"""The seeded synthetic inputs against their specification (host code only,
no GPU): grid shape, road link counts and symmetric weights, R-MAT
dedupe / symmetrisation and reproducibility."""
import numpy as np
import pytest

from vertexapi2_amd import workloads


def _pairs(s, d):
    return s.astype(np.int64) * (1 << 32) + d.astype(np.int64)


def test_grid64_shape():
    el = workloads.grid64()
    s, d, w = el.arrays()
    assert el.n_vertices == 4096 and len(s) == 16128
    r, c = s // 64, s % 64
    r2, c2 = d // 64, d % 64
    assert np.all(np.abs(r - r2) + np.abs(c - c2) == 1)  # 4-neighbour moves only
    assert len(np.unique(_pairs(s, d))) == len(s)


@pytest.mark.parametrize("rows,cols", [(64, 64), (128, 96)])
def test_road_counts_and_symmetry(rows, cols):
    el = workloads.road(rows, cols, seed=3)
    s, d, w = el.arrays()
    n = rows * cols
    assert el.n_vertices == n
    links = rows * (cols - 1) + cols * (rows - 1)
    # drop floor(3%) of the grid links, add floor(1%) long links, both directions
    assert len(s) == 2 * (links - int(0.03 * links) + int(0.01 * links))
    assert np.all(s != d)
    key = _pairs(s, d)
    assert len(np.unique(key)) == len(s)
    # every edge has its reverse with the same weight
    order = np.argsort(key)
    rev = np.searchsorted(key[order], _pairs(d, s))
    assert np.all(key[order][rev] == _pairs(d, s))
    assert np.array_equal(w[order][rev], w)
    assert w.min() >= 1 and w.max() <= 64


def test_road_reproducible_and_seeded():
    a = workloads.road(48, 48, seed=5).arrays()
    b = workloads.road(48, 48, seed=5).arrays()
    c = workloads.road(48, 48, seed=6).arrays()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not (len(a[0]) == len(c[0]) and np.array_equal(a[0], c[0])
                and np.array_equal(a[1], c[1]))


@pytest.mark.parametrize("scale", [10, 12])
def test_rmat_dedupe_and_symmetrise(scale):
    el = workloads.rmat(scale)
    s, d, w = el.arrays()
    assert el.n_vertices == 1 << scale
    assert w is None
    assert len(s) <= 16 << scale
    assert np.all(s != d)
    assert len(np.unique(_pairs(s, d))) == len(s)
    # R-MAT with a = .57 favours low ids: vertex 0's degree is far above mean
    deg = np.bincount(s, minlength=el.n_vertices)
    assert deg[0] > 10 * deg.mean()
    sym = workloads.rmat(scale, symmetrize=True, weights=True)
    s2, d2, w2 = sym.arrays()
    k2 = _pairs(s2, d2)
    assert len(np.unique(k2)) == len(s2)
    order = np.argsort(k2)
    rev = np.searchsorted(k2[order], _pairs(d2, s2))
    assert np.all(k2[order][rev] == _pairs(d2, s2))
    assert np.array_equal(w2[order][rev], w2)  # one weight per undirected pair
    assert w2.min() >= 1 and w2.max() <= 255
    # the symmetrised graph holds every directed edge of the original
    assert np.all(np.isin(_pairs(s, d), k2))

