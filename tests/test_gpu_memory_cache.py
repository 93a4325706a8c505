"""GPU tests of the device memory cache: blocks released by destroyed handles are reused by later handles,
results do not depend on what a reused block held before, and egt_release_cached_memory hands every
cached block back.
"""
import numpy as np
import pytest

from multigpuegtsmoothing_amd import datasets as D

pytestmark = pytest.mark.gpu


def _run(E, pts, k):
    g = E.GridStructure(pts, 1, 1e-6, 1024, 4)
    g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
    rows = g.GetKNNIndexes()
    E.EGT_CUDA(g).PerformSmoothing(20, 0.5, -0.53, 0.65)
    nrm = np.stack(E.Normal_CUDA(g).GetNormals(), 1)
    out = np.stack(E.getOutputPoints(g), 1)
    g.close()
    return rows, out, nrm


def test_reused_blocks_give_identical_results(egt_mod):
    E = egt_mod
    E.release_cached_memory()
    a = D.sphere(30_000, seed=71)
    b = D.torus(45_000, seed=72)  # a different cloud dirties the cached blocks in between
    first = _run(E, a, 40)
    _, cached = E.release_cached_memory()
    print("cached after one handle", cached, flush=True)
    assert cached > 0
    first = _run(E, a, 40)
    _run(E, b, 16)
    again = _run(E, a, 40)
    for x, y in zip(first, again):
        np.testing.assert_array_equal(x, y)
    released, cached = E.release_cached_memory()
    assert released == cached and cached > 0
    assert E.release_cached_memory() == (0, 0)


def test_batch_after_single_cloud_reuse(egt_mod):
    E = egt_mod
    clouds = [D.sphere(8_000, seed=81), D.torus(12_000, seed=82)]
    _run(E, D.torus(40_000, seed=83), 40)
    res = []
    for _ in range(2):
        bg = E.BatchGrid(clouds)
        bg.knn([16, 40])
        E.EGT_CUDA(bg).PerformSmoothing(10, 0.5, -0.53, 0.65)
        res.append((bg.GetKNNIndexes(0), bg.GetKNNIndexes(1), np.stack(E.getOutputPoints(bg), 1)))
        bg.close()
        _run(E, D.sphere(20_000, seed=84), 64)
    for x, y in zip(res[0], res[1]):
        np.testing.assert_array_equal(x, y)
