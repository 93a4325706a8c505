"""The allocation-free query search (egt_grid_knn_query_device_noalloc) through the C ABI against the allocating
entry point (egt_grid_knn_query_device) on the same inputs: bit-identical rows, on two calls in a row, the
second with a smaller input and another k; plus its size limit. Its temporaries live in a fixed device array
that keeps the last call's contents, so the second call also checks that nothing stale leaks into its rows."""
import ctypes

import numpy as np
import pytest

from multigpuegtsmoothing_amd import datasets as D

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def _noalloc(E, torch, g, q, k, thres=1e8):
    nq = q[0].numel()
    rows = torch.full((max(nq * k, 1),), -7, dtype=torch.int32, device="cuda")  # every entry must be written
    st = E.lib().egt_grid_knn_query_device_noalloc(g.handle, q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), nq,
                                                    k, ctypes.c_float(thres), rows.data_ptr())
    assert st == 0, E.lib().egt_last_error_string()
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipStreamSynchronize(ctypes.c_void_p(g.stream())) == 0
    return rows[:nq * k].cpu().numpy()


def _allocating(E, g, q, k, thres=1e8):
    g.GRIDCUDAKNNCompact(q, k, thres)
    p, nq, kk = g.GetKNNIndexesQueryDevice()
    assert nq == q[0].numel() and kk == k
    return E.device_array(p, nq * k, np.int32)


def test_noalloc_query_matches_allocating_entry_point(egt_mod):
    import torch

    E = egt_mod
    ref = tuple(torch.from_numpy(a).cuda() for a in D.torus(200_000, seed=61))
    g = E.GridStructure(ref, 1, 1e-6, 1024, 4)
    cases = []
    qa = tuple(torch.from_numpy(a).cuda() for a in D.torus(300_000, seed=62))
    cases.append(("300k queries, k=16", qa, 16))
    # second call: fewer queries (a tenth of them outside the reference cube) and a longer list
    x, y, z = D.torus(5_000, seed=63)
    x[::10] *= 3.0
    qb = tuple(torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z))
    cases.append(("5k queries, k=40", qb, 40))
    for label, q, k in cases:
        _log("noalloc case:", label)
        got = _noalloc(E, torch, g, q, k)
        want = _allocating(E, g, q, k)
        assert (got != -7).all()
        np.testing.assert_array_equal(got, want)
        _log("  rows equal;", "valid entries", int((got >= 0).sum()), "of", got.size)
    g.close()


def test_noalloc_query_limits(egt_mod):
    import torch

    E = egt_mod
    ref = tuple(torch.from_numpy(a).cuda() for a in D.sphere(20_000, seed=64))
    g = E.GridStructure(ref, 1, 1e-6, 1024, 4)
    q = tuple(torch.zeros(8, dtype=torch.float32, device="cuda") for _ in range(3))
    rows = torch.empty(8 * 4, dtype=torch.int32, device="cuda")
    L = E.lib()
    # over the scratch's query capacity: refused before any work is queued
    assert L.egt_grid_knn_query_device_noalloc(g.handle, q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(),
                                               (1 << 20) + 1, 4, ctypes.c_float(1e8), rows.data_ptr()) == 4
    assert L.egt_grid_knn_query_device_noalloc(g.handle, q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), 8, 65,
                                               ctypes.c_float(1e8), rows.data_ptr()) == 4
    assert L.egt_grid_knn_query_device_noalloc(g.handle, q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), 8, 4,
                                               ctypes.c_float(1e8), None) == 1
    # an empty query set is a no-op
    assert L.egt_grid_knn_query_device_noalloc(g.handle, None, None, None, 0, 4, ctypes.c_float(1e8), None) == 0
    # after smoothing moved the points the search structure is stale: refused until a search rebuilds it
    g.GRIDCUDAKNNSELF_COMPACT(8, 1e8)
    E.EGT_CUDA(g).PerformSmoothing(2, 0.5, -0.53, 0.65)
    assert L.egt_grid_knn_query_device_noalloc(g.handle, q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), 8, 4,
                                               ctypes.c_float(1e8), rows.data_ptr()) == 6
    g.close()
