"""Parity at the workloads' full sizes (the bench's configurations), against the oracle run on the host's
cores: the 4M-point torus (k=40, 300 iterations, normals), the 655,362-vertex icosphere mesh (8 subdivisions,
150 iterations, normals), 2M points of the 16M-point mixture's generator with the learned start at k=40 and
k=64, and 64 clouds of the uneven batch with the learned start. Bars as in the small parity tests: kNN rows,
EGG rows and FP32 coordinates bit-exact; FP32 coordinates within 1e-5 of the bounding-box diagonal of the FP64
oracle run; normals bit-exact (and so sign agreement on every point). Each test stays within about two minutes
on 16 host cores."""
import numpy as np
import pytest

from multigpuegtsmoothing_amd import datasets as D

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torus4m(oracle_mod):
    """The torus workload (4,000,000 points, seed 2) through the oracle's kNN and EGG (sorted space)."""
    O = oracle_mod
    pts = D.torus(4_000_000, seed=2)
    og = O.Grid(*pts)
    knn, _, inv, _ = og.knn(40)
    egg = O.egg(og.xs, og.ys, og.zs, knn, 0.65)
    _log("torus 4M oracle: invalid rows", int(inv.sum()), "kept", int((egg >= 0).sum()))
    return pts, og, knn, egg


@pytest.fixture(scope="module")
def torus4m_gpu(egt_mod, torus4m):
    """The library's run of the same workload (autotuned start level, as the bench): kNN rows before
    smoothing, EGG rows, output coordinates, sorted-space coordinates and normals."""
    E = egt_mod
    pts = torus4m[0]
    g = E.GridStructure(pts, 1, 1e-6, 1024, 4)
    lvl = g.autotune(40)[0]
    g.set_option(E.OPT_START_LEVEL, lvl)
    g.GRIDCUDAKNNSELF_COMPACT(40, 1e8)
    knn = g.GetKNNIndexes()
    E.EGT_CUDA(g).PerformSmoothing(300, 0.5, -0.53, 0.65)
    egg = g.GetKNNIndexes()
    out = E.getOutputPoints(g)
    px, py, pz = g.GetRefPointsDevice()
    n = g.pointsRefCount()
    sorted_xyz = tuple(E.device_array(p, n, np.float32) for p in (px, py, pz))
    nrm = E.Normal_CUDA(g).GetNormals()
    st = g.stats()
    g.close()
    _log("torus 4M gpu: start level", lvl, "taubin tiles: overflow", st["taubin_overflow_tiles"], "CSR",
         st["taubin_fallback_tiles"])
    return {"knn": knn, "egg": egg, "out": out, "sorted": sorted_xyz, "normals": nrm, "stats": st}


def test_torus_4m_rows_coordinates_normals(oracle_mod, torus4m, torus4m_gpu):
    O = oracle_mod
    pts, og, knn, egg = torus4m
    G = torus4m_gpu
    np.testing.assert_array_equal(G["knn"], og.to_original(knn))
    np.testing.assert_array_equal(G["egg"], og.to_original(egg))
    assert G["stats"]["egg_kept"] == int((egg >= 0).sum())
    sx, sy, sz = O.taubin_f32(og.xs, og.ys, og.zs, egg, 300, 0.5, -0.53)
    for a, b in zip(G["sorted"], (sx, sy, sz)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(G["out"], og.output(sx, sy, sz)):
        np.testing.assert_array_equal(a, b)
    nx, ny, nz = O.pca_normals(sx, sy, sz, egg)
    onx, ony, onz, ostats, _ = O.orient_normals(egg, nx, ny, nz)
    want = O.normals_original(og, onx, ony, onz)
    dot = sum(a * b for a, b in zip(G["normals"], want))
    _log("normals: sign agreement", float(np.mean(dot > 0)), "msf edges", ostats[1], "trees", ostats[2])
    assert np.mean(dot > 0) >= 0.999
    for a, b in zip(G["normals"], want):
        np.testing.assert_array_equal(a, b)
    assert G["stats"]["msf_edges"] == ostats[1] and G["stats"]["msf_trees"] == ostats[2]


def test_torus_4m_fp64_bound(oracle_mod, torus4m, torus4m_gpu):
    O = oracle_mod
    _, og, _, egg = torus4m
    dx, dy, dz = O.taubin_f64(og.xs, og.ys, og.zs, egg, 300, 0.5, -0.53)
    gx, gy, gz = torus4m_gpu["sorted"]
    err = np.sqrt((gx - dx) ** 2 + (gy - dy) ** 2 + (gz - dz) ** 2)
    diag = np.sqrt(3.0)  # normalized cube [0, 1]^3
    _log("torus 4M: max err / diag", err.max() / diag, "rms / diag", np.sqrt((err ** 2).mean()) / diag)
    assert err.max() / diag <= 1e-5


def test_icosphere_8_mesh_full(egt_mod, oracle_mod):
    E, O = egt_mod, oracle_mod
    x, y, z, tri = D.icosphere(8, seed=3)
    assert x.size == 655_362 and tri.shape[0] == 1_310_720
    k, iters, alpha = 40, 150, 1e-8
    m = E.GridMeshCuda((x, y, z), tri)
    m.GetNeighborsFromTriangles(k)
    rows = m.GetKNNIndexes()
    E.EGT_CUDA(m).PerformSmoothing(iters, 0.5, -0.53, alpha)
    egg = m.GetKNNIndexes()
    out = E.getOutputPoints(m)
    nrm = E.Normal_CUDA(m).GetNormals()
    m.close()
    ref = O.pipeline(x, y, z, k, iters, 0.5, -0.53, alpha, normals=True, mesh_tri=tri)
    og = ref["grid"]
    np.testing.assert_array_equal(rows, og.to_original(ref["knn"]))
    np.testing.assert_array_equal(egg, og.to_original(ref["egg"]))
    for a, b in zip(out, ref["output"]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(nrm, ref["normals"]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("k", [40, 64])
def test_mixture_2m_learned_rows(egt_mod, oracle_mod, k):
    """2M points of the mixture generator (seed 4), the learned start (level and radius bound) of a model
    trained on the GPU on 500k points of seed 1000: kNN and EGG rows bit-exact against the oracle."""
    E, O = egt_mod, oracle_mod
    pts = D.mixture(2_000_000, seed=4)
    model = E.RadiusModel(seed=17)
    gt = E.GridStructure(D.mixture(500_000, seed=1000), 1, 1e-6, 1024, 4)
    model.train(gt, k, epochs=4, batch=4096, max_samples=500_000, lr=2e-3, quantile=0.9)
    gt.close()
    g = E.GridStructure(pts, 1, 1e-6, 1024, 4)
    model.propose(g, k)
    g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
    st = g.stats()
    rows = g.GetKNNIndexes()
    E.EGT_CUDA(g).PerformSmoothing(1, 0.5, -0.53, 0.65)
    egg = g.GetKNNIndexes()
    g.close()
    model.close()
    n = pts[0].size
    _log(f"mixture 2M k={k}: fallbacks {st['learned_fallbacks'] / n:.4f} handed back "
         f"{st['learned_bound_retries'] / n:.4f} invalid {st['invalid_rows']}")
    assert st["learned_queries"] == n
    og = O.Grid(*pts)
    knn, _, inv, _ = og.knn(k)
    np.testing.assert_array_equal(rows, og.to_original(knn))
    assert st["invalid_rows"] == int(inv.sum())
    np.testing.assert_array_equal(egg, og.to_original(O.egg(og.xs, og.ys, og.zs, knn, 0.65)))


def test_batch_64_clouds_learned(egt_mod, oracle_mod):
    """The first 64 clouds of the uneven batch (seeds 5000..5063, 5k-500k points each, k from {16, 40, 64}),
    learned start (model trained on 16 clouds of seeds 9000..), 100 iterations: per cloud, kNN rows, EGG rows
    and coordinates bit-exact against the oracle run of that cloud alone."""
    E, O = egt_mod, oracle_mod
    clouds, ks = D.batched_clouds(count=64, seed=5000)
    ks = np.ascontiguousarray(ks, np.int32)
    tclouds, tks = D.batched_clouds(count=16, seed=9000)
    model = E.RadiusModel(seed=13)
    tb = E.BatchGrid(tclouds)
    model.train(tb, np.ascontiguousarray(tks, np.int32), epochs=4, batch=4096, max_samples=1_000_000, lr=2e-3,
                quantile=0.9)
    tb.close()
    b = E.BatchGrid(clouds)
    model.propose(b, ks)
    b.knn(ks)
    st = b.stats()
    total = sum(c[0].size for c in clouds)
    _log("batch 64 clouds:", total, "points; fallbacks", st["learned_fallbacks"], "handed back",
         st["learned_bound_retries"])
    rows = [b.GetKNNIndexes(c) for c in range(len(clouds))]
    its = 100
    E.EGT_CUDA(b).PerformSmoothing(its, 0.5, -0.53, 0.65)
    eggs = [b.GetKNNIndexes(c) for c in range(len(clouds))]
    out = b.split(np.stack(E.getOutputPoints(b), 1))
    b.close()
    model.close()
    for c, pts in enumerate(clouds):
        k = int(ks[c])
        og = O.Grid(*pts)
        knn, _, _, _ = og.knn(k)
        np.testing.assert_array_equal(rows[c], og.to_original(knn), err_msg=f"knn {c}")
        egg = O.egg(og.xs, og.ys, og.zs, knn, 0.65)
        np.testing.assert_array_equal(eggs[c], og.to_original(egg), err_msg=f"egg {c}")
        sx, sy, sz = O.taubin_f32(og.xs, og.ys, og.zs, egg, its, 0.5, -0.53)
        np.testing.assert_array_equal(out[c], np.stack(og.output(sx, sy, sz), 1), err_msg=f"coords {c}")
