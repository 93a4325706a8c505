"""Batches of independent clouds (one handle, one launch per stage and neighbour-count class) against the
oracle run on every cloud on its own: kNN rows, invalidation counts, EGG rows, smoothed coordinates
bit-exact per cloud, normals within the smoke criterion."""
import numpy as np
import pytest

from multigpuegtsmoothing_amd import datasets as D

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def _clouds():
    specs = [("sphere", 3000, 16), ("superquadric", 700, 40), ("torus", 9000, 64), ("mixture", 5000, 40),
             ("superquadric", 65, 8), ("sphere", 20000, 40), ("superquadric", 2500, 16), ("torus", 1200, 24),
             ("superquadric", 12000, 64), ("sphere", 130, 40), ("mixture", 4000, 16), ("superquadric", 6100, 40)]
    clouds, ks = [], []
    for i, (gen, n, k) in enumerate(specs):
        clouds.append(getattr(D, gen)(n, seed=900 + i))
        ks.append(k)
    return clouds, np.array(ks, np.int32)


@pytest.mark.parametrize("persistent,lanes,cap", [(0, 1, None), (1, 1, None), (0, 64, None), (0, 1, 48)])
def test_batch_matches_separate_oracle_runs(egt_mod, oracle_mod, persistent, lanes, cap):
    """The kNN as one wave per 64-query chunk with a lane per query (default), as a work-queue launch, with
    a wave per query, or with most queries handed on to the wave-per-query pass (cap of 48 candidates; the
    pass then takes queries of every cloud from one list)."""
    E, O = egt_mod, oracle_mod
    clouds, ks = _clouds()
    lam, mu, alpha, its = 0.5, -0.53, 0.65, 12
    _log("batch of", len(clouds), "clouds, sizes", [c[0].size for c in clouds], "persistent", persistent,
         "lanes", lanes, "cap", cap)
    b = E.BatchGrid(clouds)
    b.set_option(E.OPT_KNN_PERSISTENT, persistent)
    b.set_option(E.OPT_KNN_LANES, lanes)
    if cap is not None:
        b.set_option(E.OPT_KNN_LONG_QUERY, cap)
    b.knn(ks)
    if cap is not None:
        long_q = b.stats()["knn_long_queries"]
        _log("queries handed on", long_q)
        assert long_q > int(b.sizes.sum()) // 4
    refs = []
    for c, pts in enumerate(clouds):
        ref = O.pipeline(*pts, int(ks[c]), its, lam, mu, alpha, normals=True)
        refs.append(ref)
        og = ref["grid"]
        n, k, mn, mx = b.cloud_info(c)
        assert (n, k) == (pts[0].size, ks[c])
        assert mn == np.float32(og.min_extent) and mx == np.float32(og.max_extent)
        np.testing.assert_array_equal(b.GetKNNIndexes(c), og.to_original(ref["knn"]), err_msg=f"cloud {c}")
    E.EGT_CUDA(b).PerformSmoothing(its, lam, mu, alpha)
    for c in range(len(clouds)):
        og = refs[c]["grid"]
        np.testing.assert_array_equal(b.GetKNNIndexes(c), og.to_original(refs[c]["egg"]), err_msg=f"egg {c}")
    out = np.stack(E.getOutputPoints(b), 1)
    nrm = np.stack(E.Normal_CUDA(b).GetNormals(), 1)
    for c, (o, nn) in enumerate(zip(b.split(out), b.split(nrm))):
        want = np.stack(refs[c]["output"], 1)
        np.testing.assert_array_equal(o, want, err_msg=f"coords {c}")
        dot = np.sum(nn * np.stack(refs[c]["normals"], 1), 1)
        assert np.mean(dot > 0.99) >= 0.999, c
    b.close()


def test_batch_same_k_neighbouring_clouds(egt_mod, oracle_mod):
    """Neighbouring clouds with the same k and sizes that are not multiples of 16: the CSR fill reads the 16 rows
    of a chunk that crosses a cloud boundary as one block (the rows of consecutive Morton positions lie back to
    back across clouds). EGG rows and smoothed coordinates equal the oracle's, cloud by cloud."""
    E, O = egt_mod, oracle_mod
    clouds = [D.sphere(1003, seed=950), D.torus(2011, seed=951), D.superquadric(517, seed=952), D.sphere(77, seed=953)]
    ks = np.array([40, 40, 40, 16], np.int32)
    lam, mu, alpha, its = 0.5, -0.53, 0.65, 8
    b = E.BatchGrid(clouds)
    b.knn(ks)
    E.EGT_CUDA(b).PerformSmoothing(its, lam, mu, alpha)
    out = np.stack(E.getOutputPoints(b), 1)
    for c, (pts, o) in enumerate(zip(clouds, b.split(out))):
        _log("cloud", c, "n", pts[0].size, "k", int(ks[c]))
        ref = O.pipeline(*pts, int(ks[c]), its, lam, mu, alpha, normals=False)
        og = ref["grid"]
        np.testing.assert_array_equal(b.GetKNNIndexes(c), og.to_original(ref["egg"]), err_msg=f"egg {c}")
        np.testing.assert_array_equal(o, np.stack(ref["output"], 1), err_msg=f"coords {c}")
    b.close()


def test_batch_with_one_point_clouds(egt_mod, oracle_mod):
    """Clouds of 1 and 2 points (and a 700-point sphere whose rows the k=40 certification blanks) between clouds
    that keep neighbours: kNN rows, EGG rows and smoothed coordinates equal the oracle's cloud by cloud, and
    the blanked clouds' normals are the fallback x axis as in the oracle."""
    E, O = egt_mod, oracle_mod
    clouds = [D.sphere(1, seed=960), D.torus(3000, seed=961), D.sphere(2, seed=962), D.superquadric(1500, seed=963),
              D.sphere(1, seed=964), D.sphere(700, seed=965)]
    ks = np.array([40, 16, 8, 40, 16, 40], np.int32)
    lam, mu, alpha, its = 0.5, -0.53, 0.65, 6
    b = E.BatchGrid(clouds)
    b.knn(ks)
    refs = []
    for c, pts in enumerate(clouds):
        _log("cloud", c, "n", pts[0].size, "k", int(ks[c]))
        ref = O.pipeline(*pts, int(ks[c]), its, lam, mu, alpha, normals=True)
        refs.append(ref)
        np.testing.assert_array_equal(b.GetKNNIndexes(c), ref["grid"].to_original(ref["knn"]), err_msg=f"cloud {c}")
    E.EGT_CUDA(b).PerformSmoothing(its, lam, mu, alpha)
    out = np.stack(E.getOutputPoints(b), 1)
    nrm = np.stack(E.Normal_CUDA(b).GetNormals(), 1)
    for c, (o, nn) in enumerate(zip(b.split(out), b.split(nrm))):
        ref = refs[c]
        np.testing.assert_array_equal(b.GetKNNIndexes(c), ref["grid"].to_original(ref["egg"]), err_msg=f"egg {c}")
        np.testing.assert_array_equal(o, np.stack(ref["output"], 1), err_msg=f"coords {c}")
        want = np.stack(ref["normals"], 1)
        if (ref["egg"] < 0).all():
            np.testing.assert_array_equal(nn, want, err_msg=f"normals {c}")
        else:
            assert np.mean(np.sum(nn * want, 1) > 0.99) >= 0.999, c
    b.close()


def test_batch_class_shape_tuner_keeps_rows(egt_mod):
    """The per-class shape tuner times each neighbour-count class in each launch shape and keeps the fastest;
    rows after it, and on a new handle given its choices through OPT_KNN_CLASS_SHAPES, equal the default
    batch's (itself checked against the oracle above). Every packed field is a shape the tuner ran."""
    E = egt_mod
    clouds, ks = _clouds()
    ref = E.BatchGrid(clouds)
    ref.knn(ks)
    want = [ref.GetKNNIndexes(c) for c in range(len(clouds))]
    ref.close()
    b = E.BatchGrid(clouds)
    packed, ms = b.autotune_shape(ks)
    _log("class shapes", hex(packed), "ms per class (lane, persistent, wave)", np.round(ms, 3).tolist())
    classes = sorted({(int(k) + 7) // 8 - 1 for k in ks})
    for ci in range(8):
        field = (packed >> (2 * ci)) & 3
        if ci in classes:
            assert field in (1, 2, 3) and ms[ci, field - 1] == ms[ci][ms[ci] >= 0].min()
        else:
            assert field == 0 and (ms[ci] < 0).all()
    b.knn(ks)
    for c in range(len(clouds)):
        np.testing.assert_array_equal(b.GetKNNIndexes(c), want[c], err_msg=f"cloud {c}")
    b.close()
    # every class forced to each shape through the option
    for shape in (1, 2, 3):
        f = E.BatchGrid(clouds)
        f.set_option(E.OPT_KNN_CLASS_SHAPES, sum(shape << (2 * ci) for ci in range(8)))
        f.knn(ks)
        for c in range(len(clouds)):
            np.testing.assert_array_equal(f.GetKNNIndexes(c), want[c], err_msg=f"shape {shape} cloud {c}")
        f.close()
    bad = E.BatchGrid(clouds[:2])
    with pytest.raises(E.EGTError):
        bad.set_option(E.OPT_KNN_CLASS_SHAPES, 1 << 16)
    bad.close()


def test_batch_invalid_rows_and_single_cloud(egt_mod, oracle_mod):
    E, O = egt_mod, oracle_mod
    # a dense blob (cells above the 1e5 rule are not needed: an isolated outlier is blanked too)
    pts = D.sphere(4000, seed=31)
    x, y, z = (np.concatenate([a, [np.float32(40.0)]]) for a in pts)
    clouds = [(x, y, z), D.torus(3000, seed=32)]
    b = E.BatchGrid(clouds)
    b.knn([12, 20])
    for c, cl in enumerate(clouds):
        og = O.Grid(*cl)
        knn, _, inv, _ = og.knn(int(b.k[c]))
        np.testing.assert_array_equal(b.GetKNNIndexes(c), og.to_original(knn))
    one = E.BatchGrid([clouds[1]])
    one.knn(20)
    np.testing.assert_array_equal(one.GetKNNIndexes(0), b.GetKNNIndexes(1))
    b.close()
    one.close()


@pytest.mark.parametrize("persistent", [0, 1])
def test_persistent_launch_same_rows(egt_mod, oracle_mod, persistent):
    E, O = egt_mod, oracle_mod
    pts = D.mixture(30_000, seed=33)
    g = E.GridStructure(pts, 1, 1e-6, 1024, 4)
    g.set_option(E.OPT_KNN_PERSISTENT, persistent)
    g.GRIDCUDAKNNSELF_COMPACT(40, 1e8)
    og = O.Grid(*pts)
    np.testing.assert_array_equal(g.GetKNNIndexes(), og.to_original(og.knn(40)[0]))
    g.close()


def test_batch_errors(egt_mod):
    E = egt_mod
    b = E.BatchGrid([D.sphere(500, seed=1), D.sphere(600, seed=2)])
    with pytest.raises(E.EGTError) as ei:
        E._check(E.lib().egt_grid_knn_self(b.handle, 8, 1e8))
    assert ei.value.status == 6
    with pytest.raises(E.EGTError):
        b.knn([8, 65])
    b.close()


def test_learned_proposal_on_batch(egt_mod, oracle_mod):
    E, O = egt_mod, oracle_mod
    train = [D.superquadric(n, seed=700 + i) for i, n in enumerate([20000, 8000, 30000, 5000, 12000, 25000])]
    ktrain = np.array([16, 40, 64, 16, 40, 64], np.int32)
    evalc, keval = _clouds()
    model = E.RadiusModel(seed=5)
    bt = E.BatchGrid(train)
    loss = model.train(bt, ktrain, epochs=6, batch=2048, max_samples=0, lr=2e-3, quantile=0.9)
    _log("batch training mse", loss, "safety", model.safety)
    assert np.isfinite(loss) and loss < 1.0
    bt.close()
    plain = E.BatchGrid(evalc)
    plain.knn(keval)
    learned = E.BatchGrid(evalc)
    model.propose(learned, keval)
    learned.knn(keval)
    st = learned.stats()
    n = int(plain.sizes.sum())
    _log("learned batch: queries", st["learned_queries"], "fallbacks", st["learned_fallbacks"],
         "cells/query", st["knn_cells"] / n, "plain cells/query", plain.stats()["knn_cells"] / n)
    assert st["learned_queries"] == n
    # the same proposal with a wave per query: each query starts at its own proposed level
    wave = E.BatchGrid(evalc)
    wave.set_option(E.OPT_KNN_LANES, 64)
    model.propose(wave, keval)
    wave.knn(keval)
    _log("learned batch, wave per query: fallbacks", wave.stats()["learned_fallbacks"])
    for c in range(len(evalc)):
        np.testing.assert_array_equal(learned.GetKNNIndexes(c), plain.GetKNNIndexes(c), err_msg=f"cloud {c}")
        np.testing.assert_array_equal(wave.GetKNNIndexes(c), plain.GetKNNIndexes(c), err_msg=f"wave {c}")
    wave.close()
    og = O.Grid(*evalc[5])
    np.testing.assert_array_equal(learned.GetKNNIndexes(5), og.to_original(og.knn(int(keval[5]))[0]))
    plain.close()
    learned.close()
    model.close()
