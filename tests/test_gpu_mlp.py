"""GPU tests of the learned radius proposal: the library's BF16-MFMA forward / backward and Adam kernels
against the CPU restatement (oracle/mlp_ref.py), training on the GPU, and neighbour rows of the learned
start identical to the plain search's (and to the oracle's).

Tolerances (stated here, derived from the rounding): both sides round the same operands to bfloat16, so
differences come from FP32 accumulation order on the device (about 1e-7 relative) and from the rare
activation whose FP32 value lies within that distance of a bfloat16 rounding boundary (one bf16 ulp,
2^-8 relative, on that activation). Outputs: mean |dy| <= 1e-4 and max |dy| <= 2e-2; loss within 1e-3
relative; gradient within 5e-3 relative in the 2-norm, per parameter block. Adam: float32 ops in the same
order on both sides, max |dp| <= 1e-6 (powf may differ by an ulp).
"""
import numpy as np
import pytest

from multigpuegtsmoothing_amd import datasets as D
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2), "W2": (M.OFF_W2, M.OFF_B2),
          "b2": (M.OFF_B2, M.OFF_W3), "W3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.NPARAM)}


def _log(*a):
    print(*a, flush=True)


def _check_grad(g, gr, tol=5e-3):
    assert np.all(np.isfinite(g))
    assert np.linalg.norm(g - gr) <= tol * np.linalg.norm(gr)
    for name, (a, b) in BLOCKS.items():
        ref = np.linalg.norm(gr[a:b])
        assert np.linalg.norm(g[a:b] - gr[a:b]) <= tol * ref + 1e-7, name


@pytest.mark.parametrize("n", [1, 77, 2500, 40000])
def test_forward_and_gradient_match_oracle(egt_mod, n):
    E = egt_mod
    _log("mlp eval n", n)
    model = E.RadiusModel(seed=1)
    params = M.init_params(5)
    model.set_params(params)
    np.testing.assert_array_equal(model.get_params(), params)
    X, T = M.synthetic_batch(n, 100 + n)
    y, g, loss = model.eval(X, T)
    yr, lr_, gr = M.forward_backward(params, X, T, rounding=M.bf16)
    dy = np.abs(y - yr)
    _log("  max|dy|", dy.max(), "mean|dy|", dy.mean(), "loss", loss, lr_,
         "grad rel", np.linalg.norm(g - gr) / np.linalg.norm(gr))
    assert dy.max() <= 2e-2 and dy.mean() <= 1e-4
    assert abs(loss - lr_) <= 1e-3 * lr_
    _check_grad(g, gr)
    # the device gradient also tracks the FP64 gradient pinned by finite differences (bf16 rounding only)
    _, _, g64 = M.forward_backward(params, X, T, rounding=None)
    assert np.linalg.norm(g - g64) <= 0.05 * np.linalg.norm(g64)
    model.close()


def test_adam_steps_match_oracle(egt_mod):
    E = egt_mod
    model = E.RadiusModel(seed=2)
    params = M.init_params(9)
    model.set_params(params)
    X, T = M.synthetic_batch(1000, 31)
    lr = 1e-3
    g1, l1 = model.adam_step(X, T, lr)
    _, _, gr = M.forward_backward(params, X, T, rounding=M.bf16)
    _check_grad(g1, gr)
    p1 = model.get_params()
    p1r, m, v = M.adam(params, np.zeros(M.NPARAM), np.zeros(M.NPARAM), g1, 1, lr)
    assert np.max(np.abs(p1 - p1r)) <= 1e-6
    g2, l2 = model.adam_step(X, T, lr)
    p2 = model.get_params()
    p2r, _, _ = M.adam(p1, m, v, g2, 2, lr)
    assert np.max(np.abs(p2 - p2r)) <= 1e-6
    assert np.all(p2[M.OFF_W1 + M.NF * M.NH:M.OFF_B1] == 0.0)  # padding rows stay zero
    # a few more steps lower the loss on the batch
    for _ in range(40):
        model.adam_step(X, T, 3e-3)
    _, _, l3 = model.eval(X, T)
    _log("adam loss", l1, "->", l3)
    assert l3 < 0.5 * l1
    model.close()


def test_learned_start_is_bit_exact_and_reports_fallbacks(egt_mod, oracle_mod):
    E, O = egt_mod, oracle_mod
    k = 40
    train = D.mixture(200_000, seed=41)
    evalc = D.mixture(200_000, seed=42)
    model = E.RadiusModel(seed=3)
    gt = E.GridStructure(train, 1, 1e-6, 1024, 4)
    gt.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
    X, T, ref = E.model_features(gt, k)
    ok = np.isfinite(T)
    _, _, loss0 = model.eval(X[ok], T[ok])
    loss = model.train(gt, k, epochs=6, batch=4096, max_samples=0, lr=2e-3, quantile=0.9)
    var = float(np.var(T[ok]))
    _log("train: ref level", ref, "initial mse", loss0, "final mse", loss, "target variance", var,
         "safety", model.safety)
    assert loss < 0.5 * var and loss < loss0

    ga = E.GridStructure(evalc, 1, 1e-6, 1024, 4)
    tuned = int(ga.autotune(k)[0])
    ga.close()
    rows, stats = {}, {}
    # learned_level: start at the proposed level; learned_bound: the autotuned level, candidates bounded by
    # the proposed radius; learned: both; learned_tight: both with the bound shrunk 8x, so most queries hold
    # fewer than k points inside it and are searched again without it; learned_wave*: the same with a wave per
    # query
    modes = ("own_cell", "occupancy", "autotuned", "learned_level", "learned_bound", "learned", "learned_tight",
             "learned_wave", "learned_wave_bound", "learned_wave_tight")
    bsafe = model.bound_safety
    for mode in modes:
        g = E.GridStructure(evalc, 1, 1e-6, 1024, 4)
        g.set_option(E.OPT_KNN_LONG_QUERY, 0)  # every query's whole work in the counters compared below
        if mode == "occupancy":
            g.set_option(E.OPT_START_LEVEL, E.START_LEVEL_OCCUPANCY)  # one global level from the mean occupancy
        if mode in ("autotuned", "learned_bound"):
            g.set_option(E.OPT_START_LEVEL, tuned)
        if mode.startswith("learned"):
            model.bound_safety = bsafe - 3.0 if mode.endswith("tight") else bsafe
            model.propose(g, k)
            g.set_option(E.OPT_LEARNED_MODE, {"learned_level": 1, "learned_bound": 2, "learned_wave": 1}.get(mode, 3))
            if mode.startswith("learned_wave"):  # a wave per query: each query starts at its own proposed level
                g.set_option(E.OPT_KNN_LANES, 64)
            if mode == "learned":
                lv = E.model_levels(g)
                _log("  proposed levels histogram (from level -3)",
                     np.bincount(lv.astype(np.int64) + E.SUB_LEVELS, minlength=13).tolist())
        g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
        rows[mode] = g.GetKNNIndexes()
        stats[mode] = g.stats()
        g.close()
    model.bound_safety = bsafe
    og = O.Grid(*evalc)
    want = og.to_original(og.knn(k)[0])
    for mode in rows:
        np.testing.assert_array_equal(rows[mode], want, err_msg=mode)
    st = stats["learned"]
    n = evalc[0].size
    assert st["learned_queries"] == n
    rate = st["learned_fallbacks"] / n
    for mode, s in stats.items():
        _log(f"  {mode}: candidates/query {s['knn_candidates'] / n:.1f} cells/query {s['knn_cells'] / n:.1f}"
             f" insertions/query {s['knn_insertions'] / n:.1f} knn_ms {s['knn_ms']:.3f}"
             f" propose_ms {s['propose_ms']:.3f} fallbacks {s['learned_fallbacks']}"
             f" handed back {s['learned_bound_retries']}")
    _log("  fallback rate", rate, "bound safety", bsafe, "autotuned level", tuned)
    assert rate < 0.5
    # the shrunk bound hands most queries back, and their second search still gives the exact rows
    assert stats["learned_tight"]["learned_bound_retries"] > n // 2
    assert stats["learned_wave_tight"]["learned_bound_retries"] > n // 2
    assert stats["learned_wave_bound"]["knn_insertions"] < stats["learned_wave"]["knn_insertions"]
    assert stats["learned"]["learned_bound_retries"] < 0.25 * n
    # the bound admits only candidates within the proposed radius into the lists
    assert stats["learned_bound"]["knn_insertions"] < stats["autotuned"]["knn_insertions"]
    # gain over the method without the model, against the best start the plain search has for this cloud
    # (the autotuner's level, on the autotuner's own work measure: candidates + 3 x cell lookups): the
    # per-query level does less work than the one tuned global level, and unlike the occupancy level it does
    # not over-scan dense regions (far fewer candidates). Wall-clock on the full workloads is the bench's.
    def work(s):
        return s["knn_candidates"] + 3 * s["knn_cells"]

    assert work(stats["learned_level"]) < work(stats["autotuned"])
    assert stats["learned"]["knn_candidates"] < 0.5 * stats["occupancy"]["knn_candidates"]
    # the learned-start tuner (level offset x proposal use x shape) and level offsets set by hand
    g = E.GridStructure(evalc, 1, 1e-6, 1024, 4)
    model.propose(g, k)
    off, lmode, shape, ms = g.autotune_learned(k)
    _log("  learned tuner: offset", off, "mode", lmode, "shape", shape, "ms", np.round(ms, 3).tolist())
    assert off in (0, 1) and lmode in (1, 3) and shape in (0, 2)
    assert ms[off, lmode // 2, shape // 2] == ms.min()
    g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
    np.testing.assert_array_equal(g.GetKNNIndexes(), want, err_msg="tuned learned start")
    for o in (-2, 2):
        for lanes in (1, 64):
            g.set_option(E.OPT_LEARNED_LEVEL_OFFSET, o)
            g.set_option(E.OPT_KNN_LANES, lanes)
            g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
            np.testing.assert_array_equal(g.GetKNNIndexes(), want, err_msg=f"offset {o} lanes {lanes}")
    with pytest.raises(E.EGTError):
        g.set_option(E.OPT_LEARNED_LEVEL_OFFSET, 4)
    g.close()
    model.close()


def test_model_argument_errors(egt_mod):
    E = egt_mod
    model = E.RadiusModel(seed=4)
    with pytest.raises(E.EGTError):
        model.set_params(np.zeros(10, np.float32))
    pts = D.sphere(2_000, seed=5)
    g = E.GridStructure(pts, 1, 1e-6, 512, 4)
    with pytest.raises(E.EGTError) as ei:
        model.propose(g, 16)  # features live on the 1024-cell pyramid
    assert ei.value.status == 4
    g.close()
    model.close()
