"""Benchmark: the grid-kNN -> EGG -> Taubin path on one MI355X, learned radius proposal vs plain search.

Default (headline) workload, the one the metric is quoted on: 16,000,000 points from the 12-cluster / 4-plane
mixture (seed 4), k=40, alpha 0.65, 150 Taubin iterations (lambda 0.5, mu -0.53). One step = grid build +
(learned proposal) + kNN + EGG + Taubin, coordinates resident in HBM before timing. Three paths are timed,
each W warm-up + K timed steps: the plain search from the query's own cell (the reference schedule), the
plain search from the autotuned start level, and the learned path (a model trained on the GPU on 2M points of
the same generator with another seed: proposed start level and radius bound per query). Their neighbour rows
are compared on the GPU. value = the learned path's end-to-end points/s; the best plain path and the speed-ups
ride along, with the batched workload (512 uneven clouds) and the 4M torus (with normals) as extra keys.

Prints one JSON line. Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload headline|torus|batch|mixture|mesh|query]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k=40 kNN queries/s + Taubin point-iters/s; learned-proposal vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=4_000_000)
    p.add_argument("--k", type=int, default=40)
    p.add_argument("--iters", type=int, default=300)
    p.add_argument("--no-normals", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=250_000)
    p.add_argument("--no-learned", action="store_true", help="skip the learned-proposal vs plain comparison")
    p.add_argument("--no-query", action="store_true", help="skip the query-vs-reference search section")
    p.add_argument("--no-autotune", action="store_true",
                   help="start every kNN query at its own cell (the reference schedule) instead of the autotuned level")
    p.add_argument("--no-graph", action="store_true", help="launch the Taubin passes one by one (profiling)")
    p.add_argument("--workload", choices=["headline", "torus", "batch", "query", "mesh", "mixture"],
                   default="headline",
                   help="headline (default): the 16M-point mixture at k=40, plain own-cell / plain autotuned / "
                        "learned, with the 512-cloud batch and the 4M torus as extra keys; torus: one 4M-point "
                        "cloud with normals; batch: 512 uneven superquadric clouds; query: only the "
                        "query-vs-reference search against the torus; mesh: the 655k-vertex icosphere mesh; "
                        "mixture: the 16M-point cluster mixture, plain vs learned vs autotuned at k=40 and k=64")
    p.add_argument("--clouds", type=int, default=512)
    p.add_argument("--mixture-ks", default="40,64", help="mixture workload: neighbour counts to run")
    p.add_argument("--no-extras", action="store_true", help="headline: skip the k=64, batch and torus keys")
    p.add_argument("--no-per-cloud", action="store_true",
                   help="headline / batch: skip the one-handle-per-cloud plain path")
    p.add_argument("--mixture-iters", type=int, default=150, help="headline: Taubin iterations (150)")
    p.add_argument("--modes", default="plain_own_cell,plain_autotuned,learned",
                   help="headline: paths to time (profiling runs time one)")
    return p.parse_args()


TRAFFIC_FILES = {"torus": os.path.join(ROOT, "profiles", "k_taubin_tile_torus4M_k40_traffic.md"),
                 "mixture": os.path.join(ROOT, "profiles", "k_taubin_tile_mixture16M_k40_traffic.md")}


def measured_traffic(which="torus"):
    """HBM bytes per Taubin pass from the committed rocprofv3 PMC summary of that workload (FETCH_SIZE +
    WRITE_SIZE, scaled by the same counters on a copy kernel of known size), or None when it is absent."""
    try:
        for line in open(TRAFFIC_FILES[which]):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def gpu_step(E, torch, dev_xyz, args, start_level):
    g = E.GridStructure(dev_xyz, 1, 1e-6, 1024, 4)
    if start_level is not None:
        g.set_option(E.OPT_START_LEVEL, start_level)
    if args.no_graph:
        g.set_option(E.OPT_TAUBIN_GRAPH, 0)
    g.GRIDCUDAKNNSELF_COMPACT(args.k, 1e8)
    E.EGT_CUDA(g).PerformSmoothing(args.iters, 0.5, -0.53, 0.65)
    if not args.no_normals:
        E.Normal_CUDA(g).GetNormals()
    st = g.stats()
    g.close()
    return st


def learned_vs_plain(E, dev_xyz, args, start_level=None):
    """kNN stage of the same workload with the plain start (the query's own cell, as the reference), with the
    learned proposal (model trained on the GPU on a torus drawn with a different seed; level offset, proposal
    use and launch shape tuned by the library's learned-start tuner) and, when given, with the autotuned start
    level. Rows must be identical; the learned time includes the proposal kernel."""
    from multigpuegtsmoothing_amd import datasets as D

    k = args.k
    model = E.RadiusModel(seed=11)
    train = E.GridStructure(D.torus(2_000_000, seed=1002), 1, 1e-6, 1024, 4)
    t0 = time.perf_counter()
    mse = model.train(train, k, epochs=6, batch=4096, max_samples=2_000_000, lr=2e-3, quantile=0.9)
    train_s = time.perf_counter() - t0
    train.close()
    res = {}
    rows = {}
    # the learned start's level offset, proposal use and launch shape, tuned once by the library
    g = E.GridStructure(dev_xyz, 1, 1e-6, 1024, 4)
    model.propose(g, k)
    level_offset, learned_mode, lshape, _ = g.autotune_learned(k)
    g.close()
    modes = ("plain", "learned") + (("autotuned",) if start_level is not None else ())
    for mode in modes:
        g = E.GridStructure(dev_xyz, 1, 1e-6, 1024, 4)
        if mode == "autotuned":
            g.set_option(E.OPT_START_LEVEL, start_level)
        if mode == "learned":
            g.set_option(E.OPT_KNN_LANES, 64 if lshape == 2 else 1)
        best = None
        for _ in range(3):
            prop = 0.0
            if mode == "learned":
                model.propose(g, k)
                g.set_option(E.OPT_LEARNED_MODE, learned_mode)
                g.set_option(E.OPT_LEARNED_LEVEL_OFFSET, level_offset)
                prop = g.stats()["propose_ms"]
            g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
            st = g.stats()
            t = st["knn_ms"] + prop
            best = t if best is None else min(best, t)
        res[mode] = {"knn_ms": best, "candidates_per_query": st["knn_candidates"] / args.n,
                     "cells_per_query": st["knn_cells"] / args.n,
                     "fallback_rate": st["learned_fallbacks"] / args.n if mode == "learned" else None}
        rows[mode] = g.GetKNNIndexes()
        g.close()
    model.close()
    res["learned"].update(level_offset=level_offset, learned_mode=learned_mode,
                          shape="wave per query" if lshape == 2 else "lane per query")
    out = {"plain": res["plain"], "learned": res["learned"],
           "knn_speedup": res["plain"]["knn_ms"] / res["learned"]["knn_ms"],
           "bit_exact": bool(np.array_equal(rows["plain"], rows["learned"])),
           "train": {"cloud": "torus 2M pts, seed 1002", "seconds": train_s, "mse": mse}}
    if "autotuned" in res:
        out["autotuned"] = dict(res["autotuned"], start_level=start_level,
                                bit_exact=bool(np.array_equal(rows["plain"], rows["autotuned"])))
    return out


def cpu_baseline(args):
    from multigpuegtsmoothing_amd import datasets as D
    from oracle import oracle as O

    n = args.cpu_sample
    pts = D.torus(n, seed=2)
    t0 = time.perf_counter()
    O.pipeline(*pts, args.k, args.iters, 0.5, -0.53, 0.65, normals=not args.no_normals)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "points/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"oracle (host C++/OpenMP restatement) on a {n}-point torus from the same generator, "
                      f"same k/alpha/{args.iters} iterations/normals; {dt:.2f} s"}


def query_search(E, dev_xyz, args):
    """Query-vs-reference search (GRIDCUDAKNNCompact) against the bench's reference cloud: args.n queries drawn
    from the same torus generator with another seed (a second scan of the same surface, as in ICP), device
    input, k = 1 and k = 16. Rates use the library's device time of the search (query normalization and
    sort included). The CPU figure is the oracle's query search on a 200k-query sample against the same
    reference cloud (its grid build is not timed)."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D

    qpts = D.torus(args.n, seed=3)
    dq = tuple(torch.from_numpy(a).cuda() for a in qpts)
    g = E.GridStructure(dev_xyz, 1, 1e-6, 1024, 4)
    res = {"queries": f"torus {args.n} pts, seed 3 (same generator as the reference cloud)"}
    for k in (1, 16):
        times = []
        for _ in range(3):
            g.GRIDCUDAKNNCompact(dq, k, 1e8)
            times.append(g.stats()["query_ms"])
        st = g.stats()
        ms = float(np.median(times))
        res[f"k{k}"] = {"queries_per_s": args.n / (ms * 1e-3), "ms": ms,
                        "candidates_per_query": st["query_candidates"] / args.n,
                        "invalid_rows": st["query_invalid_rows"],
                        "achieved_GBs": args.n * (12 + 4 * k) / (ms * 1e-3) / 1e9}
    g.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        og = O.Grid(*D.torus(args.n, seed=2))
        m = 200_000
        sample = tuple(a[:m] for a in qpts)
        t0 = time.perf_counter()
        og.knn_query(*sample, 16)
        dt = time.perf_counter() - t0
        res["cpu_baseline_k16"] = {"value": m / dt, "unit": "queries/s", "cores": O.num_threads(), "kind": "port",
                                   "sample": f"oracle query search, first {m} queries against the same reference "
                                             f"cloud; {dt:.2f} s"}
    return res


def _device_rows(torch, ptr, count):
    """Device-to-device copy of a neighbour-row buffer into a torch int32 tensor (row comparison on the GPU)."""
    import ctypes

    t = torch.empty(count, dtype=torch.int32, device="cuda")
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(ptr), ctypes.c_size_t(count * 4), 3)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    return t


def mesh_main(args):
    """Mesh mode: icosphere subdivided 8 times (655,362 vertices, 1,310,720 triangles), radial noise +-0.01,
    seed 3; k=40, alpha=1e-8, 150 Taubin iterations on the mesh adjacency, then PCA normals with MST
    orientation. The mesh provider takes host arrays (as the reference's GridMeshCuda does), so each step
    includes uploading the coordinates and triangles (7.9 MB + 15.7 MB)."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    x, y, z, tri = D.icosphere(8, seed=3)
    n = x.size
    iters = 150

    def step():
        m = E.GridMeshCuda((x, y, z), tri)
        m.GetNeighborsFromTriangles(40)
        E.EGT_CUDA(m).PerformSmoothing(iters, 0.5, -0.53, 1e-8)
        E.Normal_CUDA(m).GetNormals()
        st = m.stats()
        m.close()
        return st

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = stats[-1]
    out = {"metric": "mesh-mode points/s (grid + mesh adjacency + EGG + Taubin + normals)", "value": n / dt,
           "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"icosphere 8 subdivisions ({n} vertices, {tri.shape[0]} triangles), radial noise "
                                  "0.01, seed 3, k=40, alpha=1e-8, 150 Taubin its, PCA normals + MST orientation",
                      "n_points": n},
           "stage_ms": {"grid": st["grid_ms"], "mesh_adjacency": st["knn_ms"], "egg_csr": st["egg_ms"],
                        "taubin": st["taubin_ms"], "normals": st["normals_ms"]},
           "taubin_point_iters_per_s": n * iters / (st["taubin_ms"] * 1e-3), "egg_kept_per_point": st["egg_kept"] / n}
    print(json.dumps(out), flush=True)


def mixture_main(args):
    """Learned radius proposal on 16,000,000 points from the 12-cluster / 4-plane mixture (seed 4), k=40 and
    k=64, alpha 0.65, 150 Taubin iterations. The model is trained on the GPU (outside the timing) on the plain
    search's radii of a 2M-point mixture drawn with seed 1000. One step = grid + (proposal) + kNN + EGG +
    Taubin; plain = own-cell start, learned = proposal, autotuned = the tuner's level (tuned outside the
    timing). Rows of the three are compared."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    pts = D.mixture(16_000_000, seed=4)
    n = pts[0].size
    dev = tuple(torch.from_numpy(a).cuda() for a in pts)
    del pts
    train = D.mixture(2_000_000, seed=1000)
    iters = 150
    res = {}
    for k in [int(v) for v in args.mixture_ks.split(",")]:
        model = E.RadiusModel(seed=17)
        gt = E.GridStructure(train, 1, 1e-6, 1024, 4)
        t0 = time.perf_counter()
        mse = model.train(gt, k, epochs=6, batch=4096, max_samples=2_000_000, lr=2e-3, quantile=0.9)
        train_s = time.perf_counter() - t0
        gt.close()
        g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
        lvl = int(g.autotune(k)[0])
        g.close()
        per = {}
        rows = {}
        for mode in ("plain", "learned", "autotuned"):
            def step():
                g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
                if mode == "learned":
                    model.propose(g, k)
                if mode == "autotuned":
                    g.set_option(E.OPT_START_LEVEL, lvl)
                g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
                r = g.GetKNNIndexesDevice()
                E.EGT_CUDA(g).PerformSmoothing(iters, 0.5, -0.53, 0.65)
                st = g.stats()
                return st, g, r
            g0 = step()[1]
            torch.cuda.synchronize()
            g0.close()
            ts, st = [], None
            for rep in range(2):
                t0 = time.perf_counter()
                st, g, r = step()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
                if rep == 1:
                    rows[mode] = _device_rows(torch, r, n * k)
                g.close()
            dt = min(ts)
            print(f"mixture k={k} {mode}: {dt * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
            per[mode] = {"ms_per_step": dt * 1e3, "points_per_s": n / dt,
                         "knn_ms": st["knn_ms"] + st["propose_ms"], "egg_ms": st["egg_ms"], "taubin_ms": st["taubin_ms"],
                         "grid_ms": st["grid_ms"], "candidates_per_query": st["knn_candidates"] / n,
                         "invalid_rows": st["invalid_rows"],
                         "fallback_rate": st["learned_fallbacks"] / n if mode == "learned" else None}
        res[f"k{k}"] = dict(per, start_level_autotuned=lvl,
                            rows_equal=bool(torch.equal(rows["plain"], rows["learned"])
                                            and torch.equal(rows["plain"], rows["autotuned"])),
                            learned_end_to_end_speedup=per["plain"]["ms_per_step"] / per["learned"]["ms_per_step"],
                            learned_knn_speedup=per["plain"]["knn_ms"] / per["learned"]["knn_ms"],
                            train={"cloud": "mixture 2M pts, seed 1000", "seconds": train_s, "mse": mse})
        model.close()
        del rows
    out = {"metric": "learned-proposal vs plain, mixture 16M (points/s end to end, grid+kNN+EGG+Taubin)",
           "value": next(iter(res.values()))["learned"]["points_per_s"],
           "unit": f"points/s at k={next(iter(res))[1:]}, learned", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"mixture {n} pts (12 anisotropic clusters + 4 planes, seed 4), k={args.mixture_ks}, "
                                  f"alpha 0.65, {iters} Taubin its", "n_points": n, "k": list(res)},
           "results": res}
    print(json.dumps(out), flush=True)


def query_main(args):
    """Only the query-vs-reference search (query_search) against the torus reference cloud; one JSON line."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    dev = tuple(torch.from_numpy(a).cuda() for a in D.torus(args.n, seed=2))
    res = query_search(E, dev, args)
    out = {"metric": "query-vs-reference kNN queries/s (GRIDCUDAKNNCompact)", "value": res["k1"]["queries_per_s"],
           "unit": "queries/s at k=1", "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"reference torus {args.n} pts (seed 2), {res['queries']}", "k": [1, 16]},
           "query_knn": res}
    print(json.dumps(out), flush=True)


def timed_steps(torch, step, steps, warmup):
    """W untimed warm-up steps, then exactly K steps bracketed by device synchronization. Returns the seconds
    per step and each timed step's result."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    res = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res.append(step())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / max(steps, 1), res


def batch_result(args, steps, warmup, per_cloud=True):
    """Batched uneven workload: args.clouds superquadric clouds, sizes log-uniform in [5k, 500k], k drawn
    from {16, 40, 64}, alpha 0.65, 100 iterations (lambda 0.5, mu -0.53). One step = batch build (cube, sort,
    cell tables for all clouds) + (learned proposal) + kNN (one launch per neighbour-count class over 64-query
    chunks) + EGG + Taubin (one captured graph over all clouds). Coordinates are resident in HBM before timing.
    The learned batch (model trained on the GPU on 64 clouds of other seeds) and the plain batch (every query
    from its own cell), each with the kNN launch shape per neighbour-count class that the class tuner picks
    for it (outside the timing), are each timed over `steps` steps; beside them, when per_cloud, the plain GPU path of
    the reference interface (one handle per cloud, plain search, EGG, Taubin, cloud after cloud), timed once
    after a warm-up pass."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    iters = 100 if args.iters == 300 else args.iters
    clouds, ks = D.batched_clouds(count=args.clouds, seed=5000)
    ks = np.ascontiguousarray(ks, np.int32)
    sizes = [c[0].size for c in clouds]
    dev = tuple(torch.from_numpy(np.concatenate([c[i] for c in clouds])).cuda() for i in range(3))
    n = int(sum(sizes))
    del clouds
    tclouds, tks = D.batched_clouds(count=64, seed=9000)
    model = E.RadiusModel(seed=13)
    tb = E.BatchGrid(tclouds)
    t0 = time.perf_counter()
    mse = model.train(tb, np.ascontiguousarray(tks, np.int32), epochs=4, batch=4096, max_samples=2_000_000,
                      lr=2e-3, quantile=0.9)
    train_s = time.perf_counter() - t0
    tb.close()
    del tclouds
    # launch shape per neighbour-count class, tuned once for the learned starts and once for the plain search,
    # outside the timed steps
    # (the learned batch: the proposal's level alone or with its radius bound, whichever totals less)
    shapes = {}
    b = E.BatchGrid(sizes, device=dev)
    model.propose(b, ks)
    tried = {}
    for off in (0, 1):
        for lmode in (1, 3):
            b.set_option(E.OPT_LEARNED_MODE, lmode)
            b.set_option(E.OPT_LEARNED_LEVEL_OFFSET, off)
            packed, ms = b.autotune_shape(ks)
            tried[(off, lmode)] = (float(sum(ms[ci][ms[ci] >= 0].min() for ci in range(8) if (ms[ci] >= 0).any())),
                                  packed, ms)
    b.close()
    level_offset, learned_mode = min(tried, key=lambda m: tried[m][0])
    shapes[True] = tried[(level_offset, learned_mode)][1:]
    b = E.BatchGrid(sizes, device=dev)
    shapes[False] = b.autotune_shape(ks)
    b.close()
    torch.cuda.synchronize()

    def step(learned=True):
        b = E.BatchGrid(sizes, device=dev)
        b.set_option(E.OPT_KNN_CLASS_SHAPES, shapes[learned][0])
        if learned:
            model.propose(b, ks)
            b.set_option(E.OPT_LEARNED_MODE, learned_mode)
            b.set_option(E.OPT_LEARNED_LEVEL_OFFSET, level_offset)
        b.knn(ks)
        E.EGT_CUDA(b).PerformSmoothing(iters, 0.5, -0.53, 0.65)
        st = b.stats()
        b.close()
        return st

    dt, stats = timed_steps(torch, step, steps, warmup)
    print(f"batch learned: {dt * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    plain_dt, pstats = timed_steps(torch, lambda: step(learned=False), steps, min(warmup, 1))
    print(f"batch plain: {plain_dt * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
    st, pst = stats[-1], pstats[-1]
    stage = {"grid": st["grid_ms"], "propose": st["propose_ms"], "knn": st["knn_ms"], "egg_csr": st["egg_ms"],
             "taubin": st["taubin_ms"]}
    out = {"metric": METRIC, "value": n / dt, "unit": "points/s (end-to-end batch build+proposal+kNN+EGG+Taubin)",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batch of {len(sizes)} superquadric clouds, sizes log-uniform in [5k, 500k] "
                                  f"({n} pts), k in {{16, 40, 64}}, alpha=0.65, {iters} Taubin its, learned proposal",
                      "n_points": n, "clouds": len(sizes), "iterations": iters, "path": "batch+learned"},
           "stage_ms": stage,
           "knn_queries_per_s": n / ((st["knn_ms"] + st["propose_ms"]) * 1e-3),
           "taubin_point_iters_per_s": n * iters / (st["taubin_ms"] * 1e-3),
           "learned_fallback_rate": st["learned_fallbacks"] / n,
           "learned_bound_retry_rate": st["learned_bound_retries"] / n,
           "plain_batch_ms_per_step": plain_dt * 1e3,
           "plain_knn_ms": pst["knn_ms"], "learned_knn_ms": st["knn_ms"] + st["propose_ms"],
           "knn_speedup_learned": pst["knn_ms"] / (st["knn_ms"] + st["propose_ms"]),
           "learned_end_to_end_speedup": plain_dt / dt,
           "train": {"clouds": 64, "seconds": train_s, "mse": mse},
           "learned_mode": {1: "proposed level", 3: "proposed level and radius bound"}[learned_mode],
           "learned_level_offset": level_offset,
           "knn_class_shapes": {path: {f"k<={8 * (ci + 1)}": {"shape": ("lane per query", "persistent", "wave per query")[
               ((packed >> (2 * ci)) & 3) - 1], "ms": [round(float(x), 3) for x in ms[ci]]}
               for ci in range(8) if (packed >> (2 * ci)) & 3}
               for path, (packed, ms) in (("learned", shapes[True]), ("plain", shapes[False]))}}
    if per_cloud:
        # The plain GPU path on the same clouds: one GridStructure per cloud, the plain (own-cell) search, EGG
        # and Taubin, cloud after cloud (what a caller of the reference's interface does with many clouds).
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

        def plain_path():
            for c in range(len(sizes)):
                a, b_ = int(offs[c]), int(offs[c + 1])
                g = E.GridStructure(tuple(d[a:b_] for d in dev), 1, 1e-6, 1024, 4)
                g.GRIDCUDAKNNSELF_COMPACT(int(ks[c]), 1e8)
                E.EGT_CUDA(g).PerformSmoothing(iters, 0.5, -0.53, 0.65)
                g.close()

        per_dt, _ = timed_steps(torch, plain_path, 1, 1)
        out["per_cloud_plain_path_ms"] = per_dt * 1e3
        out["speedup_vs_per_cloud_plain_path"] = per_dt / dt
    model.close()
    del dev
    return out


def batch_main(args):
    print(json.dumps(batch_result(args, args.steps, args.warmup, per_cloud=not args.no_per_cloud)), flush=True)


def torus_result(args):
    """The 4M-point torus workload (k=40, 300 iterations, normals): one JSON-ready dict."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    x, y, z = D.torus(args.n, seed=2)
    dev = tuple(torch.from_numpy(a).cuda() for a in (x, y, z))
    torch.cuda.synchronize()
    # The search's start level (the cell size of the first box; rows do not depend on it) is chosen once for
    # the workload by the library's autotuner, outside the timed steps. --no-autotune keeps the reference
    # schedule's own-cell start.
    start_level, tune = None, None
    if not args.no_autotune:
        g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
        t0 = time.perf_counter()
        lvl, ms, cost = g.autotune(args.k)
        tune = {"start_level": int(lvl), "seconds": time.perf_counter() - t0,
                "trial_ms": {str(i - E.SUB_LEVELS): round(float(v), 3) for i, v in enumerate(ms) if v >= 0},
                "trial_cost": {str(i - E.SUB_LEVELS): round(float(v), 1) for i, v in enumerate(cost) if v >= 0}}
        g.close()
        start_level = int(lvl)
    for _ in range(args.warmup):
        gpu_step(E, torch, dev, args, start_level)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(gpu_step(E, torch, dev, args, start_level))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    def med(key):
        return float(np.median([s[key] for s in stats]))

    knn_ms, egg_ms, taub_ms, nrm_ms, grid_ms = (med(k) for k in ("knn_ms", "egg_ms", "taubin_ms", "normals_ms",
                                                                 "grid_ms"))
    n, k = args.n, args.k
    nnz = stats[-1]["egg_kept"]
    passes = 2 * args.iters
    # Taubin pass, algorithmic bytes per launch: own xyz read 12 + write 12 + row ptr 4 per point, 4 B per
    # kept neighbour index
    taub_bytes = 28 * n + 4 * nnz
    # one pass: the captured lambda/mu passes alone, timed by the library with HIP events on its stream (the
    # once-per-call tile set-up and graph build are reported apart as taubin_setup_ms)
    taub_pass_ms = med("taubin_pass_ms") if passes else float("nan")
    # kNN, algorithmic bytes: query xyz 12 + 4k output per query
    knn_bytes = n * (12 + 4 * k)
    stage = {"grid": grid_ms, "knn": knn_ms, "egg_csr": egg_ms, "taubin": taub_ms, "normals": nrm_ms}
    dominant = max(stage, key=stage.get)
    if dominant == "taubin":
        ach = taub_bytes / (taub_pass_ms * 1e-3) / 1e9
        roof = {"kernel": f"k_taubin_tile<{stats[-1]['taubin_tile']}>", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic("torus"), "bytes_per_launch": taub_bytes,
                "avg_launch_ms": taub_pass_ms}
    else:
        ach = knn_bytes / (knn_ms * 1e-3) / 1e9
        roof = {"kernel": "k_knn", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": knn_bytes, "avg_launch_ms": knn_ms}
    out = {
        "metric": METRIC,
        "value": n / dt,
        "unit": "points/s (end-to-end grid+kNN+EGG+Taubin+normals)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"torus {n} pts (R=1, r=0.3, density 1-10x, sigma 0.003, seed 2), cloud mode, "
                               f"k={k}, alpha=0.65, {args.iters} Taubin its (lambda 0.5, mu -0.53)"
                               + ("" if args.no_normals else ", PCA normals + MST orientation"),
                   "n_points": n, "k": k, "iterations": args.iters, "path": "plain",
                   "knn_start_level": "own cell (reference schedule)" if start_level is None
                   else f"{start_level} (autotuned once per workload, outside the timed steps)"},
        "autotune": tune,
        "knn_queries_per_s": n / (knn_ms * 1e-3),
        "taubin_point_iters_per_s": n * args.iters / (taub_ms * 1e-3) if taub_ms > 0 else None,
        "taubin_setup_ms": med("taubin_setup_ms"),
        "stage_ms": stage,
        "egg_kept_per_point": nnz / n,
        "knn_candidates_per_query": stats[-1]["knn_candidates"] / n,
        "invalid_rows": stats[-1]["invalid_rows"],
        "roofline": roof,
    }
    if not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain(E, dev, args, start_level)
    if not args.no_query:
        out["query_knn"] = query_search(E, dev, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    del dev
    return out


def mixture_modes(args, k=40, steps=5, warmup=2):
    """The headline workload: 16M mixture points (seed 4) at neighbour count k, alpha 0.65, 150 Taubin
    iterations. Paths (each warm-up + `steps` timed steps of grid + (proposal) + kNN + EGG + Taubin):
    plain_own_cell (every query from its own cell, the reference schedule, one lane per query), plain_autotuned
    (the level the autotuner picks for this cloud and the launch shape the shape tuner picks for it, tuned once
    outside the timing) and learned (the model's start level per query, offset by 0 or +1, alone or with its
    radius bound, in the launch shape: whichever the library's learned-start tuner times fastest;
    model trained on the GPU, outside the timing, on the plain search's radii over 2M mixture points of seed
    1000). Rows of the three are compared on the GPU after the timing."""
    import torch

    from multigpuegtsmoothing_amd import datasets as D
    from multigpuegtsmoothing_amd import egt as E

    torch.cuda.set_device(0)
    pts = D.mixture(16_000_000, seed=4)
    n = pts[0].size
    dev = tuple(torch.from_numpy(a).cuda() for a in pts)
    del pts
    iters = args.mixture_iters
    modes = [m for m in args.modes.split(",") if m]
    model = E.RadiusModel(seed=17)
    gt = E.GridStructure(D.mixture(2_000_000, seed=1000), 1, 1e-6, 1024, 4)
    t0 = time.perf_counter()
    mse = model.train(gt, k, epochs=6, batch=4096, max_samples=2_000_000, lr=2e-3, quantile=0.9)
    train_s = time.perf_counter() - t0
    gt.close()
    g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
    t0 = time.perf_counter()
    lvl = int(g.autotune(k)[0])
    # launch shape (lane per query / persistent / wave per query) for the tuned level and for the learned
    # starts, each timed on the tuner's sample, outside the timed steps
    g.set_option(E.OPT_START_LEVEL, lvl)
    shapes = {"plain_own_cell": (0, None), "plain_autotuned": g.autotune_shape(k)}
    g.close()
    # the learned path: level offset, the proposal's level alone or with its radius bound, and the shape,
    # tuned together by the library's learned-start tuner
    g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
    model.propose(g, k)
    level_offset, learned_mode, lshape, lms = g.autotune_learned(k)
    shapes["learned"] = (lshape, None)
    tune_s = time.perf_counter() - t0
    g.close()
    holder = {}

    def make_step(mode):
        shape = shapes[mode][0]

        def step():
            g = E.GridStructure(dev, 1, 1e-6, 1024, 4)
            if args.no_graph:
                g.set_option(E.OPT_TAUBIN_GRAPH, 0)
            if mode == "learned":
                model.propose(g, k)
                g.set_option(E.OPT_LEARNED_MODE, learned_mode)
                g.set_option(E.OPT_LEARNED_LEVEL_OFFSET, level_offset)
            elif mode == "plain_autotuned":
                g.set_option(E.OPT_START_LEVEL, lvl)
            g.set_option(E.OPT_KNN_PERSISTENT, 1 if shape == 1 else 0)
            g.set_option(E.OPT_KNN_LANES, 64 if shape == 2 else 1)
            g.GRIDCUDAKNNSELF_COMPACT(k, 1e8)
            E.EGT_CUDA(g).PerformSmoothing(iters, 0.5, -0.53, 0.65)
            st = g.stats()
            if holder.get("keep"):
                holder["rows"] = _device_rows(torch, g.GetKNNIndexesDevice(), n * k)
            g.close()
            return st
        return step

    per, rows = {}, {}
    for mode in modes:
        dt, stats = timed_steps(torch, make_step(mode), steps, warmup)
        print(f"mixture k={k} {mode}: {dt * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
        holder["keep"] = True  # one more, untimed step keeps the EGG-filtered rows for the comparison
        make_step(mode)()
        holder["keep"] = False
        rows[mode] = holder.pop("rows")
        st = stats[-1]
        knn_ms = st["knn_ms"] + st["propose_ms"]
        per[mode] = {"ms_per_step": dt * 1e3, "points_per_s": n / dt, "knn_ms": knn_ms,
                     "knn_queries_per_s": n / (knn_ms * 1e-3), "propose_ms": st["propose_ms"],
                     "egg_ms": st["egg_ms"], "taubin_ms": st["taubin_ms"], "grid_ms": st["grid_ms"],
                     "taubin_pass_ms": st["taubin_pass_ms"], "taubin_tile": st["taubin_tile"],
                     "egg_kept": st["egg_kept"],
                     "candidates_per_query": st["knn_candidates"] / n, "cells_per_query": st["knn_cells"] / n,
                     "insertions_per_query": st["knn_insertions"] / n, "invalid_rows": st["invalid_rows"],
                     "knn_shape": ("lane per query", "persistent, lane per query", "wave per query")[shapes[mode][0]]}
        if shapes[mode][1] is not None:
            per[mode]["shape_tuner_ms"] = [round(float(x), 4) for x in shapes[mode][1]]
        if mode == "learned":
            per[mode]["fallback_rate"] = st["learned_fallbacks"] / n
            per[mode]["bound_retry_rate"] = st["learned_bound_retries"] / n
            per[mode]["learned_mode"] = {1: "proposed level", 3: "proposed level and radius bound"}[learned_mode]
            per[mode]["learned_level_offset"] = level_offset
            per[mode]["learned_tuner_ms"] = {f"offset {o}, {u}, {sh}": round(float(lms[o, mi, si]), 3)
                                             for o in range(2) for mi, u in enumerate(("level", "level+bound"))
                                             for si, sh in enumerate(("lane", "wave"))}
    first = next(iter(rows))
    equal = bool(all(torch.equal(rows[first], rows[m]) for m in rows))
    del rows
    res = {"n": n, "k": k, "iters": iters, "modes": per, "rows_equal": equal, "start_level_autotuned": lvl,
           "autotune_seconds": tune_s,
           "train": {"cloud": "mixture 2M pts, seed 1000", "seconds": train_s, "mse": mse,
                     "level_margin": model.safety, "bound_margin": model.bound_safety}}
    plains = [m for m in ("plain_own_cell", "plain_autotuned") if m in per]
    if plains and "learned" in per:
        best_plain = min(plains, key=lambda m: per[m]["ms_per_step"])
        best_plain_knn = min(plains, key=lambda m: per[m]["knn_ms"])
        res.update(best_plain=best_plain,
                   learned_end_to_end_speedup=per[best_plain]["ms_per_step"] / per["learned"]["ms_per_step"],
                   learned_knn_speedup=per[best_plain_knn]["knn_ms"] / per["learned"]["knn_ms"])
    model.close()
    del dev
    torch.cuda.empty_cache()
    E.release_cached_memory()
    return res


def cpu_baseline_mixture(k=40, iters=150, sample=1_000_000, taubin_iters=2):
    """The oracle (host C++/OpenMP restatement) timed on a bounded sample of the headline workload: the grid
    build of all 16M points; kNN + EGG of a seeded 1M-query subset, searched against the whole cloud; and
    `taubin_iters` Taubin iterations over a 2M-point mixture of the same generator (Taubin needs every row, and
    only the subset's rows exist; that cloud's own kNN and EGG are not timed), scaled per point-iteration.
    points/s = 1 / (grid + kNN + EGG per point + 150 x Taubin per point-iteration)."""
    from multigpuegtsmoothing_amd import datasets as D
    from oracle import oracle as O

    pts = D.mixture(16_000_000, seed=4)
    t0 = time.perf_counter()
    og = O.Grid(*pts)
    t_grid = time.perf_counter() - t0
    n = og.n
    sel = np.sort(np.random.default_rng(1234).choice(n, sample, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    rows, _, _, _ = og.knn_subset(k, sel)
    t_knn = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.egg_subset(og.xs, og.ys, og.zs, sel, rows, 0.65)
    t_egg = time.perf_counter() - t0
    del rows, og, pts
    small = O.Grid(*D.mixture(2_000_000, seed=4))
    krows, _, _, _ = small.knn(k)
    erows = O.egg(small.xs, small.ys, small.zs, krows, 0.65)
    del krows
    t0 = time.perf_counter()
    O.taubin_f32(small.xs, small.ys, small.zs, erows, taubin_iters, 0.5, -0.53)
    t_taub = time.perf_counter() - t0
    per_point = t_grid / n + (t_knn + t_egg) / sample + iters * t_taub / (taubin_iters * small.n)
    return {"value": 1.0 / per_point, "unit": "points/s (end-to-end grid+kNN+EGG+Taubin)", "cores": O.num_threads(),
            "kind": "port",
            "sample": f"oracle on the headline cloud: grid build of all {n} points {t_grid:.2f} s; kNN + EGG of a "
                      f"seeded {sample}-query subset against the whole cloud {t_knn:.2f} + {t_egg:.2f} s; Taubin "
                      f"{taubin_iters} iterations over a 2M-point mixture of the same generator {t_taub:.2f} s, "
                      f"scaled to {iters} iterations per point",
            "stage_s_per_million_points": {"grid": t_grid / n * 1e6, "knn": t_knn / sample * 1e6,
                                           "egg": t_egg / sample * 1e6,
                                           "taubin_150_its": iters * t_taub / (taubin_iters * small.n) * 1e6}}


def headline_main(args):
    steps, warmup = args.steps, args.warmup
    res = mixture_modes(args, 40, steps, warmup)
    n, k, iters = res["n"], res["k"], res["iters"]
    main_mode = "learned" if "learned" in res["modes"] else next(iter(res["modes"]))
    L = res["modes"][main_mode]
    # Taubin pass (dominant kernel), algorithmic bytes per launch: own xyz read 12 + write 12 + row ptr 4 per
    # point, 4 B per kept neighbour index; kNN: query xyz 12 + 4k output per query; grid: 32 B per point
    taub_bytes = 28 * n + 4 * L["egg_kept"]
    ach = taub_bytes / (L["taubin_pass_ms"] * 1e-3) / 1e9
    knn_gbs = n * (12 + 4 * k) / (L["knn_ms"] * 1e-3) / 1e9
    grid_gbs = n * 32 / (L["grid_ms"] * 1e-3) / 1e9
    roof = {"kernel": f"k_taubin_tile<{L['taubin_tile']}>", "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": measured_traffic("mixture"),
            "bytes_per_launch": taub_bytes, "avg_launch_ms": L["taubin_pass_ms"],
            "knn": {"kernel": ("k_knn_wave" if L.get("knn_shape") == "wave per query" else "k_knn<40>")
                              + (" (+ proposal)" if main_mode == "learned" else ""),
                    "shape": L.get("knn_shape"), "bound": "valu/latency", "achieved_GBs": knn_gbs,
                    "frac": knn_gbs / HBM_PEAK_GBS, "bytes_per_query": 12 + 4 * k},
            "grid": {"kernels": "grid build (cube, key, sorts, gathers, level tables)", "bound": "hbm",
                     "achieved_GBs": grid_gbs, "frac": grid_gbs / HBM_PEAK_GBS, "bytes_per_point": 32}}
    out = {"metric": METRIC, "value": n / (L["ms_per_step"] * 1e-3),
           "unit": f"points/s (end-to-end grid+proposal+kNN+EGG+Taubin, {main_mode} path)", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": L["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"mixture {n} pts (12 anisotropic clusters + 4 planes, seed 4), k={k}, alpha 0.65, "
                                  f"{iters} Taubin its (lambda 0.5, mu -0.53), "
                                  + ("learned radius proposal" if main_mode == "learned" else main_mode),
                      "n_points": n, "k": k, "iterations": iters, "path": main_mode},
           "knn_queries_per_s": L["knn_queries_per_s"],
           "taubin_point_iters_per_s": n * iters / (L["taubin_ms"] * 1e-3),
           "learned_end_to_end_speedup": res.get("learned_end_to_end_speedup"),
           "learned_knn_speedup": res.get("learned_knn_speedup"), "best_plain": res.get("best_plain"),
           "rows_equal": res["rows_equal"], "mixture": res, "roofline": roof}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_mixture(k, iters)
    if not args.no_extras:
        # the same cloud at k=64 (the metric's second neighbour count), fewer steps
        r64 = mixture_modes(args, 64, min(steps, 3), min(warmup, 1))
        out["mixture_k64"] = {"ms_per_step": {m: v["ms_per_step"] for m, v in r64["modes"].items()},
                              "knn_ms": {m: v["knn_ms"] for m, v in r64["modes"].items()},
                              "knn_shape": {m: v["knn_shape"] for m, v in r64["modes"].items()},
                              "learned_end_to_end_speedup": r64.get("learned_end_to_end_speedup"),
                              "learned_knn_speedup": r64.get("learned_knn_speedup"),
                              "best_plain": r64.get("best_plain"), "rows_equal": r64["rows_equal"],
                              "start_level_autotuned": r64["start_level_autotuned"]}
        out["batch_512_clouds"] = batch_result(args, min(steps, 3), min(warmup, 1),
                                               per_cloud=not args.no_per_cloud)
        for key in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"):
            out["batch_512_clouds"].pop(key, None)
        targs = argparse.Namespace(**vars(args))
        targs.steps, targs.warmup = min(steps, 3), min(warmup, 1)
        targs.no_cpu_baseline, targs.no_query = True, True
        out["torus_4M"] = torus_result(targs)
        for key in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data"):
            out["torus_4M"].pop(key, None)
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)
    if args.workload == "headline":
        headline_main(args)
    elif args.workload == "batch":
        batch_main(args)
    elif args.workload == "query":
        query_main(args)
    elif args.workload == "mesh":
        mesh_main(args)
    elif args.workload == "mixture":
        mixture_main(args)
    else:
        print(json.dumps(torus_result(args)), flush=True)


if __name__ == "__main__":
    main()
