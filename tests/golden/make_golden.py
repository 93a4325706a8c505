"""Writes the golden fixtures in this directory from the oracle (host C++ restatement of the reference).

The reference ships no tests or vectors of its own, so these fixtures are oracle outputs on small seeded
inputs. The oracle itself is pinned by tests/test_oracle_*.py. The inputs are stored with the outputs, so
the fixtures do not depend on the generators. Run from the repository root:

    python tests/golden/make_golden.py

Each case is one compressed .npz (no pickles). manifest.json records the parameters and sha256 of each file.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from multigpuegtsmoothing_amd import datasets as D  # noqa: E402
from oracle import oracle as O  # noqa: E402

LAM, MU, ALPHA = 0.5, -0.53, 0.65

CASES = {
    "cloud_sphere64_k8": dict(gen="sphere", n=64, seed=101, k=8),
    "cloud_torus1000_k16": dict(gen="torus", n=1000, seed=102, k=16),
    "cloud_mixture2000_k40": dict(gen="mixture", n=2000, seed=103, k=40),
    "mesh_icosphere2_k16": dict(gen="icosphere", levels=2, seed=104, k=16),
}


def inputs(spec):
    if spec["gen"] == "icosphere":
        return D.icosphere(spec["levels"], noise=0.01, seed=spec["seed"])
    gen = getattr(D, spec["gen"])
    x, y, z = gen(spec["n"], seed=spec["seed"])
    return x, y, z, None


def outputs(x, y, z, tri, k):
    """Original-order outputs of the whole path for one input."""
    mesh = tri is not None
    alpha = 1e-8 if mesh else ALPHA
    res = {}
    g = O.Grid(x, y, z)
    if mesh:
        knn = g.mesh_adjacency(tri, k)
    else:
        knn, kth, inv, st = g.knn(k)
        res["invalid_rows"] = np.array([int(inv.sum())], np.int64)
    res["knn"] = g.to_original(knn)
    egg_rows = O.egg(g.xs, g.ys, g.zs, knn, alpha)
    res["egg"] = g.to_original(egg_rows)
    for its in (1, 10):
        sx, sy, sz = O.taubin_f32(g.xs, g.ys, g.zs, egg_rows, its, LAM, MU)
        res[f"taubin{its}_f32"] = np.stack(g.output(sx, sy, sz), 1)
    d64 = np.stack(O.taubin_f64(g.xs, g.ys, g.zs, egg_rows, 10, LAM, MU), 1)
    ext = float(g.max_extent) - float(g.min_extent)
    out64 = np.empty_like(d64)
    out64[g.perm] = d64 * ext + float(g.min_extent)
    res["taubin10_f64"] = out64
    if not mesh:
        sx, sy, sz = O.taubin_f32(g.xs, g.ys, g.zs, egg_rows, 10, LAM, MU)
        nx, ny, nz = O.pca_normals(sx, sy, sz, egg_rows)
        onx, ony, onz, ostats, w = O.orient_normals(egg_rows, nx, ny, nz)
        res["normals10"] = np.stack(O.normals_original(g, onx, ony, onz), 1)
    return res


def main():
    manifest = {"lambda": LAM, "mu": MU, "alpha_cloud": ALPHA, "alpha_mesh": 1e-8, "eps": 1e-6,
                "bucket_count": 1024, "thres": 1e8, "cases": {}}
    for name, spec in CASES.items():
        x, y, z, tri = inputs(spec)
        arrays = {"x": x, "y": y, "z": z}
        if tri is not None:
            arrays["tri"] = tri
        arrays.update(outputs(x, y, z, tri, spec["k"]))
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        manifest["cases"][name] = dict(spec, file=name + ".npz", sha256=digest, n=int(x.size))
        print(name, x.size, os.path.getsize(path), "bytes")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
