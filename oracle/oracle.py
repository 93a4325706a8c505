"""ctypes binding of the host oracle (liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only as the
checker or the timed CPU baseline. Build with ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        i64, i32, f32 = ctypes.c_int64, ctypes.c_int, ctypes.c_float
        L.orc_grid_build.argtypes = [F32P, F32P, F32P, i64, f32, i32, F32P, F32P, F32P, I32P, I32P, F32P]
        L.orc_knn_self.argtypes = [F32P, F32P, F32P, I32P, i64, f32, i32, i32, f32, i32, I32P, F32P, I32P, I64P]
        L.orc_knn_self_subset.argtypes = [F32P, F32P, F32P, I32P, i64, f32, i32, i32, f32, i32, I32P, i64, I32P,
                                          F32P, I32P, I64P]
        L.orc_egg_subset.argtypes = [F32P, F32P, F32P, I32P, i64, i32, f32, I32P]
        L.orc_knn_brute.argtypes = [F32P, F32P, F32P, i64, i32, f32, I32P]
        L.orc_knn_query.argtypes = [F32P, F32P, F32P, I32P, i64, F32P, F32P, F32P, i64, f32, f32, f32, i32, i32,
                                    f32, i32, I32P, F32P, I32P, I64P]
        L.orc_knn_query_brute.argtypes = [F32P, F32P, F32P, i64, F32P, F32P, F32P, i64, f32, f32, i32, f32, I32P]
        L.orc_egg.argtypes = [F32P, F32P, F32P, i64, i32, f32, I32P]
        L.orc_taubin_f32.argtypes = [F32P, F32P, F32P, I32P, i64, i32, i32, f32, f32, i32]
        L.orc_taubin_f64.argtypes = [F64P, F64P, F64P, I32P, i64, i32, i32, ctypes.c_double, ctypes.c_double, i32]
        L.orc_output.argtypes = [F32P, F32P, F32P, I32P, i64, f32, f32, F32P, F32P, F32P]
        L.orc_knn_to_original.argtypes = [I32P, I32P, i64, i32, I32P]
        L.orc_mesh_adjacency.argtypes = [F32P, F32P, F32P, i64, U32P, i64, I32P, i32, I32P]
        L.orc_pca_normals.argtypes = [F32P, F32P, F32P, I32P, i64, i32, F32P, F32P, F32P]
        L.orc_orient_normals.argtypes = [I32P, i64, i32, F32P, F32P, F32P, I64P,
                                         ctypes.POINTER(ctypes.c_double)]
        L.orc_num_threads.restype = ctypes.c_int
        L.egt_expf.argtypes = [f32]
        L.egt_expf.restype = f32
        L.orc_eig3.argtypes = [F32P, F32P, F32P]
        _LIB = L
    return _LIB


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Grid:
    """Sorted-space grid (GridStructure's constructor state): normalized sorted points, perm, keys."""

    def __init__(self, x, y, z, eps=1e-6, B=1024):
        x, y, z = _f32(x), _f32(y), _f32(z)
        n = x.size
        self.n, self.eps, self.B = n, float(np.float32(eps)), B
        self.xs = np.empty(n, np.float32)
        self.ys = np.empty(n, np.float32)
        self.zs = np.empty(n, np.float32)
        self.perm = np.empty(n, np.int32)
        self.key = np.empty(n, np.int32)
        ext = np.empty(2, np.float32)
        rc = lib().orc_grid_build(x, y, z, n, np.float32(eps), B, self.xs, self.ys, self.zs, self.perm, self.key, ext)
        if rc != 0:
            raise ValueError("orc_grid_build failed")
        self.min_extent, self.max_extent = float(ext[0]), float(ext[1])
        self.inv_perm = np.empty(n, np.int32)
        self.inv_perm[self.perm] = np.arange(n, dtype=np.int32)

    def thres_n(self, thres=1e8):
        t = np.float32(thres) / np.float32(np.float32(self.max_extent) - np.float32(self.min_extent))
        return np.float32(t * t)

    def knn(self, k, thres=1e8, maxc=100000):
        n = self.n
        knn = np.empty(n * k, np.int32)
        kth = np.empty(n, np.float32)
        inv = np.empty(n, np.int32)
        stats = np.zeros(8, np.int64)
        rc = lib().orc_knn_self(self.xs, self.ys, self.zs, self.key, n, np.float32(self.eps), self.B, k,
                                self.thres_n(thres), maxc, knn, kth, inv, stats)
        if rc != 0:
            raise ValueError("orc_knn_self failed")
        return knn.reshape(n, k), kth, inv, stats

    def knn_subset(self, k, qsel, thres=1e8, maxc=100000):
        """Rows of the sorted-space queries qsel only (row t is query qsel[t]), as knn() computes them."""
        qsel = np.ascontiguousarray(qsel, np.int32)
        m = qsel.size
        knn = np.empty(max(m, 1) * k, np.int32)
        kth = np.empty(max(m, 1), np.float32)
        inv = np.empty(max(m, 1), np.int32)
        stats = np.zeros(8, np.int64)
        rc = lib().orc_knn_self_subset(self.xs, self.ys, self.zs, self.key, self.n, np.float32(self.eps), self.B, k,
                                       self.thres_n(thres), maxc, qsel, m, knn, kth, inv, stats)
        if rc != 0:
            raise ValueError("orc_knn_self_subset failed")
        return knn[:m * k].reshape(m, k), kth[:m], inv[:m], stats

    def knn_brute(self, k, thres=1e8):
        knn = np.empty(self.n * k, np.int32)
        lib().orc_knn_brute(self.xs, self.ys, self.zs, self.n, k, self.thres_n(thres), knn)
        return knn.reshape(self.n, k)

    def knn_query(self, qx, qy, qz, k, thres=1e8, maxc=100000):
        """Query-vs-reference rows (GRIDCUDAKNNCompact) in original query order, reference sorted indices."""
        qx, qy, qz = _f32(qx), _f32(qy), _f32(qz)
        nq = qx.size
        knn = np.empty(max(nq, 1) * k, np.int32)
        kth = np.empty(max(nq, 1), np.float32)
        inv = np.empty(max(nq, 1), np.int32)
        stats = np.zeros(8, np.int64)
        rc = lib().orc_knn_query(self.xs, self.ys, self.zs, self.key, self.n, qx, qy, qz, nq,
                                 np.float32(self.min_extent), np.float32(self.max_extent), np.float32(self.eps),
                                 self.B, k, self.thres_n(thres), maxc, knn, kth, inv, stats)
        if rc != 0:
            raise ValueError("orc_knn_query failed")
        return knn[:nq * k].reshape(nq, k), kth[:nq], inv[:nq], stats

    def knn_query_brute(self, qx, qy, qz, k, thres=1e8):
        qx, qy, qz = _f32(qx), _f32(qy), _f32(qz)
        nq = qx.size
        knn = np.empty(max(nq, 1) * k, np.int32)
        lib().orc_knn_query_brute(self.xs, self.ys, self.zs, self.n, qx, qy, qz, nq, np.float32(self.min_extent),
                                  np.float32(self.max_extent), k, self.thres_n(thres), knn)
        return knn[:nq * k].reshape(nq, k)

    def ref_to_original(self, rows):
        """Maps reference sorted indices in query rows to original reference indices (-1 kept)."""
        rows = np.asarray(rows, np.int32)
        return np.where(rows >= 0, self.perm[np.maximum(rows, 0)], -1).astype(np.int32)

    def to_original(self, knn):
        n, k = knn.shape
        out = np.empty(n * k, np.int32)
        lib().orc_knn_to_original(np.ascontiguousarray(knn, np.int32).ravel(), self.perm, n, k, out)
        return out.reshape(n, k)

    def output(self, xs, ys, zs):
        ox = np.empty(self.n, np.float32)
        oy = np.empty(self.n, np.float32)
        oz = np.empty(self.n, np.float32)
        lib().orc_output(_f32(xs), _f32(ys), _f32(zs), self.perm, self.n, np.float32(self.min_extent),
                         np.float32(self.max_extent), ox, oy, oz)
        return ox, oy, oz

    def mesh_adjacency(self, tri, k):
        tri = np.ascontiguousarray(tri, dtype=np.uint32).ravel()
        knn = np.empty(self.n * k, np.int32)
        lib().orc_mesh_adjacency(self.xs, self.ys, self.zs, self.n, tri, tri.size // 3, self.inv_perm, k, knn)
        return knn.reshape(self.n, k)


def egg(xs, ys, zs, knn, alpha):
    n, k = knn.shape
    out = np.ascontiguousarray(knn, np.int32).copy()
    lib().orc_egg(_f32(xs), _f32(ys), _f32(zs), n, k, np.float32(alpha), out.ravel())
    return out


def egg_subset(xs, ys, zs, qsel, rows, alpha):
    """EGG filter of the rows of the points qsel (row t is point qsel[t]'s), as egg() filters them."""
    qsel = np.ascontiguousarray(qsel, np.int32)
    m, k = rows.shape
    out = np.ascontiguousarray(rows, np.int32).copy()
    lib().orc_egg_subset(_f32(xs), _f32(ys), _f32(zs), qsel, m, k, np.float32(alpha), out.ravel())
    return out


def taubin_f32(xs, ys, zs, knn, iters, lam, mu, regularized=False):
    n, k = knn.shape
    x, y, z = _f32(xs).copy(), _f32(ys).copy(), _f32(zs).copy()
    lib().orc_taubin_f32(x, y, z, np.ascontiguousarray(knn, np.int32).ravel(), n, k, iters, np.float32(lam),
                         np.float32(mu), int(regularized))
    return x, y, z


def taubin_f64(xs, ys, zs, knn, iters, lam, mu, regularized=False):
    n, k = knn.shape
    x = np.ascontiguousarray(xs, np.float64).copy()
    y = np.ascontiguousarray(ys, np.float64).copy()
    z = np.ascontiguousarray(zs, np.float64).copy()
    lib().orc_taubin_f64(x, y, z, np.ascontiguousarray(knn, np.int32).ravel(), n, k, iters, float(lam), float(mu),
                         int(regularized))
    return x, y, z


def pca_normals(xs, ys, zs, knn):
    n, k = knn.shape
    nx = np.empty(n, np.float32)
    ny = np.empty(n, np.float32)
    nz = np.empty(n, np.float32)
    lib().orc_pca_normals(_f32(xs), _f32(ys), _f32(zs), np.ascontiguousarray(knn, np.int32).ravel(), n, k,
                          nx, ny, nz)
    return nx, ny, nz


def orient_normals(knn, nx, ny, nz):
    n, k = knn.shape
    nx, ny, nz = _f32(nx).copy(), _f32(ny).copy(), _f32(nz).copy()
    stats = np.zeros(8, np.int64)
    w = ctypes.c_double(0.0)
    lib().orc_orient_normals(np.ascontiguousarray(knn, np.int32).ravel(), n, k, nx, ny, nz, stats, ctypes.byref(w))
    return nx, ny, nz, stats, w.value


def num_threads():
    return int(lib().orc_num_threads())


def expf(x):
    return float(lib().egt_expf(np.float32(x)))


def eig3(a6):
    ev = np.empty(3, np.float32)
    V = np.empty(9, np.float32)
    idx = lib().orc_eig3(np.ascontiguousarray(a6, np.float32), ev, V)
    return idx, ev, V.reshape(3, 3)  # V[c] = eigenvector c


def normals_original(grid, nx, ny, nz):
    """Scatter sorted-space normals to original order (Normal_CUDA.cu:751-759)."""
    ox = np.empty(grid.n, np.float32)
    oy = np.empty(grid.n, np.float32)
    oz = np.empty(grid.n, np.float32)
    ox[grid.perm] = nx
    oy[grid.perm] = ny
    oz[grid.perm] = nz
    return ox, oy, oz


def pipeline(x, y, z, k, iters, lam, mu, alpha, normals=True, mesh_tri=None, eps=1e-6, B=1024, thres=1e8):
    """Full reference path on the CPU: provider -> EGG -> Taubin -> (normals) -> output (main.cpp:106-134)."""
    g = Grid(x, y, z, eps=eps, B=B)
    if mesh_tri is None:
        knn, kth, inv, stats = g.knn(k, thres=thres)
    else:
        knn = g.mesh_adjacency(mesh_tri, k)
        stats = None
    egg_rows = egg(g.xs, g.ys, g.zs, knn, alpha)
    sx, sy, sz = taubin_f32(g.xs, g.ys, g.zs, egg_rows, iters, lam, mu)
    res = {"grid": g, "knn": knn, "egg": egg_rows, "smoothed_sorted": (sx, sy, sz), "knn_stats": stats}
    if normals:
        nx, ny, nz = pca_normals(sx, sy, sz, egg_rows)
        res["pca_sorted"] = (nx, ny, nz)
        onx, ony, onz, ostats, w = orient_normals(egg_rows, nx, ny, nz)
        res["normals_sorted"] = (onx, ony, onz)
        res["orient_stats"] = ostats
        res["msf_weight"] = w
        res["normals"] = normals_original(g, onx, ony, onz)
    res["output"] = g.output(sx, sy, sz)
    return res
