"""Python mirror of the reference's provider / consumer classes over the gfx950 C ABI (include/egt.h).

Class and method names follow the reference so call sites read alike:

    GridStructure(points, num_cards=1, eps=1e-6, bucketCount=1024, maxDivision=4)   GridKNNCuda.cuh:15-16
        .GRIDCUDAKNNSELF_COMPACT(k, thres)                                            GridKNNCuda.cuh:22
    GridMeshCuda(points, triangles, num_cards=1, eps=1e-6, bucketCount=1024)        GridMeshCuda1.cuh:16
        .GetNeighborsFromTriangles(k)                                                 GridMeshCuda1.cuh:17
    EGT_CUDA(provider).PerformSmoothing(iterations, lambda_=0.63, mu=-0.64, alpha=1e-8)  EGT_CUDA.cuh:13-14
    Normal_CUDA(provider).GetNormals()                                                Normal_CUDA.cuh:12-13
    getOutputPoints(provider)                                                         getOutputPoints.cuh:10

Errors raise EGTError carrying the library's status code and message. The library is loaded from
multigpuegtsmoothing_amd/libegt.so; importing this module without it raises immediately (there is no CPU
fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EGT_LIB_PATH") or os.path.join(_HERE, "libegt.so")

EGT_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "HIP_ERROR", 4: "UNSUPPORTED", 5: "IO_ERROR", 6: "STATE"}
OPT_START_LEVEL, OPT_LEARNED, OPT_TAUBIN_GRAPH, OPT_MAX_PTS_IN_CUBE, OPT_KNN_PERSISTENT, OPT_TAUBIN_TILE = 0, 1, 2, 3, 4, 5
OPT_TAUBIN_REGULARIZED = 6
OPT_LEARNED_MODE = 7         # learned proposal: 1 = start level (default), 2 = radius bound, 3 = both
OPT_KNN_LANES = 8            # kNN self search: 1 = one lane per query (default), 64 = one wave per query
OPT_KNN_LONG_QUERY = 9       # lane-per-query search: candidates after which a query goes to a wave (2048; 0 = never)
OPT_KNN_CLASS_SHAPES = 10    # per neighbour-count class launch shapes, as BatchGrid.autotune_shape returns them
OPT_LEARNED_LEVEL_OFFSET = 11  # added to the proposed start levels (-3..3)
START_LEVEL_OCCUPANCY = 127  # OPT_START_LEVEL value: one level from the mean cell occupancy
SUB_LEVELS = 3               # start levels -1, -2, -3: boxes of cells 1/2, 1/4, 1/8 of a grid cell
AUTOTUNE_SLOTS = 14          # autotuner trial arrays: levels -3 .. 10


class EGTError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"egt status {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [("n_points", ctypes.c_int64), ("k", ctypes.c_int32), ("bucket_count", ctypes.c_int32),
                ("invalid_rows", ctypes.c_int64), ("knn_candidates", ctypes.c_int64), ("egg_kept", ctypes.c_int64),
                ("learned_queries", ctypes.c_int64), ("learned_fallbacks", ctypes.c_int64),
                ("msf_edges", ctypes.c_int64), ("msf_trees", ctypes.c_int64), ("msf_weight", ctypes.c_double),
                ("knn_ms", ctypes.c_float), ("egg_ms", ctypes.c_float), ("taubin_ms", ctypes.c_float),
                ("normals_ms", ctypes.c_float), ("grid_ms", ctypes.c_float), ("propose_ms", ctypes.c_float),
                ("knn_cells", ctypes.c_int64), ("taubin_tile_ms", ctypes.c_float * 2),
                ("taubin_tile", ctypes.c_int32), ("taubin_fallback_tiles", ctypes.c_int32),
                ("query_count", ctypes.c_int64), ("query_candidates", ctypes.c_int64),
                ("query_invalid_rows", ctypes.c_int64), ("query_ms", ctypes.c_float), ("query_k", ctypes.c_int32),
                ("taubin_pass_ms", ctypes.c_float), ("taubin_setup_ms", ctypes.c_float),
                ("taubin_overflow_tiles", ctypes.c_int32), ("learned_bound_retries", ctypes.c_int64),
                ("knn_insertions", ctypes.c_int64), ("knn_max_candidates", ctypes.c_int64),
                ("knn_long_queries", ctypes.c_int64)]

    def as_dict(self):
        d = {f[0]: getattr(self, f[0]) for f in self._fields_}
        d["taubin_tile_ms"] = list(d["taubin_tile_ms"])
        return d


_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_H = ctypes.c_void_p


class _OptionalF32:
    """ctypes argument type: a C-contiguous float32 array or None (NULL)."""

    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        if not (isinstance(obj, np.ndarray) and obj.dtype == np.float32 and obj.flags["C_CONTIGUOUS"]):
            raise TypeError("expected a C-contiguous float32 array or None")
        return obj.ctypes.data_as(ctypes.c_void_p)


_F32N = _OptionalF32
_PH = ctypes.POINTER(ctypes.c_void_p)
_PF = ctypes.POINTER(ctypes.c_void_p)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C "
                          "multigpuegtsmoothing_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    i64, i32, f32, st = ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_int
    sig = {
        "egt_last_error_string": ([], ctypes.c_char_p),
        "egt_grid_create": ([_F32P, _F32P, _F32P, i64, i32, f32, i32, i32, _PH], st),
        "egt_grid_create_device": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, i64, i32, f32, i32, i32, _PH], st),
        "egt_grid_knn_self": ([_H, i32, f32], st),
        "egt_grid_knn_query": ([_H, _F32P, _F32P, _F32P, i64, i32, f32], st),
        "egt_grid_knn_query_device": ([_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, i64, i32, f32], st),
        "egt_grid_knn_query_device_noalloc": ([_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, i64, i32, f32,
                                               ctypes.c_void_p], st),
        "egt_get_knn_query_original": ([_H, _I32P], st),
        "egt_get_device_query_knn": ([_H, _PF, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int32)], st),
        "egt_mesh_create": ([_F32P, _F32P, _F32P, i64, _U32P, i64, i32, f32, i32, _PH], st),
        "egt_mesh_neighbors": ([_H, i32], st),
        "egt_smooth": ([_H, i32, f32, f32, f32], st),
        "egt_normals": ([_H, _F32P, _F32P, _F32P], st),
        "egt_get_output_points": ([_H, _F32P, _F32P, _F32P], st),
        "egt_get_knn_original": ([_H, _I32P], st),
        "egt_get_device_points": ([_H, _PF, _PF, _PF], st),
        "egt_get_device_knn": ([_H, _PF], st),
        "egt_get_device_original_indexes": ([_H, _PF], st),
        "egt_get_extent": ([_H, ctypes.POINTER(f32), ctypes.POINTER(f32)], st),
        "egt_get_counts": ([_H, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                            ctypes.POINTER(i64)], st),
        "egt_get_stats": ([_H, ctypes.POINTER(Stats)], st),
        "egt_set_option": ([_H, i32, i64], st),
        "egt_release_cached_memory": ([ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)], st),
        "egt_destroy": ([_H], st),
        "egt_get_stream": ([_H], ctypes.c_void_p),
        "egt_autotune": ([_H, i32, f32, ctypes.POINTER(ctypes.c_int32), _F32N, _F32N], st),
        "egt_autotune_shape": ([_H, i32, f32, ctypes.POINTER(ctypes.c_int32), _F32N], st),
        "egt_batch_autotune_shape": ([_H, _I32P, f32, ctypes.POINTER(i64), _F32N], st),
        "egt_autotune_learned": ([_H, i32, f32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int32), _F32N], st),
        "egt_batch_create": ([_F32P, _F32P, _F32P, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"),
                              ctypes.c_int32, f32, i32, _PH], st),
        "egt_batch_create_device": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS"), ctypes.c_int32, f32, i32,
                                     _PH], st),
        "egt_batch_knn": ([_H, _I32P, f32], st),
        "egt_batch_get_knn_original": ([_H, ctypes.c_int32, _I32P], st),
        "egt_batch_cloud_info": ([_H, ctypes.c_int32, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(f32), ctypes.POINTER(f32)], st),
        "egt_model_param_count": ([], ctypes.c_int32),
        "egt_model_feature_count": ([], ctypes.c_int32),
        "egt_model_create": ([ctypes.c_uint64, _PH], st),
        "egt_model_destroy": ([_H], st),
        "egt_model_get_params": ([_H, _F32P, i64], st),
        "egt_model_set_params": ([_H, _F32P, i64], st),
        "egt_model_eval": ([_H, _F32P, _F32P, i64, _F32N, _F32N, ctypes.POINTER(f32)], st),
        "egt_model_adam_step": ([_H, _F32P, _F32P, i64, f32, _F32N, ctypes.POINTER(f32)], st),
        "egt_model_train": ([_H, _H, i32, i32, i32, i64, f32, f32, ctypes.POINTER(f32)], st),
        "egt_model_get_safety": ([_H, ctypes.POINTER(f32)], st),
        "egt_model_set_safety": ([_H, f32], st),
        "egt_model_get_bound_safety": ([_H, ctypes.POINTER(f32)], st),
        "egt_model_set_bound_safety": ([_H, f32], st),
        "egt_model_propose": ([_H, _H, i32], st),
        "egt_model_train_batch": ([_H, _H, _I32P, i32, i32, i64, f32, f32, ctypes.POINTER(f32)], st),
        "egt_model_propose_batch": ([_H, _H, _I32P], st),
        "egt_model_features": ([_H, i32, _F32N, _F32N, ctypes.POINTER(ctypes.c_int32)], st),
        "egt_model_levels": ([_H, np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")], st),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


_lib = _load()


def lib():
    return _lib


def _check(status):
    if status != EGT_OK:
        raise EGTError(status, _lib.egt_last_error_string().decode())


def _xyz(points):
    if isinstance(points, (tuple, list)) and len(points) == 3:
        x, y, z = points
    else:
        p = np.asarray(points, dtype=np.float32)
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return (np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
            np.ascontiguousarray(z, np.float32))


class _Provider:
    """KNNInterface (KNNInterface.cuh:8-31) over one handle."""

    def __init__(self):
        self._h = None

    @property
    def handle(self):
        if self._h is None:
            raise EGTError(6, "handle destroyed")
        return self._h

    def close(self):
        if self._h is not None:
            _lib.egt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, opt, value):
        _check(_lib.egt_set_option(self.handle, opt, int(value)))

    # KNNInterface accessors
    def GetNumberOfCards(self):
        return 1

    def pointsRefCount(self):
        n = ctypes.c_int64()
        _check(_lib.egt_get_counts(self.handle, ctypes.byref(n), None, None, None))
        return n.value

    def NeighborCount(self):
        k = ctypes.c_int32()
        _check(_lib.egt_get_counts(self.handle, None, ctypes.byref(k), None, None))
        return k.value

    def GetPartitionSize(self):
        return self.pointsRefCount()

    def GetMinExtent(self):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        _check(_lib.egt_get_extent(self.handle, ctypes.byref(mn), ctypes.byref(mx)))
        return mn.value

    def GetMaxExtent(self):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        _check(_lib.egt_get_extent(self.handle, ctypes.byref(mn), ctypes.byref(mx)))
        return mx.value

    def GetRefPointsDevice(self):
        """Device pointers (int addresses) of the sorted, normalized x, y, z arrays (GetRefPointsX/Y/Z)."""
        px, py, pz = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _check(_lib.egt_get_device_points(self.handle, ctypes.byref(px), ctypes.byref(py), ctypes.byref(pz)))
        return px.value, py.value, pz.value

    def GetKNNIndexesDevice(self):
        p = ctypes.c_void_p()
        _check(_lib.egt_get_device_knn(self.handle, ctypes.byref(p)))
        return p.value

    def GetOriginalRefIndexesDevice(self):
        p = ctypes.c_void_p()
        _check(_lib.egt_get_device_original_indexes(self.handle, ctypes.byref(p)))
        return p.value

    def GetKNNIndexes(self):
        """Neighbour rows in original index space, (n, k) int32 (GridKNNCuda.cu:1941-1968)."""
        n, k = self.pointsRefCount(), self.NeighborCount()
        out = np.empty(n * k, np.int32)
        _check(_lib.egt_get_knn_original(self.handle, out))
        return out.reshape(n, k)

    def stats(self):
        st = Stats()
        _check(_lib.egt_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def stream(self):
        return _lib.egt_get_stream(self.handle)


class GridStructure(_Provider):
    """Point-cloud provider (GridKNNCuda.cuh:15-16). points: (n, 3) array or (x, y, z); a torch CUDA tensor
    triple is taken as device input."""

    def __init__(self, points, num_cards=1, eps=1e-6, bucketCount=1024, maxDivision=4):
        super().__init__()
        h = ctypes.c_void_p()
        if isinstance(points, (tuple, list)) and len(points) == 3 and hasattr(points[0], "data_ptr"):
            x, y, z = points
            n = x.numel()
            self._keep = (x, y, z)
            _check(_lib.egt_grid_create_device(x.data_ptr(), y.data_ptr(), z.data_ptr(), n, num_cards, eps,
                                               bucketCount, maxDivision, ctypes.byref(h)))
        else:
            x, y, z = _xyz(points)
            _check(_lib.egt_grid_create(x, y, z, x.size, num_cards, eps, bucketCount, maxDivision, ctypes.byref(h)))
        self._h = h

    def autotune(self, k, thres=1e8):
        """Picks the start level with the least measured work per query for this cloud and k, on a sample of
        queries (no reference counterpart). Returns (level, per-level sample times in ms, per-level scores):
        entry i of the arrays is level i - 3, -1 where untried."""
        lvl = ctypes.c_int32()
        ms = np.empty(AUTOTUNE_SLOTS, np.float32)
        cost = np.empty(AUTOTUNE_SLOTS, np.float32)
        _check(_lib.egt_autotune(self.handle, k, thres, ctypes.byref(lvl), ms, cost))
        return lvl.value, ms, cost

    def autotune_shape(self, k, thres=1e8):
        """Times the current search (start level, learned proposal) in each kNN launch shape over all queries
        and keeps the fastest (no reference counterpart). Returns (shape, ms per shape): 0 = one lane per
        query, 1 = the same as one persistent launch, 2 = one wave per query."""
        shape = ctypes.c_int32()
        ms = np.empty(3, np.float32)
        _check(_lib.egt_autotune_shape(self.handle, k, thres, ctypes.byref(shape), ms))
        return shape.value, ms

    def autotune_learned(self, k, thres=1e8):
        """With a proposal in place: times level offsets (0, +1) x proposal use (level; level and radius
        bound) x shapes (lane, wave per query) over the full search and keeps the fastest (no reference
        counterpart). Returns (offset, learned mode, shape, ms array [offset, mode 1/3, lane/wave])."""
        off, mode, shape = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        ms = np.empty(8, np.float32)
        _check(_lib.egt_autotune_learned(self.handle, k, thres, ctypes.byref(off), ctypes.byref(mode),
                                         ctypes.byref(shape), ms))
        return off.value, mode.value, shape.value, ms.reshape(2, 2, 2)

    def GRIDCUDAKNNSELF_COMPACT(self, k, thres=1e8):
        _check(_lib.egt_grid_knn_self(self.handle, int(k), float(thres)))
        self._last_query = False

    def GRIDCUDAKNNCompact(self, queries, k, thres=1e8):
        """k nearest reference points of each query (GridKNNCuda.cu:1645-1939). queries: (nq, 3) array or
        (x, y, z) in input units; a torch CUDA tensor triple is taken as device input. Afterwards
        GetKNNIndexes() returns the query rows, as in the reference."""
        if isinstance(queries, (tuple, list)) and len(queries) == 3 and hasattr(queries[0], "data_ptr"):
            x, y, z = queries
            _check(_lib.egt_grid_knn_query_device(self.handle, x.data_ptr(), y.data_ptr(), z.data_ptr(), x.numel(),
                                                  int(k), float(thres)))
        else:
            x, y, z = _xyz(queries)
            _check(_lib.egt_grid_knn_query(self.handle, x, y, z, x.size, int(k), float(thres)))
        self._last_query = True

    def pointsQuerCount(self):
        if not getattr(self, "_last_query", False):
            return self.pointsRefCount()
        nq = ctypes.c_int64()
        _check(_lib.egt_get_device_query_knn(self.handle, None, ctypes.byref(nq), None))
        return nq.value

    def GetKNNIndexesQuery(self):
        """Rows of the last query search: (nq, k) int32, original query order, original reference indices."""
        nq, k = ctypes.c_int64(), ctypes.c_int32()
        _check(_lib.egt_get_device_query_knn(self.handle, None, ctypes.byref(nq), ctypes.byref(k)))
        out = np.empty(max(nq.value * k.value, 1), np.int32)
        _check(_lib.egt_get_knn_query_original(self.handle, out))
        return out[:nq.value * k.value].reshape(nq.value, k.value)

    def GetKNNIndexesQueryDevice(self):
        """Device address of the query rows (reference sorted indices), with their count and width."""
        p, nq, k = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int32()
        _check(_lib.egt_get_device_query_knn(self.handle, ctypes.byref(p), ctypes.byref(nq), ctypes.byref(k)))
        return p.value, nq.value, k.value

    def GetKNNIndexes(self):
        """Neighbour rows in original index space (GridKNNCuda.cu:1941-1968): the query rows after a query
        search, else the self rows."""
        if getattr(self, "_last_query", False):
            return self.GetKNNIndexesQuery()
        return super().GetKNNIndexes()


class GridMeshCuda(_Provider):
    """Mesh provider (GridMeshCuda1.cuh:16). triangles: (F, 3) uint32 original vertex indices."""

    def __init__(self, points, triangles, num_cards=1, eps=1e-6, bucketCount=1024):
        super().__init__()
        x, y, z = _xyz(points)
        tri = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1)
        h = ctypes.c_void_p()
        _check(_lib.egt_mesh_create(x, y, z, x.size, tri, tri.size // 3, num_cards, eps, bucketCount,
                                    ctypes.byref(h)))
        self._h = h

    def GetNeighborsFromTriangles(self, k):
        _check(_lib.egt_mesh_neighbors(self.handle, int(k)))


class BatchGrid(_Provider):
    """Many independent clouds in one handle (no reference counterpart). Each cloud behaves exactly like its
    own GridStructure; EGT_CUDA / Normal_CUDA / getOutputPoints work on the whole batch and use the
    concatenated input order (split() cuts such an array back into clouds)."""

    def __init__(self, clouds, eps=1e-6, bucketCount=1024, device=None):
        """clouds: list of point sets, or (with device=(dx, dy, dz) torch tensors holding the clouds one
        after another) the list of cloud sizes."""
        super().__init__()
        h = ctypes.c_void_p()
        if device is not None:
            self.sizes = np.asarray(clouds, np.int64)
            self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
            ptrs = [ctypes.c_void_p(t.data_ptr()) for t in device]
            _check(_lib.egt_batch_create_device(*ptrs, self.offsets, len(self.sizes), eps, bucketCount,
                                                ctypes.byref(h)))
        else:
            parts = [_xyz(c) for c in clouds]
            self.sizes = np.array([p[0].size for p in parts], np.int64)
            self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
            x, y, z = (np.ascontiguousarray(np.concatenate([p[i] for p in parts]), np.float32) for i in range(3))
            _check(_lib.egt_batch_create(x, y, z, self.offsets, len(parts), eps, bucketCount, ctypes.byref(h)))
        self._h = h.value
        self.k = None

    def knn(self, k, thres=1e8):
        """GRIDCUDAKNNSELF_COMPACT for every cloud; k is one count or one per cloud."""
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (len(self.sizes),)), np.int32)
        _check(_lib.egt_batch_knn(self.handle, ks, thres))
        self.k = ks

    def autotune_shape(self, k, thres=1e8):
        """Times each neighbour-count class's search in each kNN launch shape and keeps the fastest per class
        (no reference counterpart). Returns (packed choices for OPT_KNN_CLASS_SHAPES, ms array [8 classes, 3
        shapes: lane per query, persistent, wave per query], -1 where not run). Run knn() afterwards."""
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (len(self.sizes),)), np.int32)
        packed = ctypes.c_int64()
        ms = np.empty(24, np.float32)
        _check(_lib.egt_batch_autotune_shape(self.handle, ks, thres, ctypes.byref(packed), ms))
        return packed.value, ms.reshape(8, 3)

    def GetKNNIndexes(self, cloud):
        out = np.empty((int(self.sizes[cloud]), int(self.k[cloud])), np.int32)
        _check(_lib.egt_batch_get_knn_original(self.handle, cloud, out))
        return out

    def cloud_info(self, cloud):
        n, k, mn, mx = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_float(), ctypes.c_float()
        _check(_lib.egt_batch_cloud_info(self.handle, cloud, ctypes.byref(n), ctypes.byref(k), ctypes.byref(mn),
                                         ctypes.byref(mx)))
        return n.value, k.value, mn.value, mx.value

    def split(self, arr):
        return [arr[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.sizes))]


class EGT_CUDA:
    """EGG filter + Taubin smoother (EGT_CUDA.cuh:11-14)."""

    def __init__(self, provider):
        self.provider = provider

    def PerformSmoothing(self, iterationCount, lambda_=0.63, mu=-0.64, alpha=1e-8):
        _check(_lib.egt_smooth(self.provider.handle, int(iterationCount), float(lambda_), float(mu), float(alpha)))


class Normal_CUDA:
    """PCA normals + MST/BFS orientation (Normal_CUDA.cuh:10-13)."""

    def __init__(self, provider):
        self.provider = provider

    def GetNormals(self):
        n = self.provider.pointsRefCount()
        nx, ny, nz = (_host_out(n) for _ in range(3))
        _check(_lib.egt_normals(self.provider.handle, nx, ny, nz))
        return nx, ny, nz


_PINNED = None


def _host_out(n):
    """float32 output array for a device-to-host copy of n values. With a GPU present it is page-locked memory
    from torch's caching pinned allocator (plumbing only), so the library's copy runs as one DMA instead of
    through the runtime's staging of pageable memory (48 MB of normals: about 1 ms instead of 3); otherwise a
    plain numpy array. The array keeps its memory alive."""
    global _PINNED
    if _PINNED is None:
        try:
            import torch

            _PINNED = torch.cuda.is_available()
        except Exception:  # no torch: pageable output
            _PINNED = False
    if _PINNED:
        import torch

        return torch.empty(max(n, 1), dtype=torch.float32, pin_memory=True).numpy()[:n]
    return np.empty(n, np.float32)


def getOutputPoints(provider):
    """De-normalized points in original order (getOutputPoints.cu:34-63)."""
    n = provider.pointsRefCount()
    x, y, z = (_host_out(n) for _ in range(3))
    _check(_lib.egt_get_output_points(provider.handle, x, y, z))
    return x, y, z


class RadiusModel:
    """Learned search-radius proposal (no reference counterpart): MLP 28 -> 64 -> 64 -> 1 trained on the GPU
    by the library's BF16-MFMA forward / backward and Adam kernels. A proposal only moves where each query's
    search starts; the exactness test keeps neighbour rows identical to the plain search."""

    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        _check(_lib.egt_model_create(seed, ctypes.byref(h)))
        self._h = h.value

    @staticmethod
    def param_count():
        return int(_lib.egt_model_param_count())

    @staticmethod
    def feature_count():
        return int(_lib.egt_model_feature_count())

    @property
    def handle(self):
        if self._h is None:
            raise EGTError(6, "model already destroyed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.egt_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.param_count(), np.float32)
        _check(_lib.egt_model_get_params(self.handle, p, p.size))
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(_lib.egt_model_set_params(self.handle, p, p.size))

    def eval(self, features, targets):
        """Forward outputs, gradient of the batch MSE and the loss for host features (n x 28) and targets."""
        X = np.ascontiguousarray(features, np.float32)
        T = np.ascontiguousarray(targets, np.float32)
        y = np.empty(len(T), np.float32)
        g = np.empty(self.param_count(), np.float32)
        loss = ctypes.c_float()
        _check(_lib.egt_model_eval(self.handle, X, T, len(T), y, g, ctypes.byref(loss)))
        return y, g, loss.value

    def adam_step(self, features, targets, lr):
        X = np.ascontiguousarray(features, np.float32)
        T = np.ascontiguousarray(targets, np.float32)
        g = np.empty(self.param_count(), np.float32)
        loss = ctypes.c_float()
        _check(_lib.egt_model_adam_step(self.handle, X, T, len(T), lr, g, ctypes.byref(loss)))
        return g, loss.value

    def train(self, provider, k, epochs=4, batch=4096, max_samples=2_000_000, lr=2e-3, quantile=0.9):
        """Trains on a GridStructure (one k) or a BatchGrid (k per cloud); returns the final training MSE."""
        loss = ctypes.c_float()
        if isinstance(provider, BatchGrid):
            ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (len(provider.sizes),)), np.int32)
            _check(_lib.egt_model_train_batch(self.handle, provider.handle, ks, epochs, batch, max_samples, lr,
                                              quantile, ctypes.byref(loss)))
            provider.k = ks
        else:
            _check(_lib.egt_model_train(self.handle, provider.handle, k, epochs, batch, max_samples, lr, quantile,
                                        ctypes.byref(loss)))
        return loss.value

    @property
    def safety(self):
        v = ctypes.c_float()
        _check(_lib.egt_model_get_safety(self.handle, ctypes.byref(v)))
        return v.value

    @safety.setter
    def safety(self, value):
        _check(_lib.egt_model_set_safety(self.handle, value))

    @property
    def bound_safety(self):
        """Margin (log2 units) of the radius bound over the predicted radius."""
        v = ctypes.c_float()
        _check(_lib.egt_model_get_bound_safety(self.handle, ctypes.byref(v)))
        return v.value

    @bound_safety.setter
    def bound_safety(self, value):
        _check(_lib.egt_model_set_bound_safety(self.handle, value))

    def propose(self, provider, k):
        if isinstance(provider, BatchGrid):
            ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.int32), (len(provider.sizes),)), np.int32)
            _check(_lib.egt_model_propose_batch(self.handle, provider.handle, ks))
        else:
            _check(_lib.egt_model_propose(self.handle, provider.handle, k))


def model_features(provider, k):
    """(features n x 28, targets n, reference level) of a provider, sorted order (diagnostics)."""
    n = provider.pointsRefCount()
    X = np.empty((n, RadiusModel.feature_count()), np.float32)
    T = np.empty(n, np.float32)
    ref = ctypes.c_int32()
    _check(_lib.egt_model_features(provider.handle, k, X, T, ctypes.byref(ref)))
    return X, T, ref.value


def model_levels(provider):
    out = np.empty(provider.pointsRefCount(), np.int8)
    _check(_lib.egt_model_levels(provider.handle, out))
    return out


def release_cached_memory():
    """Returns the device blocks that destroyed handles left cached (egt_release_cached_memory);
    gives (bytes released, bytes that were cached)."""
    rel, cached = ctypes.c_uint64(), ctypes.c_uint64()
    _check(_lib.egt_release_cached_memory(ctypes.byref(rel), ctypes.byref(cached)))
    return rel.value, cached.value


def device_array(ptr, count, dtype):
    """Copy a device buffer handed out by an accessor into a numpy array (tests / debugging)."""
    import torch  # plumbing only

    torch_dtype = {np.float32: torch.float32, np.int32: torch.int32}[np.dtype(dtype).type]
    t = torch.empty(count, dtype=torch_dtype, device="cuda")
    nbytes = count * np.dtype(dtype).itemsize
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(ptr), ctypes.c_size_t(nbytes), 3)
    if rc != 0:
        raise EGTError(3, f"hipMemcpy failed ({rc})")
    return t.cpu().numpy()
