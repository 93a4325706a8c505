// extern "C" entry points of include/egt.h: argument checks, handle lifecycle, status codes.
#include <cmath>
#include <cstring>
#include <new>
#include "egt_internal.h"

namespace {
thread_local std::string g_err;

egt_status fail(egt_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

template <class F>
egt_status guarded(F&& f) {
    try {
        f();
        return EGT_OK;
    } catch (const egt::ArgError& e) {
        return fail(e.st, e.what);
    } catch (const egt::HipError& e) {
        return fail(e.err == hipErrorOutOfMemory ? EGT_ERR_OOM : EGT_ERR_HIP, e.what);
    } catch (const std::bad_alloc&) {
        return fail(EGT_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(EGT_ERR_STATE, e.what());
    }
}

bool pow2_in_range(int b) { return b >= 1 && b <= 1024 && (b & (b - 1)) == 0; }

egt_status check_common(int64_t n, int num_cards, int bucket_count, egt_handle* out) {
    if (!out) return fail(EGT_ERR_INVALID_ARG, "null handle pointer");
    if (n < 1) return fail(EGT_ERR_INVALID_ARG, "point count must be >= 1");
    if (n >= (int64_t)1 << 31) return fail(EGT_ERR_UNSUPPORTED, "point count must fit in int32");
    if (num_cards != 1) return fail(EGT_ERR_UNSUPPORTED, "num_cards must be 1: one GPU holds the whole cloud");
    if (!pow2_in_range(bucket_count)) return fail(EGT_ERR_INVALID_ARG, "bucketCount must be a power of two in [1, 1024]");
    return EGT_OK;
}

egt_status make_handle(const float* x, const float* y, const float* z, int64_t n, float eps, int bucket_count,
                       bool device_input, int kind, egt_handle* out) {
    if (!x || !y || !z) return fail(EGT_ERR_INVALID_ARG, "null coordinate array");
    if (!(eps >= 0.0f) || !std::isfinite(eps)) return fail(EGT_ERR_INVALID_ARG, "eps must be finite and >= 0");
    egt_ctx* h = new (std::nothrow) egt_ctx();
    if (!h) return fail(EGT_ERR_OOM, "host allocation failed");
    egt_status st = guarded([&] {
        HIP_CHECK(hipGetDevice(&h->device));
        HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->kind = kind;
        h->n = n;
        h->B = bucket_count;
        h->eps = eps;
        h->stats.n_points = n;
        h->stats.bucket_count = bucket_count;
        if (device_input) {
            egt::build_grid(h, x, y, z);
        } else {
            float* tmp = h->scratch.get<float>(16, 3 * (size_t)n);
            HIP_CHECK(hipMemcpyAsync(tmp, x, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + n, y, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + 2 * n, z, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            egt::build_grid(h, tmp, tmp + n, tmp + 2 * n);
        }
    });
    if (st != EGT_OK) {
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return st;
    }
    *out = h;
    return EGT_OK;
}

egt_status need(egt_handle h) {
    if (!h) return fail(EGT_ERR_INVALID_ARG, "null handle");
    return EGT_OK;
}
}  // namespace

extern "C" {

const char* egt_last_error_string(void) { return g_err.c_str(); }

egt_status egt_grid_create(const float* x, const float* y, const float* z, int64_t n, int num_cards, float eps,
                           int bucket_count, int max_division, egt_handle* out) {
    egt_status st = check_common(n, num_cards, bucket_count, out);
    if (st != EGT_OK) return st;
    if (max_division < 1) return fail(EGT_ERR_INVALID_ARG, "maxDivision must be >= 1");
    return make_handle(x, y, z, n, eps, bucket_count, false, 0, out);
}

egt_status egt_grid_create_device(const float* dx, const float* dy, const float* dz, int64_t n, int num_cards,
                                  float eps, int bucket_count, int max_division, egt_handle* out) {
    egt_status st = check_common(n, num_cards, bucket_count, out);
    if (st != EGT_OK) return st;
    if (max_division < 1) return fail(EGT_ERR_INVALID_ARG, "maxDivision must be >= 1");
    return make_handle(dx, dy, dz, n, eps, bucket_count, true, 0, out);
}

egt_status egt_grid_knn_self(egt_handle h, int k, float thres) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind == 2) return fail(EGT_ERR_STATE, "a batch handle searches with egt_batch_knn");
    if (h->kind != 0) return fail(EGT_ERR_STATE, "kNN search needs a point-cloud handle");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    return guarded([&] {
        // GridKNNCuda.cu:2324-2325
        float t = thres / (h->max_extent - h->min_extent);
        t *= t;
        if (h->search_stale) egt::refresh_search_structure(h);
        egt::run_knn(h, k, t);
        h->csr_valid = false;
        h->smoothed = false;
    });
}

namespace {
egt_status knn_query(egt_handle h, const float* qx, const float* qy, const float* qz, int64_t nq, int k, float thres,
                     bool device) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind == 2) return fail(EGT_ERR_STATE, "a batch handle has no query search");
    if (h->kind != 0) return fail(EGT_ERR_STATE, "the query search needs a point-cloud handle");
    if (nq < 0) return fail(EGT_ERR_INVALID_ARG, "query count must be >= 0");
    if (nq > 0 && (!qx || !qy || !qz)) return fail(EGT_ERR_INVALID_ARG, "null query array");
    if (nq > INT32_MAX) return fail(EGT_ERR_UNSUPPORTED, "more than 2^31 - 1 queries");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    if (!device) {
        for (int64_t i = 0; i < nq; ++i)
            if (!std::isfinite(qx[i]) || !std::isfinite(qy[i]) || !std::isfinite(qz[i]))
                return fail(EGT_ERR_INVALID_ARG, "query coordinates must be finite");
    }
    return guarded([&] {
        // GridKNNCuda.cu:1649-1650
        float t = thres / (h->max_extent - h->min_extent);
        t *= t;
        if (h->search_stale) egt::refresh_search_structure(h);
        const float *dx = qx, *dy = qy, *dz = qz;
        if (!device && nq > 0) {
            float* tmp = h->scratch.get<float>(16, 3 * (size_t)nq);
            HIP_CHECK(hipMemcpyAsync(tmp, qx, nq * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + nq, qy, nq * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + 2 * nq, qz, nq * sizeof(float), hipMemcpyHostToDevice, h->stream));
            dx = tmp;
            dy = tmp + nq;
            dz = tmp + 2 * nq;
        }
        egt::run_knn_query(h, dx, dy, dz, nq, k, t);
    });
}
}  // namespace

egt_status egt_grid_knn_query(egt_handle h, const float* qx, const float* qy, const float* qz, int64_t nq, int k,
                              float thres) {
    return knn_query(h, qx, qy, qz, nq, k, thres, false);
}

egt_status egt_grid_knn_query_device(egt_handle h, const float* qx, const float* qy, const float* qz, int64_t nq,
                                     int k, float thres) {
    return knn_query(h, qx, qy, qz, nq, k, thres, true);
}

egt_status egt_grid_knn_query_device_noalloc(egt_handle h, const float* qx, const float* qy, const float* qz,
                                             int64_t nq, int k, float thres, int32_t* rows) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 0) return fail(EGT_ERR_STATE, "the query search needs a one-cloud point-cloud handle");
    if (nq < 0) return fail(EGT_ERR_INVALID_ARG, "query count must be >= 0");
    if (nq > egt::QNA_MAX_QUERIES)
        return fail(EGT_ERR_UNSUPPORTED, "more than 1048576 queries in one allocation-free call");
    if (nq > 0 && (!qx || !qy || !qz || !rows)) return fail(EGT_ERR_INVALID_ARG, "null query or row array");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    if (h->search_stale)
        return fail(EGT_ERR_STATE, "the points moved since the search structure was built; run egt_grid_knn_query "
                                   "or egt_grid_knn_self once to rebuild it");
    return guarded([&] {
        int dev = -1;
        HIP_CHECK(hipGetDevice(&dev));
        if (dev != h->device) throw egt::ArgError{EGT_ERR_STATE, "the handle's device is not the current device"};
        float t = thres / (h->max_extent - h->min_extent);  // GridKNNCuda.cu:1649-1650
        t *= t;
        egt::run_knn_query_noalloc(h, qx, qy, qz, nq, k, t, rows);
    });
}

egt_status egt_get_knn_query_original(egt_handle h, int32_t* out) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!h->have_query) return fail(EGT_ERR_STATE, "no query search yet");
    if (!out && h->nq > 0) return fail(EGT_ERR_INVALID_ARG, "null output array");
    return guarded([&] { egt::run_knn_query_original(h, out); });
}

egt_status egt_get_device_query_knn(egt_handle h, int32_t** rows, int64_t* nq, int32_t* k) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (rows) *rows = h->have_query ? h->qknn.p : nullptr;
    if (nq) *nq = h->have_query ? h->nq : 0;
    if (k) *k = h->have_query ? h->kq : 0;
    return EGT_OK;
}

egt_status egt_mesh_create(const float* x, const float* y, const float* z, int64_t n, const uint32_t* tri,
                           int64_t ntri, int num_cards, float eps, int bucket_count, egt_handle* out) {
    egt_status st = check_common(n, num_cards, bucket_count, out);
    if (st != EGT_OK) return st;
    if (ntri < 0 || (ntri > 0 && !tri)) return fail(EGT_ERR_INVALID_ARG, "bad triangle array");
    if (const char* err = egt::mesh_triangles_error(ntri)) return fail(EGT_ERR_UNSUPPORTED, err);
    st = make_handle(x, y, z, n, eps, bucket_count, false, 1, out);
    if (st != EGT_OK) return st;
    egt_handle h = *out;
    st = guarded([&] {
        h->ntri = ntri;
        h->tri.alloc(3 * (size_t)std::max<int64_t>(ntri, 1));
        if (ntri > 0)
            HIP_CHECK(hipMemcpyAsync(h->tri.p, tri, 3 * ntri * sizeof(uint32_t), hipMemcpyHostToDevice, h->stream));
        HIP_CHECK(hipStreamSynchronize(h->stream));
    });
    if (st != EGT_OK) {
        egt_destroy(h);
        *out = nullptr;
    }
    return st;
}

egt_status egt_mesh_neighbors(egt_handle h, int k) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 1) return fail(EGT_ERR_STATE, "mesh neighbours need a mesh handle");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    return guarded([&] {
        egt::run_mesh_adjacency(h, k);
        h->csr_valid = false;
        h->smoothed = false;
    });
}

egt_status egt_smooth(egt_handle h, int iterations, float lambda, float mu, float alpha) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!h->have_knn) return fail(EGT_ERR_STATE, "run the kNN / mesh neighbour step first");
    if (iterations < 0) return fail(EGT_ERR_INVALID_ARG, "iterations must be >= 0");
    if (!std::isfinite(lambda) || !std::isfinite(mu) || !std::isfinite(alpha))
        return fail(EGT_ERR_INVALID_ARG, "lambda, mu and alpha must be finite");
    return guarded([&] {
        egt::run_egg(h, alpha);
        egt::run_taubin(h, iterations, lambda, mu);
        h->smoothed = true;
    });
}

egt_status egt_normals(egt_handle h, float* nx, float* ny, float* nz) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!h->have_knn) return fail(EGT_ERR_STATE, "run the kNN / mesh neighbour step first");
    if (!nx || !ny || !nz) return fail(EGT_ERR_INVALID_ARG, "null normal array");
    return guarded([&] { egt::run_normals(h, nx, ny, nz); });
}

egt_status egt_get_output_points(egt_handle h, float* x, float* y, float* z) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!x || !y || !z) return fail(EGT_ERR_INVALID_ARG, "null output array");
    return guarded([&] { egt::run_output(h, x, y, z); });
}

egt_status egt_get_knn_original(egt_handle h, int32_t* out) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind == 2) return fail(EGT_ERR_STATE, "a batch handle exports rows per cloud (egt_batch_get_knn_original)");
    if (!h->have_knn) return fail(EGT_ERR_STATE, "no neighbour rows yet");
    if (!out) return fail(EGT_ERR_INVALID_ARG, "null output array");
    return guarded([&] { egt::run_knn_original(h, out); });
}

egt_status egt_get_device_points(egt_handle h, float** x, float** y, float** z) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (x) *x = h->x.p;
    if (y) *y = h->y.p;
    if (z) *z = h->z.p;
    return EGT_OK;
}

egt_status egt_get_device_knn(egt_handle h, int32_t** knn) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (knn) *knn = h->have_knn ? h->knn.p : nullptr;
    return EGT_OK;
}

egt_status egt_get_device_original_indexes(egt_handle h, int32_t** idx) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (idx) *idx = h->perm.p;
    return EGT_OK;
}

egt_status egt_get_extent(egt_handle h, float* mn, float* mx) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (mn) *mn = h->min_extent;
    if (mx) *mx = h->max_extent;
    return EGT_OK;
}

egt_status egt_get_counts(egt_handle h, int64_t* n, int32_t* k, int32_t* cards, int64_t* part) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (n) *n = h->n;
    if (k) *k = h->k;
    if (cards) *cards = 1;
    if (part) *part = h->n;
    return EGT_OK;
}

egt_status egt_get_stats(egt_handle h, egt_stats* st) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!st) return fail(EGT_ERR_INVALID_ARG, "null stats pointer");
    *st = h->stats;
    st->k = h->k;
    return EGT_OK;
}

egt_status egt_release_cached_memory(uint64_t* released, uint64_t* cached) {
    return guarded([&] {
        if (cached) *cached = egt::device_cache_bytes();
        const size_t r = egt::device_cache_release_all();
        if (released) *released = r;
    });
}

egt_status egt_set_option(egt_handle h, egt_option opt, int64_t value) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    switch (opt) {
        case EGT_OPT_START_LEVEL:
            if (value != EGT_START_LEVEL_OCCUPANCY && (value < -egt::SUB || value >= h->nlev))
                return fail(EGT_ERR_INVALID_ARG, "start level out of range");
            h->start_level = (int)value;
            return EGT_OK;
        case EGT_OPT_LEARNED:
            h->learned = value ? 1 : 0;
            return EGT_OK;
        case EGT_OPT_TAUBIN_GRAPH:
            h->taubin_graph = value ? 1 : 0;
            return EGT_OK;
        case EGT_OPT_MAX_PTS_IN_CUBE:
            if (value < 1 || value > INT32_MAX) return fail(EGT_ERR_INVALID_ARG, "max points in cube must be >= 1");
            h->maxc = (int)value;
            return EGT_OK;
        case EGT_OPT_KNN_PERSISTENT:
            h->knn_persistent = value ? 1 : 0;
            return EGT_OK;
        case EGT_OPT_TAUBIN_TILE:
            if (value != 0 && value != 1024 && value != 2048)
                return fail(EGT_ERR_INVALID_ARG, "Taubin tile width must be 1024, 2048 or 0 (autotune)");
            h->taubin_tile = (int)value;
            return EGT_OK;
        case EGT_OPT_TAUBIN_REGULARIZED:
            h->taubin_regularized = value ? 1 : 0;
            return EGT_OK;
        case EGT_OPT_LEARNED_MODE:
            if (value < 1 || value > 3) return fail(EGT_ERR_INVALID_ARG, "learned mode must be 1, 2 or 3");
            h->learned_mode = (int)value;
            return EGT_OK;
        case EGT_OPT_KNN_LANES:
            if (value != 1 && value != 64) return fail(EGT_ERR_INVALID_ARG, "kNN lanes per query must be 1 or 64");
            h->knn_lanes = (int)value;
            return EGT_OK;
        case EGT_OPT_KNN_LONG_QUERY:
            if (value < 0 || value > UINT32_MAX) return fail(EGT_ERR_INVALID_ARG, "long-query cap must be in [0, 2^32)");
            h->knn_long_cap = value;
            return EGT_OK;
        case EGT_OPT_LEARNED_LEVEL_OFFSET:
            if (value < -3 || value > 3) return fail(EGT_ERR_INVALID_ARG, "learned level offset must be in [-3, 3]");
            h->learned_level_offset = (int)value;
            return EGT_OK;
        case EGT_OPT_KNN_CLASS_SHAPES:
            if (value < 0 || value >= (1 << 16)) return fail(EGT_ERR_INVALID_ARG, "class shapes: 8 fields of 2 bits");
            for (int c = 0; c < 8; ++c) h->class_shape[c] = (int)((value >> (2 * c)) & 3) - 1;
            return EGT_OK;
    }
    return fail(EGT_ERR_INVALID_ARG, "unknown option");
}

egt_status egt_autotune(egt_handle h, int k, float thres, int32_t* start_level, float* trial_ms,
                        float* trial_cost) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 0) return fail(EGT_ERR_STATE, "the autotuner needs a one-cloud point-cloud handle");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    return guarded([&] {
        float t = thres / (h->max_extent - h->min_extent);
        t *= t;
        if (h->search_stale) egt::refresh_search_structure(h);
        float ms[EGT_AUTOTUNE_SLOTS], cost[EGT_AUTOTUNE_SLOTS];
        const int best = egt::autotune_start_level(h, k, t, ms, cost);
        h->csr_valid = false;
        if (start_level) *start_level = best;
        for (int i = 0; i < EGT_AUTOTUNE_SLOTS; ++i) {
            if (trial_ms) trial_ms[i] = ms[i];
            if (trial_cost) trial_cost[i] = cost[i];
        }
    });
}

egt_status egt_autotune_shape(egt_handle h, int k, float thres, int32_t* shape, float* shape_ms) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 0) return fail(EGT_ERR_STATE, "the shape tuner needs a one-cloud point-cloud handle");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    return guarded([&] {
        float t = thres / (h->max_extent - h->min_extent);
        t *= t;
        if (h->search_stale) egt::refresh_search_structure(h);
        float ms[3];
        const int best = egt::autotune_shape(h, k, t, ms);
        h->csr_valid = false;
        if (shape) *shape = best;
        if (shape_ms)
            for (int i = 0; i < 3; ++i) shape_ms[i] = ms[i];
    });
}

namespace {
egt_status batch_create(const float* x, const float* y, const float* z, const int64_t* offsets, int32_t nclouds,
                        float eps, int bucket_count, bool device_input, egt_handle* out) {
    if (!out) return fail(EGT_ERR_INVALID_ARG, "null handle pointer");
    if (!x || !y || !z || !offsets) return fail(EGT_ERR_INVALID_ARG, "null coordinate or offset array");
    if (nclouds < 1 || nclouds > (1 << 20)) return fail(EGT_ERR_INVALID_ARG, "cloud count must be in [1, 2^20]");
    if (offsets[0] != 0) return fail(EGT_ERR_INVALID_ARG, "offsets[0] must be 0");
    for (int32_t c = 0; c < nclouds; ++c)
        if (offsets[c + 1] <= offsets[c]) return fail(EGT_ERR_INVALID_ARG, "every cloud needs at least one point");
    const int64_t n = offsets[nclouds];
    if (n >= (int64_t)1 << 31) return fail(EGT_ERR_UNSUPPORTED, "total point count must fit in int32");
    if (!pow2_in_range(bucket_count)) return fail(EGT_ERR_INVALID_ARG, "bucketCount must be a power of two in [1, 1024]");
    if (!(eps >= 0.0f) || !std::isfinite(eps)) return fail(EGT_ERR_INVALID_ARG, "eps must be finite and >= 0");
    egt_ctx* h = new (std::nothrow) egt_ctx();
    if (!h) return fail(EGT_ERR_OOM, "host allocation failed");
    egt_status st = guarded([&] {
        HIP_CHECK(hipGetDevice(&h->device));
        HIP_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->kind = 2;
        h->n = n;
        h->B = bucket_count;
        h->eps = eps;
        h->knn_persistent = 0;  // one wave per chunk: 221 ms against 300 ms persistent (512 clouds, 54.7M points)
        h->stats.n_points = n;
        h->stats.bucket_count = bucket_count;
        if (device_input) {
            egt::build_grid_batch(h, x, y, z, offsets, nclouds);
        } else {
            float* tmp = h->scratch.get<float>(16, 3 * (size_t)n);
            HIP_CHECK(hipMemcpyAsync(tmp, x, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + n, y, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            HIP_CHECK(hipMemcpyAsync(tmp + 2 * n, z, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
            egt::build_grid_batch(h, tmp, tmp + n, tmp + 2 * n, offsets, nclouds);
        }
    });
    if (st != EGT_OK) {
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return st;
    }
    *out = h;
    return EGT_OK;
}
}  // namespace

egt_status egt_batch_create(const float* x, const float* y, const float* z, const int64_t* offsets, int32_t nclouds,
                            float eps, int bucket_count, egt_handle* out) {
    return batch_create(x, y, z, offsets, nclouds, eps, bucket_count, false, out);
}

egt_status egt_batch_create_device(const float* dx, const float* dy, const float* dz, const int64_t* offsets,
                                   int32_t nclouds, float eps, int bucket_count, egt_handle* out) {
    return batch_create(dx, dy, dz, offsets, nclouds, eps, bucket_count, true, out);
}

egt_status egt_batch_knn(egt_handle h, const int32_t* k, float thres) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 2) return fail(EGT_ERR_STATE, "egt_batch_knn needs a batch handle");
    if (!k) return fail(EGT_ERR_INVALID_ARG, "null k array");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    for (int c = 0; c < h->nclouds; ++c) {
        if (k[c] < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
        if (k[c] > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    }
    return guarded([&] {
        if (h->search_stale) egt::refresh_search_structure(h);
        int kmax = 0;
        for (int c = 0; c < h->nclouds; ++c) {
            egt::CloudDev& cd = h->clouds[c];
            cd.k = k[c];
            float t = thres / cd.extent;  // GridKNNCuda.cu:2324-2325, per cloud
            cd.thres_n = t * t;
            kmax = std::max(kmax, k[c]);
        }
        egt::run_knn_clouds(h);
        h->k = kmax;
        h->csr_valid = false;
        h->smoothed = false;
    });
}

egt_status egt_autotune_learned(egt_handle h, int k, float thres, int32_t* offset, int32_t* mode, int32_t* shape,
                                float* ms) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 0) return fail(EGT_ERR_STATE, "the learned-start tuner needs a one-cloud point-cloud handle");
    if (!h->learned || !h->qlevel.p) return fail(EGT_ERR_STATE, "no learned proposal in place (egt_model_propose)");
    if (k < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
    if (k > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    return guarded([&] {
        float t = thres / (h->max_extent - h->min_extent);
        t *= t;
        if (h->search_stale) egt::refresh_search_structure(h);
        float m8[8];
        int o = 0, md = 1, sh = 0;
        egt::autotune_learned(h, k, t, &o, &md, &sh, m8);
        h->csr_valid = false;
        if (offset) *offset = o;
        if (mode) *mode = md;
        if (shape) *shape = sh;
        if (ms)
            for (int i = 0; i < 8; ++i) ms[i] = m8[i];
    });
}

egt_status egt_batch_autotune_shape(egt_handle h, const int32_t* k, float thres, int64_t* shapes, float* class_ms) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 2) return fail(EGT_ERR_STATE, "egt_batch_autotune_shape needs a batch handle");
    if (!k) return fail(EGT_ERR_INVALID_ARG, "null k array");
    if (!(thres > 0.0f)) return fail(EGT_ERR_INVALID_ARG, "thres must be > 0");
    for (int c = 0; c < h->nclouds; ++c) {
        if (k[c] < 1) return fail(EGT_ERR_INVALID_ARG, "k must be >= 1");
        if (k[c] > 64) return fail(EGT_ERR_UNSUPPORTED, "k > 64 is not supported");
    }
    return guarded([&] {
        if (h->search_stale) egt::refresh_search_structure(h);
        for (int c = 0; c < h->nclouds; ++c) {
            egt::CloudDev& cd = h->clouds[c];
            cd.k = k[c];
            float t = thres / cd.extent;
            cd.thres_n = t * t;
        }
        float ms[24];
        const int64_t packed = egt::autotune_class_shapes(h, ms);
        h->csr_valid = false;
        if (shapes) *shapes = packed;
        if (class_ms)
            for (int i = 0; i < 24; ++i) class_ms[i] = ms[i];
    });
}

egt_status egt_batch_get_knn_original(egt_handle h, int32_t cloud, int32_t* out) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 2) return fail(EGT_ERR_STATE, "egt_batch_get_knn_original needs a batch handle");
    if (cloud < 0 || cloud >= h->nclouds) return fail(EGT_ERR_INVALID_ARG, "cloud index out of range");
    if (!h->have_knn) return fail(EGT_ERR_STATE, "no neighbour rows yet");
    if (!out) return fail(EGT_ERR_INVALID_ARG, "null output array");
    return guarded([&] { egt::run_knn_original_cloud(h, cloud, out); });
}

egt_status egt_batch_cloud_info(egt_handle h, int32_t cloud, int64_t* n, int32_t* k, float* min_extent,
                                float* max_extent) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (cloud < 0 || cloud >= h->nclouds) return fail(EGT_ERR_INVALID_ARG, "cloud index out of range");
    const egt::CloudDev& c = h->clouds[cloud];
    if (n) *n = c.n;
    if (k) *k = h->have_knn ? c.k : 0;
    if (min_extent) *min_extent = c.min_extent;
    if (max_extent) *max_extent = c.max_extent;
    return EGT_OK;
}

egt_status egt_destroy(egt_handle h) {
    if (!h) return EGT_OK;
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    delete h;
    return EGT_OK;
}

void* egt_get_stream(egt_handle h) { return h ? (void*)h->stream : nullptr; }

int32_t egt_model_param_count(void) { return egt::mlp_param_count(); }
int32_t egt_model_feature_count(void) { return egt::mlp_feature_count(); }

egt_status egt_model_create(uint64_t seed, egt_model* out) {
    if (!out) return fail(EGT_ERR_INVALID_ARG, "null model pointer");
    egt_model_s* m = new (std::nothrow) egt_model_s();
    if (!m) return fail(EGT_ERR_OOM, "host allocation failed");
    egt_status st = guarded([&] { egt::mlp_init(m, seed); });
    if (st != EGT_OK) {
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        return st;
    }
    *out = m;
    return EGT_OK;
}

egt_status egt_model_destroy(egt_model m) {
    if (!m) return EGT_OK;
    if (m->stream) {
        (void)hipStreamSynchronize(m->stream);
        (void)hipStreamDestroy(m->stream);
    }
    delete m;
    return EGT_OK;
}

egt_status egt_model_get_params(egt_model m, float* params, int64_t count) {
    if (!m || !params) return fail(EGT_ERR_INVALID_ARG, "null model or parameter array");
    if (count != egt::mlp_param_count()) return fail(EGT_ERR_INVALID_ARG, "count must be egt_model_param_count()");
    return guarded([&] { egt::mlp_get_params(m, params); });
}

egt_status egt_model_set_params(egt_model m, const float* params, int64_t count) {
    if (!m || !params) return fail(EGT_ERR_INVALID_ARG, "null model or parameter array");
    if (count != egt::mlp_param_count()) return fail(EGT_ERR_INVALID_ARG, "count must be egt_model_param_count()");
    return guarded([&] { egt::mlp_set_params(m, params); });
}

egt_status egt_model_eval(egt_model m, const float* features, const float* targets, int64_t n, float* y, float* grad,
                          float* loss) {
    if (!m || !features || !targets) return fail(EGT_ERR_INVALID_ARG, "null model, features or targets");
    if (n < 1 || n > ((int64_t)1 << 26)) return fail(EGT_ERR_INVALID_ARG, "batch size must be in [1, 2^26]");
    return guarded([&] { egt::mlp_eval(m, features, targets, n, y, grad, loss); });
}

egt_status egt_model_adam_step(egt_model m, const float* features, const float* targets, int64_t n, float lr,
                               float* grad, float* loss) {
    if (!m || !features || !targets) return fail(EGT_ERR_INVALID_ARG, "null model, features or targets");
    if (n < 1 || n > ((int64_t)1 << 26)) return fail(EGT_ERR_INVALID_ARG, "batch size must be in [1, 2^26]");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(EGT_ERR_INVALID_ARG, "learning rate must be > 0");
    return guarded([&] { egt::mlp_adam_step(m, features, targets, n, lr, grad, loss); });
}

namespace {
egt_status check_train_args(int epochs, int batch, int64_t max_samples, float lr, float quantile) {
    if (epochs < 0 || batch < 32) return fail(EGT_ERR_INVALID_ARG, "epochs must be >= 0 and batch >= 32");
    if (max_samples < 0) return fail(EGT_ERR_INVALID_ARG, "max_samples must be >= 0");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(EGT_ERR_INVALID_ARG, "learning rate must be > 0");
    if (!(quantile >= 0.0f && quantile <= 1.0f)) return fail(EGT_ERR_INVALID_ARG, "quantile must be in [0, 1]");
    return EGT_OK;
}
egt_status check_batch_ks(egt_handle h, const int32_t* k) {
    if (!k) return fail(EGT_ERR_INVALID_ARG, "null k array");
    for (int c = 0; c < h->nclouds; ++c)
        if (k[c] < 1 || k[c] > 64) return fail(EGT_ERR_INVALID_ARG, "k must be in [1, 64]");
    return EGT_OK;
}
}  // namespace

egt_status egt_model_train_batch(egt_model m, egt_handle h, const int32_t* k, int epochs, int batch,
                                 int64_t max_samples, float lr, float quantile, float* loss) {
    if (!m || need(h)) return fail(EGT_ERR_INVALID_ARG, "null model or handle");
    if (h->kind != 2) return fail(EGT_ERR_STATE, "egt_model_train_batch needs a batch handle");
    if (egt_status st = check_batch_ks(h, k)) return st;
    if (egt_status st = check_train_args(epochs, batch, max_samples, lr, quantile)) return st;
    return guarded([&] {
        h->csr_valid = false;
        h->smoothed = false;
        egt::mlp_train(m, h, k, epochs, batch, max_samples, lr, quantile, loss);
    });
}

egt_status egt_model_propose_batch(egt_model m, egt_handle h, const int32_t* k) {
    if (!m || need(h)) return fail(EGT_ERR_INVALID_ARG, "null model or handle");
    if (h->kind != 2) return fail(EGT_ERR_STATE, "egt_model_propose_batch needs a batch handle");
    if (egt_status st = check_batch_ks(h, k)) return st;
    return guarded([&] { egt::mlp_propose(m, h, k); });
}

egt_status egt_model_train(egt_model m, egt_handle h, int k, int epochs, int batch, int64_t max_samples, float lr,
                           float quantile, float* loss) {
    if (!m || need(h)) return fail(EGT_ERR_INVALID_ARG, "null model or handle");
    if (h->kind != 0) return fail(EGT_ERR_UNSUPPORTED, "training needs a one-cloud point-cloud handle");
    if (k < 1 || k > 64 || k >= h->n) return fail(EGT_ERR_INVALID_ARG, "k must be in [1, min(64, n - 1)]");
    if (epochs < 0 || batch < 32) return fail(EGT_ERR_INVALID_ARG, "epochs must be >= 0 and batch >= 32");
    if (max_samples < 0) return fail(EGT_ERR_INVALID_ARG, "max_samples must be >= 0");
    if (!(lr > 0.0f) || !std::isfinite(lr)) return fail(EGT_ERR_INVALID_ARG, "learning rate must be > 0");
    if (!(quantile >= 0.0f && quantile <= 1.0f)) return fail(EGT_ERR_INVALID_ARG, "quantile must be in [0, 1]");
    return guarded([&] {
        h->csr_valid = false;
        h->smoothed = false;
        egt::mlp_train(m, h, &k, epochs, batch, max_samples, lr, quantile, loss);
    });
}

egt_status egt_model_get_safety(egt_model m, float* safety) {
    if (!m || !safety) return fail(EGT_ERR_INVALID_ARG, "null model or output");
    *safety = m->safety;
    return EGT_OK;
}

egt_status egt_model_set_safety(egt_model m, float safety) {
    if (!m) return fail(EGT_ERR_INVALID_ARG, "null model");
    if (!std::isfinite(safety)) return fail(EGT_ERR_INVALID_ARG, "safety must be finite");
    m->safety = safety;
    return EGT_OK;
}

egt_status egt_model_get_bound_safety(egt_model m, float* safety) {
    if (!m || !safety) return fail(EGT_ERR_INVALID_ARG, "null model or output");
    *safety = m->bound_safety;
    return EGT_OK;
}

egt_status egt_model_set_bound_safety(egt_model m, float safety) {
    if (!m) return fail(EGT_ERR_INVALID_ARG, "null model");
    if (!std::isfinite(safety)) return fail(EGT_ERR_INVALID_ARG, "safety must be finite");
    m->bound_safety = safety;
    return EGT_OK;
}

egt_status egt_model_propose(egt_model m, egt_handle h, int k) {
    if (!m || need(h)) return fail(EGT_ERR_INVALID_ARG, "null model or handle");
    if (h->kind != 0) return fail(EGT_ERR_UNSUPPORTED, "proposals apply to one-cloud point-cloud handles");
    if (k < 1 || k > 64) return fail(EGT_ERR_INVALID_ARG, "k must be in [1, 64]");
    return guarded([&] { egt::mlp_propose(m, h, &k); });
}

egt_status egt_model_features(egt_handle h, int k, float* features, float* targets, int32_t* ref_level) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (h->kind != 0) return fail(EGT_ERR_UNSUPPORTED, "feature export applies to one-cloud point-cloud handles");
    if (k < 1 || k > 64) return fail(EGT_ERR_INVALID_ARG, "k must be in [1, 64]");
    return guarded([&] {
        int r = 0;
        egt::mlp_features(h, k, features, targets, &r);
        if (ref_level) *ref_level = r;
    });
}

egt_status egt_model_levels(egt_handle h, int8_t* levels) {
    if (need(h)) return EGT_ERR_INVALID_ARG;
    if (!levels) return fail(EGT_ERR_INVALID_ARG, "null output");
    return guarded([&] { egt::mlp_levels(h, levels); });
}

}  // extern "C"
