/*
 * egt.h -- C ABI of the MI355X (gfx950) grid-kNN -> elliptic Gabriel graph -> Taubin smoothing library.
 *
 * Drop-in boundary for the reference's KNNInterface / GridStructure / GridMeshCuda / EGT_CUDA /
 * Normal_CUDA / getOutputPoints path (agalex1974/MultiGPUEGTSmoothing). Each entry point below names the
 * reference interface it replaces (file:line in the reference repository). Plain pointers and sizes only.
 *
 * Conventions
 *  - One handle = one point set (cloud or mesh) resident on the current HIP device, with its own stream.
 *    Calls on one handle must be serialized by the caller; different handles are independent.
 *  - "Sorted space": points are normalized into the reference's cube and permuted by a stable sort on the
 *    row-major cell key x + y*B + z*B*B at bucket count B (GridKNNCuda.cu:74-88, 898-936, 1107-1128).
 *    Device buffers handed out by the accessors are in that space; host outputs are in original order.
 *  - Every call returns an egt_status; egt_last_error_string() describes the last failure on this thread.
 *  - num_cards must be 1: the whole point set lives on one GPU.
 *  - Limits (each refused with EGT_ERR_UNSUPPORTED and a message, never wrapped silently):
 *      k <= 64 at every search entry point. The reference takes any k (its lists live in global memory,
 *        GridKNNCuda.cu:311-339; GridMeshCuda1.cuh:17); here a query's list is held in registers.
 *      Kept neighbour slots (the EGG adjacency, or every valid slot when normals follow the search directly)
 *        <= 2^31 - 1 per handle, all clouds together: the adjacency is a CSR with int32 offsets. egt_smooth /
 *        egt_normals check the exact count after the filter; when it does not fit they return the error with
 *        the rows already EGG-filtered (the filter's in-place side effect, EGT_CUDA.cu:185) and no smoothing
 *        done.
 *      Mesh triangles <= (2^31 - 1) / 6 (int32 corner-list offsets); points < 2^31.
 */
#ifndef EGT_H
#define EGT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EGT_OK = 0,
    EGT_ERR_INVALID_ARG = 1,
    EGT_ERR_OOM = 2,
    EGT_ERR_HIP = 3,
    EGT_ERR_UNSUPPORTED = 4,
    EGT_ERR_IO = 5,
    EGT_ERR_STATE = 6
} egt_status;

typedef struct egt_ctx* egt_handle;

/* Per-handle counters (no reference counterpart: the reference prints them to stdout,
 * GridKNNCuda.cu:2422-2425, 2558-2561; Normal_CUDA.cu:339-436). */
typedef struct {
    int64_t n_points;
    int32_t k;
    int32_t bucket_count;
    int64_t invalid_rows;      /* kNN rows blanked by the reference's validity rule */
    int64_t knn_candidates;    /* candidate distances evaluated by the kNN search */
    int64_t egg_kept;          /* neighbour slots kept by the EGG filter (CSR nnz) */
    int64_t learned_queries;   /* queries that started from the learned radius proposal */
    int64_t learned_fallbacks; /* of those, queries whose proposal failed the certificate */
    int64_t msf_edges;         /* spanning-forest edges of the orientation step */
    int64_t msf_trees;         /* trees (components with at least one edge) */
    double msf_weight;         /* total forest weight */
    float knn_ms, egg_ms, taubin_ms, normals_ms, grid_ms; /* device time of the last call of each stage */
    float propose_ms;          /* device time of the last learned proposal (features + MLP + level choice) */
    int64_t knn_cells;         /* cells looked up by the kNN search */
    float taubin_tile_ms[2];   /* Taubin tile autotuner: ms per pass at tile width 1024 and 2048 (0 if untried) */
    int32_t taubin_tile;       /* tile width of the last smoothing call */
    int32_t taubin_fallback_tiles; /* tiles of that call run as CSR rows with global gathers: halo list over
                                      its cap, or a row that lists itself (mesh with degenerate triangles) */
    int64_t query_count;       /* queries of the last query-vs-reference search */
    int64_t query_candidates;  /* candidate distances it evaluated */
    int64_t query_invalid_rows; /* its rows blanked by the reference's validity rule */
    float query_ms;            /* its device time (query sort + search) */
    int32_t query_k;           /* its neighbour count */
    float taubin_pass_ms;      /* last smoothing call: device time per lambda or mu pass (the passes alone) */
    float taubin_setup_ms;     /* its tile set-up and graph build (taubin_ms = setup + all passes) */
    int32_t taubin_overflow_tiles; /* tiles of that call whose halo only partly fit in LDS (the rest is read
                                      from global memory) */
    int64_t learned_bound_retries; /* last kNN: queries with fewer than k points within their proposed radius,
                                      searched again without the bound */
    int64_t knn_insertions;    /* last kNN: keys inserted into the neighbour lists */
    int64_t knn_max_candidates; /* last kNN: candidate distances of its costliest query */
    int64_t knn_long_queries;   /* last kNN: queries handed on to the wave-per-query pass (EGT_OPT_KNN_LONG_QUERY) */
} egt_stats;

/* EGT_OPT_START_LEVEL value: choose the level from the mean cell occupancy. */
#define EGT_START_LEVEL_OCCUPANCY 127
/* Entries of the autotuner's trial arrays: start levels -3 .. 10, entry i = level i - 3. */
#define EGT_AUTOTUNE_SLOTS 14

/* Options set through egt_set_option (no reference counterpart). */
typedef enum {
    EGT_OPT_START_LEVEL = 0,     /* kNN: pyramid level of the first search box; 0 = the query's own cell, as
                                    in the reference (default); > 0 coarser cells; -1, -2, -3 = cells of 1/2,
                                    1/4, 1/8 of a grid cell; EGT_START_LEVEL_OCCUPANCY = one level from the
                                    mean cell occupancy */
    EGT_OPT_LEARNED = 1,         /* kNN: 1 = start each query from the learned radius proposal */
    EGT_OPT_TAUBIN_GRAPH = 2,    /* Taubin: 1 = replay the lambda/mu loop as a captured hipGraph (default) */
    EGT_OPT_MAX_PTS_IN_CUBE = 3, /* kNN validity rule: cell population that invalidates a query (1e5) */
    EGT_OPT_KNN_PERSISTENT = 4,  /* kNN: 1 = one persistent work-queue launch per neighbour-count class,
                                    0 = one workgroup per 256 queries (default) */
    EGT_OPT_TAUBIN_TILE = 5,     /* Taubin: points per LDS tile, 1024 (default) or 2048; 0 = autotune on the
                                    next smoothing call (times both on copies, keeps the faster; the
                                    per-pass times land in egt_stats.taubin_tile_ms) */
    EGT_OPT_TAUBIN_REGULARIZED = 6, /* Taubin weights: 0 = exp(-d^2) (default, as the reference runs); 1 = the
                                      reference's dormant isRegularized variant, EGT_CUDA.cu:263, 544: weight
                                      1 in the lambda pass, 1 / (d^2 + 1e-8) in the mu pass */
    EGT_OPT_LEARNED_MODE = 7,    /* kNN with the learned proposal: bit 0 = start at the proposed level, bit 1 =
                                    bound the search by the proposed radius (queries with fewer than k points
                                    inside it are searched again without the bound); default 1 (the level only).
                                    Which is faster depends on the cloud and k (try both with
                                    egt_autotune_shape). Rows do not depend on it */
    EGT_OPT_KNN_LANES = 8,       /* kNN self search: lanes per query, 1 = one lane per query with the list in its
                                    registers (default), 64 = one wave per query with the list spread over the
                                    lanes. Rows do not depend on it */
    EGT_OPT_KNN_LONG_QUERY = 9,  /* kNN self search with one lane per query: a query that has tested more
                                    candidates than this is handed on to a wave-per-query pass (its wave no
                                    longer waits for it); 0 = never; default 2048. Rows do not depend on it */
    EGT_OPT_KNN_CLASS_SHAPES = 10, /* kNN self search: launch shape per neighbour-count class (k <= 8, 16, ..,
                                    64 are classes 0..7): bits 2i..2i+1 hold class i's shape + 1 (1 = one lane
                                    per query, 2 = the same as one persistent launch, 3 = one wave per query,
                                    0 = the handle-wide EGT_OPT_KNN_LANES / EGT_OPT_KNN_PERSISTENT; default 0).
                                    egt_batch_autotune_shape returns such a value. Rows do not depend on it */
    EGT_OPT_LEARNED_LEVEL_OFFSET = 11 /* kNN with the learned proposal: added to every proposed start level,
                                    -3..3 (default 0; +1 = one level coarser than proposed).
                                    egt_autotune_learned picks it. Rows do not depend on it */
} egt_option;

const char* egt_last_error_string(void);

/* GridStructure(const Mesh&, int num_cards, float eps, int bucketCount, int maxDivision)
 * GridKNNCuda.cuh:15-16, GridKNNCuda.cu:1165-1205 (+ InitStructure :1023-1133).
 * Host SoA input in original order. bucketCount must be a power of two in [1, 1024]. */
egt_status egt_grid_create(const float* x, const float* y, const float* z, int64_t n, int num_cards, float eps,
                           int bucket_count, int max_division, egt_handle* out);

/* Same, with x, y, z already in device memory on the current device (no reference counterpart: the
 * reference always copies from the host, GridKNNCuda.cu:1030-1040). The inputs are not modified. */
egt_status egt_grid_create_device(const float* dx, const float* dy, const float* dz, int64_t n, int num_cards,
                                  float eps, int bucket_count, int max_division, egt_handle* out);

/* GridStructure::GRIDCUDAKNNSELF_COMPACT(int k, float thres), GridKNNCuda.cuh:22, GridKNNCuda.cu:2322-2634.
 * thres in input units (main.cpp:109 passes 1e8). Rows are the exact k nearest neighbours ordered by
 * (distance, sorted index); rows the reference would blank are all -1. */
egt_status egt_grid_knn_self(egt_handle h, int k, float thres);

/* GridStructure::GRIDCUDAKNNCompact(float* qx, float* qy, float* qz, int pointsCountQuery, int k, float thres),
 * GridKNNCuda.cuh:19, GridKNNCuda.cu:1645-1939 (kernel construct_neighbors :340-440): the k nearest reference
 * points of each query. Host query arrays in input units and original order, finite coordinates. Queries are
 * normalized with the reference cloud's extent. Candidates need distance < thres (strict), and no point is
 * excluded. Rows are the exact k nearest reference points ordered by (distance, reference sorted index);
 * rows the reference would blank are all -1. For a query outside the reference's bounding cube, the
 * reference's ring schedule can skip cells and return a list that misses nearer points; this call returns the
 * exact list there. The query rows are kept beside the self rows: smoothing and normals keep using the self
 * rows. */
egt_status egt_grid_knn_query(egt_handle h, const float* qx, const float* qy, const float* qz, int64_t nq, int k,
                              float thres);
/* As egt_grid_knn_query with the query coordinates already in device memory (no reference counterpart:
 * the reference always copies from the host, GridKNNCuda.cu:1693-1703). */
egt_status egt_grid_knn_query_device(egt_handle h, const float* qx, const float* qy, const float* qz, int64_t nq,
                                     int k, float thres);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc): egt_grid_knn_query_device without any allocation.
 * rows: caller's device buffer of nq*k int32 that receives the rows egt_get_device_query_knn would hand out
 * (original query order, reference sorted indices, -1 holes); the handle's own query rows and stats are left
 * as they were. At most 1,048,576 queries per call (larger counts return EGT_ERR_UNSUPPORTED). Temporaries live
 * in a fixed device array of 56 MiB per GPU, allocated when the library's code loads on that GPU and held until
 * the process ends; calls that overlap on one GPU (two threads, or two streams in flight) race on it. The work
 * is queued on the handle's stream (egt_get_stream) without a host wait: the rows are ready when that stream
 * completes. Needs the handle's device to be current and an up-to-date search structure (EGT_ERR_STATE after
 * smoothing moved the points, until a query or self search rebuilds it). */
egt_status egt_grid_knn_query_device_noalloc(egt_handle h, const float* qx, const float* qy, const float* qz,
                                             int64_t nq, int k, float thres, int32_t* rows);
/* GridStructure::GetKNNIndexes(int* knn) after a query search, GridKNNCuda.cu:1941-1968: host nq*k int32,
 * original query order, original reference indices, -1 holes. */
egt_status egt_get_knn_query_original(egt_handle h, int32_t* out);
/* KNNInterface::GetKNNIndexesInPartitions() after a query search, KNNInterface.cuh:20 (device rows, reference
 * sorted indices). Borrowed pointer valid until the next query search or egt_destroy. The rows are in
 * original query order (the reference keeps them in its own query sort order, GridKNNCuda.cu:1208-1267).
 * nq and k (optional) receive the row count and width; pointsQuerCount, KNNInterface.cuh:29. */
egt_status egt_get_device_query_knn(egt_handle h, int32_t** rows, int64_t* nq, int32_t* k);

/* GridMeshCuda(const Mesh&, int num_cards, float eps, int bucketCount), GridMeshCuda1.cuh:16,
 * GridMeshCuda1.cu:247-414. tri: ntri x 3 uint32 original vertex indices (Mesh.h:15-19). */
egt_status egt_mesh_create(const float* x, const float* y, const float* z, int64_t n, const uint32_t* tri,
                           int64_t ntri, int num_cards, float eps, int bucket_count, egt_handle* out);

/* GridMeshCuda::GetNeighborsFromTriangles(int k), GridMeshCuda1.cuh:17, GridMeshCuda1.cu:416-511. */
egt_status egt_mesh_neighbors(egt_handle h, int k);

/* EGT_CUDA(KNNInterface&).PerformSmoothing(int iterationCount, float lambda, float mu, float alpha),
 * EGT_CUDA.cuh:13-14, EGT_CUDA.cu:476-648. Side effects as in the reference: the handle's neighbour rows
 * become the EGG-filtered rows and its points become the smoothed points (still sorted, normalized). */
egt_status egt_smooth(egt_handle h, int iterations, float lambda, float mu, float alpha);

/* Normal_CUDA(KNNInterface&).GetNormals(float* nx, float* ny, float* nz), Normal_CUDA.cuh:12-13,
 * Normal_CUDA.cu:304-775. Host arrays of n floats, original order. */
egt_status egt_normals(egt_handle h, float* nx, float* ny, float* nz);

/* getOutputPoints(float* x, float* y, float* z, KNNInterface*), getOutputPoints.cuh:10,
 * getOutputPoints.cu:34-63. Host arrays, original order, de-normalized. Unlike the reference this does not
 * overwrite the handle's sorted points (:52-54), so the handle stays usable. */
egt_status egt_get_output_points(egt_handle h, float* x, float* y, float* z);

/* KNNInterface::GetKNNIndexes(int* knnIndexes), KNNInterface.cuh:23, GridKNNCuda.cu:1941-1968,
 * GridMeshCuda1.cu:513-540. Host n*k int32, original index space, -1 holes. */
egt_status egt_get_knn_original(egt_handle h, int32_t* out);

/* KNNInterface::GetRefPointsX/Y/Z() -> float**, KNNInterface.cuh:10-12 (card 0). Borrowed device pointers. */
egt_status egt_get_device_points(egt_handle h, float** x, float** y, float** z);

/* KNNInterface::GetKNNIndexesInPartitions() -> int**, KNNInterface.cuh:18 (card 0). Borrowed, n*k, sorted. */
egt_status egt_get_device_knn(egt_handle h, int32_t** knn);

/* KNNInterface::GetOriginalRefIndexes() -> int**, KNNInterface.cuh:21 (card 0): sorted -> original. */
egt_status egt_get_device_original_indexes(egt_handle h, int32_t** idx);

/* KNNInterface::GetMinExtent/GetMaxExtent, KNNInterface.cuh:24-25. */
egt_status egt_get_extent(egt_handle h, float* min_extent, float* max_extent);

/* KNNInterface::pointsRefCount/NeighborCount/GetNumberOfCards/GetPartitionSize, KNNInterface.cuh:9,19,26,28. */
egt_status egt_get_counts(egt_handle h, int64_t* n_points, int32_t* k, int32_t* num_cards,
                          int64_t* partition_size);

/* Counters and stage timings (no reference counterpart). */
egt_status egt_get_stats(egt_handle h, egt_stats* st);

/* Tuning switches (no reference counterpart). */
egt_status egt_set_option(egt_handle h, egt_option opt, int64_t value);

/* Device memory cache (no reference counterpart; the reference allocates and frees per call). Memory a
 * destroyed handle held stays cached per device for later handles, up to EGT_DEVICE_CACHE_MB (environment,
 * default 32768) per device. This call returns every cached block to the driver; *released (optional)
 * receives the bytes returned. *cached (optional) receives the bytes cached before the call. */
egt_status egt_release_cached_memory(uint64_t* released, uint64_t* cached);

/* Autotuner (no reference counterpart): runs the plain search with neighbour count k and radius thres
 * (input units, as egt_grid_knn_self) on an evenly spread sample of queries for each candidate start level
 * (the cell size of the first search box), scores each level by its measured work per query (candidates
 * tested + 3 x cells looked up) and sets EGT_OPT_START_LEVEL to the cheapest. Rows do not depend on the
 * choice. trial_ms and trial_cost (optional, EGT_AUTOTUNE_SLOTS floats each) receive each level's sample time
 * in ms and its score at index level + 3, -1 where untried. The handle's neighbour rows are not valid
 * afterwards: run egt_grid_knn_self. */
egt_status egt_autotune(egt_handle h, int k, float thres, int32_t* start_level, float* trial_ms,
                        float* trial_cost);

/* Launch-shape tuner (no reference counterpart): times the handle's current search (its start level and, with
 * EGT_OPT_LEARNED, the proposal in place) with neighbour count k and radius thres over all queries in each
 * launch shape: 0 = one lane per query, one workgroup per 256 queries; 1 = one lane per query, one
 * persistent work-queue launch; 2 = one wave per query. Each shape is
 * warmed up on the autotuner's sample, then timed once on the full search; the fastest is kept through
 * EGT_OPT_KNN_PERSISTENT and EGT_OPT_KNN_LANES. Rows do not depend on the choice. *shape (optional) receives
 * it, shape_ms (optional, 3 floats) each shape's time in ms. The handle's neighbour rows
 * are not valid afterwards: run egt_grid_knn_self. */
egt_status egt_autotune_shape(egt_handle h, int k, float thres, int32_t* shape, float* shape_ms);

/* Learned-start tuner (no reference counterpart): with a proposal in place (egt_model_propose), times the full
 * search with neighbour count k and radius thres for each level offset (0, +1), each use of the proposal
 * (EGT_OPT_LEARNED_MODE 1 = its level, 3 = its level and radius bound) and each of the lane-per-query and
 * wave-per-query shapes, each after a warm-up on the autotuner's sample, and keeps the fastest through
 * EGT_OPT_LEARNED_LEVEL_OFFSET, EGT_OPT_LEARNED_MODE and EGT_OPT_KNN_LANES. Rows do not depend on the choice.
 * offset, mode, shape (optional) receive it (shape 0 = lane per query, 2 = wave per query); ms (optional, 8
 * floats) the times, index (offset x 2 + (mode == 3)) x 2 + (shape == 2). One-cloud handles. The handle's
 * neighbour rows are not valid afterwards: run egt_grid_knn_self. */
egt_status egt_autotune_learned(egt_handle h, int k, float thres, int32_t* offset, int32_t* mode, int32_t* shape,
                                float* ms);

/* Per-class launch-shape tuner for a batch (no reference counterpart): with neighbour counts k[c] per cloud
 * and radius thres (as egt_batch_knn) and the handle's start level or learned proposal in place, times each
 * neighbour-count class's full search in each launch shape (as egt_autotune_shape: lane per query,
 * persistent, wave per query), each after a warm-up launch, and keeps the fastest per class. *shapes
 * (optional) receives the choices packed as EGT_OPT_KNN_CLASS_SHAPES takes them, for later handles; class_ms
 * (optional, 24 floats) each class's time per shape (index 3 class + shape), -1 where not run. The handle's
 * neighbour rows are not valid afterwards: run egt_batch_knn. */
egt_status egt_batch_autotune_shape(egt_handle h, const int32_t* k, float thres, int64_t* shapes, float* class_ms);

/* Provider destructors, GridKNNCuda.cu:1005-1021, GridMeshCuda1.cu:542-551; also frees what EGT_CUDA
 * leaked (EGT_CUDA.cu:490-492). */
egt_status egt_destroy(egt_handle h);

/* HIP stream of the handle (hipStream_t), for callers that order their own work after the library's. */
void* egt_get_stream(egt_handle h);

/* ---- Batches of independent clouds (no reference counterpart: the reference handles one cloud per
 * provider). x, y, z hold the clouds one after another (host, each in its own original order); cloud c is
 * points offsets[c] .. offsets[c + 1] - 1 (nclouds + 1 offsets, offsets[0] = 0). Every cloud is normalized,
 * sorted and searched exactly as if it were its own GridStructure, so results equal separate runs, but each
 * stage runs over all clouds at once: the kNN as one persistent work-queue launch per neighbour-count class,
 * the EGG filter as one launch per class, Taubin as one graph over all clouds.
 * On a batch handle egt_smooth, egt_normals and egt_get_output_points work as on one cloud and use the
 * concatenated order of the input; egt_grid_knn_self and egt_get_knn_original are replaced below. */
egt_status egt_batch_create(const float* x, const float* y, const float* z, const int64_t* offsets, int32_t nclouds,
                            float eps, int bucket_count, egt_handle* out);
/* Same with x, y, z already in device memory on the current device (offsets stay on the host). */
egt_status egt_batch_create_device(const float* dx, const float* dy, const float* dz, const int64_t* offsets,
                                   int32_t nclouds, float eps, int bucket_count, egt_handle* out);
/* GRIDCUDAKNNSELF_COMPACT for every cloud: k[c] neighbours (1..64) for cloud c, radius thres in the input
 * units of each cloud (GridKNNCuda.cu:2322-2634 per cloud). */
egt_status egt_batch_knn(egt_handle h, const int32_t* k, float thres);
/* GetKNNIndexes for one cloud (GridKNNCuda.cu:1941-1968): n_c x k_c int32 in that cloud's own original
 * order and index space, -1 holes. */
egt_status egt_batch_get_knn_original(egt_handle h, int32_t cloud, int32_t* out);
/* Point count, neighbour count of the last search and bounding cube of one cloud. */
egt_status egt_batch_cloud_info(egt_handle h, int32_t cloud, int64_t* n, int32_t* k, float* min_extent,
                                float* max_extent);

/* ---- Learned search-radius proposal (no reference counterpart; the reference starts every query at its
 * own cell, GridKNNCuda.cu:2409-2474). A 28 -> 64 -> 64 -> 1 ReLU MLP maps the log-normalised occupancy of
 * the 3x3x3 cells around a query (at the level the plain search would start from) plus log2(k) to
 * log2(k-th neighbour distance / cell edge); the kNN then starts each query at the smallest level whose
 * box reaches that radius plus a safety margin. Its exactness test expands past any proposal that is too
 * small, so neighbour rows are identical to the plain search's. Forward and backward run on BF16 MFMA with
 * FP32 accumulation; Adam (beta1 0.9, beta2 0.999, eps 1e-8) updates FP32 master weights.
 * Parameter vector (FP32, egt_model_param_count() entries): W1[32][64] (input-major; rows 28..31 are
 * padding and kept at zero), b1[64], W2[64][64], b2[64], W3[64], b3.
 * Feature rows passed in or out are egt_model_feature_count() = 28 floats each. */
typedef struct egt_model_s* egt_model;

int32_t egt_model_param_count(void);
int32_t egt_model_feature_count(void);
/* New model on the current device, weights drawn He-uniform from seed. */
egt_status egt_model_create(uint64_t seed, egt_model* out);
egt_status egt_model_destroy(egt_model m);
egt_status egt_model_get_params(egt_model m, float* params, int64_t count);
/* Sets the weights and clears the Adam state. */
egt_status egt_model_set_params(egt_model m, const float* params, int64_t count);
/* Forward and gradient of the mean squared error over one batch (host arrays): y[n], grad[param_count]
 * and loss are each optional. */
egt_status egt_model_eval(egt_model m, const float* features, const float* targets, int64_t n, float* y,
                          float* grad, float* loss);
/* One Adam step on a host batch; grad (optional) receives the gradient it used. */
egt_status egt_model_adam_step(egt_model m, const float* features, const float* targets, int64_t n, float lr,
                               float* grad, float* loss);
/* Trains on a cloud: runs the plain exact search with k on h (its neighbour rows become that search's),
 * takes up to max_samples valid queries (0 = all), runs epochs of Adam over batches of batch samples, and
 * sets the safety margin to the given quantile of the training residuals. loss = final training MSE. */
egt_status egt_model_train(egt_model m, egt_handle h, int k, int epochs, int batch, int64_t max_samples, float lr,
                           float quantile, float* loss);
/* The same on a batch handle (egt_batch_create), k[c] neighbours for cloud c: one model learns all the
 * clouds and neighbour counts of the batch. */
egt_status egt_model_train_batch(egt_model m, egt_handle h, const int32_t* k, int epochs, int batch,
                                 int64_t max_samples, float lr, float quantile, float* loss);
egt_status egt_model_get_safety(egt_model m, float* safety);
egt_status egt_model_set_safety(egt_model m, float safety);
/* Margin (log2 units) of the radius bound: the bound is the predicted radius times 2^margin. Training sets it
 * to the 0.99 quantile of the residuals (at least the level margin) plus 0.6, so that about 0.05 % of the
 * queries fall outside their bound and are searched again. */
egt_status egt_model_get_bound_safety(egt_model m, float* safety);
egt_status egt_model_set_bound_safety(egt_model m, float safety);
/* Proposes a start level and a radius bound for every query of h for neighbour count k and switches h to
 * learned starts (EGT_OPT_LEARNED = 1); the next egt_grid_knn_self uses them. */
egt_status egt_model_propose(egt_model m, egt_handle h, int k);
/* Batch handle: proposals for every cloud with k[c] neighbours; the next egt_batch_knn uses them. */
egt_status egt_model_propose_batch(egt_model m, egt_handle h, const int32_t* k);
/* Diagnostics: feature rows (n x 28, sorted order), targets of the handle's last search with this k (NaN
 * where invalid or absent) and the reference level; each output optional. */
egt_status egt_model_features(egt_handle h, int k, float* features, float* targets, int32_t* ref_level);
/* Diagnostics: the proposed start level of every point (sorted order). */
egt_status egt_model_levels(egt_handle h, int8_t* levels);

#ifdef __cplusplus
}
#endif

#endif /* EGT_H */
