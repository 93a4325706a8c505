// Internal declarations of the gfx950 library behind include/egt.h.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/egt.h"
#include "egt_device.h"
#include "egt_limits.h"

#define EGT_MAX_LEVELS 11
static_assert(EGT_AUTOTUNE_SLOTS == EGT_MAX_LEVELS + egt::SUB, "autotuner slots: sub-cell levels + pyramid levels");

namespace egt {

struct HipError {
    hipError_t err;
    std::string what;
};
struct ArgError {
    egt_status st;
    std::string what;
};

#define HIP_CHECK(expr)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::egt::HipError{e_, std::string(#expr) + ": " + hipGetErrorString(e_)};       \
    } while (0)

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// A HIP event owned for one scope (stage timing): destroyed on every exit path, exceptions included.
struct Event {
    hipEvent_t e = nullptr;
    Event() { HIP_CHECK(hipEventCreate(&e)); }
    ~Event() {
        if (e) (void)hipEventDestroy(e);
    }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    operator hipEvent_t() const { return e; }
};

// An instantiated graph owned for one scope.
struct GraphExec {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec&) = delete;
    GraphExec& operator=(const GraphExec&) = delete;
    ~GraphExec() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
};

// Stream capture for one scope: if anything throws between begin and end, the destructor ends the capture
// (and drops the partial graph), so the stream leaves capture mode and the handle stays usable.
struct Capture {
    hipStream_t s;
    bool active = false;
    explicit Capture(hipStream_t st) : s(st) {
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        active = true;
    }
    Capture(const Capture&) = delete;
    Capture& operator=(const Capture&) = delete;
    // Ends the capture into out.g and instantiates out.ge.
    void end(GraphExec& out) {
        active = false;
        HIP_CHECK(hipStreamEndCapture(s, &out.g));
        HIP_CHECK(hipGraphInstantiate(&out.ge, out.g, nullptr, nullptr, 0));
    }
    ~Capture() {
        if (!active) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
    }
};

// Device memory cache (egt_memory.hip): device_alloc returns a block of at least `bytes` (its capacity in
// *capacity), reusing a block an earlier handle released; device_release hands a block back (the caller
// guarantees no work in flight uses it).
void* device_alloc(size_t bytes, size_t* capacity);
void device_release(void* p, size_t capacity);
size_t device_cache_release_all();
size_t device_cache_bytes();

// Device allocation owned by a handle.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;    // elements usable
    size_t cap = 0;  // bytes of the block
    void alloc(size_t count) {
        if (count <= n && p) return;
        free();
        if (count == 0) count = 1;
        p = static_cast<T*>(device_alloc(count * sizeof(T), &cap));
        n = cap / sizeof(T);
    }
    void free() {
        if (p) {
            (void)hipDeviceSynchronize();  // never release memory an in-flight kernel may still touch
            device_release(p, cap);
        }
        p = nullptr;
        n = 0;
        cap = 0;
    }
    ~DBuf() { free(); }
};

// One level of the cell pyramid: open-addressing hash from Morton cell code to (start, count) in the
// Morton-ordered point array.
struct LevelTable {
    uint32_t log2cap = 0;
    uint64_t offset = 0;  // slot offset in the shared table
    int32_t max_count = 0;
    int64_t cells = 0;
};

// Per-cloud parameters (a single-cloud handle has one entry). Sorted space and Morton order both hold the
// clouds one after another in input order, so a cloud's points are positions [base, base + n) in both.
struct CloudDev {
    int64_t base;     // first sorted index / Morton position of the cloud
    int64_t n;        // point count
    int64_t row_off;  // first element of the cloud's neighbour rows (rows are n x k, cloud after cloud)
    int32_t k;        // neighbour count of the last search
    float thres_n;    // squared search radius in normalized units
    float min_extent, extent;  // de-normalization: x = x' * extent + min_extent
    int32_t ref_level;         // learned proposal: feature level
    float max_extent;          // bounding cube upper end (as computed, not min_extent + extent)
};

// A neighbour-count class of a batch: its K template and its range in the chunk work list.
struct WorkClass {
    int K;
    int64_t first, count;
};

// Reusable temporaries: 0-2 scan partials, 3 scan total, 4-7 radix sort, 8-15 stage code (15 = k-th
// distance of the last search), 16 host-input staging, 17-23 EGG / Taubin, 24-26 learned proposal,
// 27-31 Taubin sliced-ELLPACK tiles, 32-35 query-vs-reference search, 36 64-bit totals, 37 queries handed
// back by a bounded kNN search, 38 the level tables' cell starts.
struct Scratch {
    DBuf<uint32_t> bufs[39];
    template <class T>
    T* get(int i, size_t count) {
        bufs[i].alloc((count * sizeof(T) + sizeof(uint32_t) - 1) / sizeof(uint32_t));
        return reinterpret_cast<T*>(bufs[i].p);
    }
};

}  // namespace egt

struct egt_ctx {
    int kind = 0;  // 0 cloud, 1 mesh
    hipStream_t stream = nullptr;
    int device = 0;
    int64_t n = 0;
    int k = 0;
    int B = 1024;
    float eps = 1e-6f;
    float min_extent = 0, max_extent = 0;
    bool have_knn = false, smoothed = false;

    // sorted space (reference layout)
    egt::DBuf<float> x, y, z;
    egt::DBuf<int32_t> perm;  // sorted -> original
    egt::DBuf<int32_t> key;   // fine row-major cell key of each sorted point
    egt::DBuf<int32_t> knn;   // n*k rows

    // search structure (Morton order)
    egt::DBuf<float4> mpts;      // x, y, z, bits(sorted index)
    egt::DBuf<uint32_t> mcell;   // packed fine cell coords cx | cy<<10 | cz<<20
    egt::DBuf<uint64_t> mvcell;  // virtual cell and sub code (vcell_pack)
    egt::DBuf<uint32_t> mcode;   // fine Morton code
    egt::DBuf<int32_t> morder;   // Morton position -> sorted index
    egt::DBuf<int32_t> minv;     // sorted index -> Morton position
    bool search_stale = false;   // points moved since the search structure was built
    egt::DBuf<uint4> htab;       // all level tables: (key = cell_key(cloud, level code), start, count) per slot
    egt::LevelTable lev[EGT_MAX_LEVELS];
    int nlev = 0;
    egt::DBuf<egt::LevelDev> dlev;  // device copy of the level descriptors
    std::vector<int64_t> cloud_cells;  // non-empty cells per cloud and level [cloud * nlev + level]

    // clouds (a batch handle holds several independent clouds; a plain handle holds one)
    int nclouds = 1;
    std::vector<egt::CloudDev> clouds;  // host copy
    egt::DBuf<egt::CloudDev> dclouds;   // device copy
    egt::DBuf<int32_t> mcloud;          // cloud of each Morton position (batch only, else empty)
    egt::DBuf<int32_t> scloud;          // cloud of each sorted index (batch only)
    egt::DBuf<int64_t> in_off;          // input offsets of the clouds (batch only)
    egt::DBuf<int64_t> work_lists;      // kNN / EGG work lists of 64-point chunks (batch only)
    std::vector<egt::WorkClass> work_classes;

    // EGG adjacency as CSR in Morton order (rows and column indices are Morton positions)
    egt::DBuf<int32_t> rowptr, col;
    int64_t nnz = 0;
    bool csr_valid = false;  // CSR matches the current neighbour rows
    int64_t taubin_fallback_tiles = 0;

    // mesh
    egt::DBuf<uint32_t> tri;
    int64_t ntri = 0;

    // learned proposal
    egt::DBuf<int8_t> qlevel;   // per Morton position: proposed start level
    egt::DBuf<float> qbound;    // per Morton position: proposed radius bound (squared, normalized units)

    // query-vs-reference search (its rows live beside the self rows, which smoothing keeps using)
    egt::DBuf<float4> qpts;       // queries in Morton order of their cell: x', y', z', bits(original index)
    egt::DBuf<uint64_t> qvcell;   // virtual cell (vcell_pack), clamped to the grid
    egt::DBuf<int32_t> qknn;      // nq*kq rows in original query order, reference sorted indices
    egt::DBuf<egt::CloudDev> dqcloud;
    int64_t nq = 0;
    int kq = 0;
    bool have_query = false;

    // options
    int start_level = 0;  // reference schedule: every query starts at its own level-0 cell; -SUB..-1 sub-cell
                          // boxes; EGT_START_LEVEL_OCCUPANCY = from the mean cell occupancy
    int learned = 0;
    int learned_mode = 1;  // with learned: 1 = start at the proposed level, 2 = bound the search by the proposed
                           // radius, 3 = both
    int knn_persistent = 0;  // kNN as one persistent work-queue launch per neighbour-count class
    int knn_lanes = 1;       // kNN self search: 1 = one lane per query, 64 = one wave per query
    int learned_level_offset = 0;  // added to the proposed start levels
    int class_shape[8] = {-1, -1, -1, -1, -1, -1, -1, -1};  // per neighbour-count class (k <= 8, 16, .., 64):
                                                           // 0 lane, 1 persistent, 2 wave; -1 handle-wide
    int64_t knn_long_cap = 2048;  // lane-per-query search: candidates after which a query goes to a wave
                                  // (0 = never)
    int taubin_graph = 1;
    int taubin_tile = 1024;  // Taubin tile width: 1024 or 2048 points; 0 = tune on the next smoothing call
    int taubin_regularized = 0;  // isRegularized weights (1 / inverse-square), off as in the reference
    int maxc = 100000;

    egt_stats stats{};
    egt::Scratch scratch;
};

// Learned proposal model (egt_mlp.hip): FP32 master weights, Adam moments, gradient partials, training set.
struct egt_model_s {
    egt::DBuf<float> P, M, V, grad, part, ploss, Y;
    egt::DBuf<float> X, T;       // host-supplied batch
    egt::DBuf<uint32_t> sel;     // training sample order
    int step = 0;
    int last_parts = 0;
    float safety = 0.0f;        // level choice: margin (log2 units) added to the predicted radius
    float bound_safety = 0.0f;  // radius bound: margin added to the predicted radius
    int64_t train_samples = 0;
    uint64_t rng = 0x243F6A8885A308D3ull;
    hipStream_t stream = nullptr;
    int device = 0;
};

namespace egt {
// scan / sort / reduce utilities (egt_prims.hip)
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, Scratch& sc,
                        uint32_t* total_host);
// Same, with the total written to device memory and no synchronization.
void exclusive_scan_u32_dev(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, Scratch& sc,
                            uint32_t* total_dev);
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s, Scratch& sc);
// The same sort over a caller's workspace of radix_sort_workspace_words(n) 32-bit words (no allocation).
size_t radix_sort_workspace_words(int64_t n);
void radix_sort_pairs_ws(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s, uint32_t* ws);
// Sum of n uint32 values in 64 bits (synchronizes the stream).
uint64_t sum_u32(const uint32_t* in, int64_t n, hipStream_t s, Scratch& sc);
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z);

// grid (egt_grid.hip)
void build_grid(egt_ctx* h, const float* dx, const float* dy, const float* dz);
void build_grid_batch(egt_ctx* h, const float* dx, const float* dy, const float* dz, const int64_t* offsets,
                      int nclouds);
void build_search_structure(egt_ctx* h);
void refresh_search_structure(egt_ctx* h);

// kNN (egt_knn.hip)
void run_knn(egt_ctx* h, int k, float thres_n);
void run_knn_query(egt_ctx* h, const float* dqx, const float* dqy, const float* dqz, int64_t nq, int k,
                   float thres_n);
void run_knn_query_original(egt_ctx* h, int32_t* out);
void run_knn_clouds(egt_ctx* h);
int auto_start_level(const egt_ctx* h, int k, int cloud = 0);
int autotune_start_level(egt_ctx* h, int k, float thres_n, float* trial_ms, float* trial_cost);
// Times the current search in each launch shape on a sample and keeps the fastest: 0 = one lane per query,
// 1 = the same as one persistent launch, 2 = one wave per query. shape_ms: 3 sample times, -1 if not offered.
int autotune_shape(egt_ctx* h, int k, float thres_n, float* shape_ms);
// The same per neighbour-count class of the handle's clouds (k and radius already set): ms is 8 x 3, the
// return value the choices packed as EGT_OPT_KNN_CLASS_SHAPES takes them.
int64_t autotune_class_shapes(egt_ctx* h, float* ms);
// With a proposal in place: level offset (0, 1) x proposal use (1, 3) x shape (lane, wave), full searches timed;
// keeps the fastest in the handle's options. ms: 8 times, index (offset * 2 + mode index) * 2 + shape index.
void autotune_learned(egt_ctx* h, int k, float thres_n, int* best_off, int* best_mode, int* best_shape, float* ms);

// learned proposal (egt_mlp.hip)
int mlp_param_count();
int mlp_feature_count();
void mlp_init(egt_model_s* m, uint64_t seed);
void mlp_set_params(egt_model_s* m, const float* params);
void mlp_get_params(egt_model_s* m, float* params);
void mlp_eval(egt_model_s* m, const float* X, const float* T, int64_t n, float* y, float* grad, float* loss);
void mlp_adam_step(egt_model_s* m, const float* X, const float* T, int64_t n, float lr, float* grad, float* loss);
void mlp_train(egt_model_s* m, egt_ctx* h, const int* ks, int epochs, int batch, int64_t max_samples, float lr,
               float quantile, float* loss);
void mlp_propose(egt_model_s* m, egt_ctx* h, const int* ks);
void mlp_features(egt_ctx* h, int k, float* X, float* T, int* ref_level);
void mlp_levels(egt_ctx* h, int8_t* out);

// EGG + CSR (egt_egg.hip)
void run_egg(egt_ctx* h, float alpha);
void build_csr_from_rows(egt_ctx* h);

// Taubin (egt_taubin.hip)
void run_taubin(egt_ctx* h, int iters, float lambda, float mu);

// mesh (egt_mesh.hip)
void run_mesh_adjacency(egt_ctx* h, int k);

// normals (egt_normals.hip)
void run_normals(egt_ctx* h, float* nx_host, float* ny_host, float* nz_host);

// output (egt_grid.hip)
void run_output(egt_ctx* h, float* x, float* y, float* z);
void run_knn_original(egt_ctx* h, int32_t* out);
void run_knn_original_cloud(egt_ctx* h, int cloud, int32_t* out);

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Query search without allocation (egt_grid_knn_query_device_noalloc): at most QNA_MAX_QUERIES queries per
// call; its temporaries live in a file-scope device array of QNA_WORDS 32-bit words (one copy per GPU).
constexpr int64_t QNA_MAX_QUERIES = (int64_t)1 << 20;
constexpr size_t QNA_WORDS = 14 * (size_t)QNA_MAX_QUERIES + 64;
void run_knn_query_noalloc(egt_ctx* h, const float* dqx, const float* dqy, const float* dqz, int64_t nq, int k,
                           float thres_n, int32_t* rows);

// One-cloud handles: mirror the handle's k into the cloud table (rows start at 0) and upload it.
inline void sync_single_cloud(egt_ctx* h) {
    if (h->nclouds != 1) return;
    h->clouds[0].k = h->k;
    h->clouds[0].row_off = 0;
    h->dclouds.alloc(1);
    HIP_CHECK(hipMemcpyAsync(h->dclouds.p, h->clouds.data(), sizeof(CloudDev), hipMemcpyHostToDevice, h->stream));
}
}  // namespace egt
