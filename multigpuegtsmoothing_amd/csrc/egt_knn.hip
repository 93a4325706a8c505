// Exact self kNN over the cell pyramid, with the reference's row-validity rule.
//
// Reference: GridStructure::GRIDCUDAKNNSELF_COMPACT, GridKNNCuda.cu:2322-2634, kernel
// construct_neighbors_self_compact :647-743, insertion appendDistanceGrid :311-339, certificate :694-741,
// invalidation :217-226, :287-309.
//
// One thread per query, queries walked in Morton order so a wave's 64 queries are spatial neighbours and
// share cache lines of candidate tiles. The search box at pyramid level l is the 3x3x3 block of level-l
// cells around the query's cell (cells of level l are contiguous ranges of the Morton-ordered point array,
// found through a per-level hash). Below the reference's fine cells the pyramid has SUB sub-cell levels
// (cells of 1/2, 1/4, 1/8 of a fine cell), whose ranges are found by binary search of the points' sub codes
// inside their fine cell's range; the search runs on internal levels il = l + SUB. Each level scans only
// cells and points outside the previous box, keeps
// a register-resident sorted top-K of 64-bit keys (distance bits << 32 | sorted index) and stops when the
// k-th distance plus a 2e-6 margin (double) is below every interior face distance, so rows are the exact
// k nearest neighbours ordered by (distance, sorted index). The reference's float certificate is then
// replayed on that exact list, stage by stage (own cell, then ring 1 at bucket counts B, B/2, ..., 8), to
// decide whether the reference would have blanked the row (cell with >= maxc points, or still searching
// after the bucket-count-8 ring).
//
// Two launch shapes run this search: one lane per query (k_knn, the list in the lane's registers) and one
// wave per query (k_knn_wave, the list spread over the lanes, one key per lane), chosen per neighbour-count
// class by the options or the tuners below. A lane-per-query search hands queries that exceed a candidate
// cap, and queries with fewer than k points inside a learned radius bound, to a wave-per-query pass.
#include <algorithm>
#include <mutex>
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int KBLK = 256;
// Cells with more points than this are scanned child by child. Measured on a 4M-point torus at start level 2
// (k=40: 14.0 ms at 48, 12.2 at 64, 11.9 at 96, 12.0 at 128; k=64: 31.7 at 48, 26.5 at 96) and on a
// 16M-point cluster mixture from the own cell (76.2 ms at 48, 77.1 at 96; 78.6 with 96 above the fine level
// and 48 below). With the final kernels on the 16M mixture at k=40 (own cell / learned proposal, the tuner's
// fastest mode): 60.3 / 53.8 ms at 64, 62.0 / 52.3 at 96, 63.9 / 55.0 at 192, 67.3 / 61.1 at 384; the
// learned start, which the mixture runs with, settles it at 96 (torus level 2: 8.34, 8.24, 8.38, 8.49).
constexpr uint32_t SPLIT_COUNT = 96;
// The same for the wave-per-query search. Its batches of 64 and constant-cost lane map scan a populous cell
// as cheaply as its children, so it splits only much larger cells. On a 16M-point cluster mixture:
//   threshold         96    192    384   1024   4096  never
//   k=40, own cell   69.9   66.9   62.5   59.8   60.5   65.9 ms
//   k=64, own cell   81.7   79.7   75.8   74.8   75.8   84.9 ms
//   k=64, learned    59.4   57.4   57.5   57.5   57.6   67.0 ms
//   512-cloud batch, learned (the class tuner's best total over all classes):
//                   127.8      -  127.4  130.5      -      - ms
// (k=40 from learned starts and a 4M-point torus at level 2: within 1 %; same rows; the columns come from
// four boxes, matched through their common runs). The sparse batch clouds start at coarse levels from
// learned proposals, where splitting lets the radius bound prune whole children: 384 keeps the batch and
// most of the gain from the own cell.
constexpr uint32_t WAVE_SPLIT_COUNT = 384;
// Wave-per-query search: a batch with more survivors than this is sorted and merged into the list, fewer are
// inserted one at a time. Measured on a 16M-point cluster mixture at 16 / 8 / 32 / never: k=64 from learned
// starts 59.0 / 60.0 / 60.0 / 75.4 ms, from the own cell 81.7 / 82.4 / 82.4 / 89.9; k=40 from the own cell
// 69.7 / 70.1 / 70.1 / 72.8, from learned starts 60.4 / 60.0 / 59.9 / 56.6 (there a lane per query is faster
// still); a 4M-point torus at level 2, k=64: 11.0 / 11.2 / 11.2 / 15.2 ms.
constexpr int WAVE_SORT_MIN = 16;
// Keys (distance bits << 32 | sorted index) of non-negative distances, read as IEEE doubles, are positive
// and order exactly as the integers do, so the list is held as doubles and merged with FP64 min / max (full
// rate on gfx950; keys of zero distance are denormal doubles, which compute kernels keep). The empty key is
// the largest finite double, above every real key (a real key's high word is below 0x7F800000).
constexpr uint64_t KEY_EMPTY = 0x7FEFFFFFFFFFFFFFull;
__device__ inline double key_d(uint64_t k) { return __builtin_bit_cast(double, k); }
__device__ inline uint64_t key_u(double d) { return __builtin_bit_cast(uint64_t, d); }
// Per-lane buffer of keys awaiting insertion (4) and candidates fetched per step (4). Measured on a 4M-point
// torus at start level 2 and on a 16M-point mixture with the FP64 insertion: k=40 8.3 ms with a buffer of 4
// against 8.7 with 8, 9.0 with 2 and 13.9 with 16 (register pressure); fetching 2 or 8 per step runs slower
// (9.0, 14.0 ms); k=16 4.4 ms fetching 4 against 4.7 fetching 2. Deferred flushes (a buffer of 5, 6 or 8
// flushed once a lane holds more than 1, 2 or 4 keys; or fetching 2) insert as many keys and run slower on the
// mixture's learned path too: 50.0-53.6 ms against 49.8.
template <int K>
constexpr int knn_buffer() { return 4; }
template <int K>
constexpr int knn_fetch() { return 4; }

struct KnnParams {
    const float4* __restrict__ mpts;
    const uint32_t* __restrict__ mcell;    // fine cell cx | cy << 10 | cz << 20 of each Morton position
    const unsigned long long* __restrict__ mvcell;  // virtual cell + sub code of each Morton position
    const uint4* __restrict__ htab;
    const int32_t* __restrict__ key;       // sorted-space fine key (member cell coordinates)
    const int8_t* __restrict__ qlevel;     // per-query start level (Morton order) or null
    const float* __restrict__ qbound;      // per-query radius bound, squared normalized units (Morton order), or
                                           // null: the list starts full of placeholder keys at that distance
    const CloudDev* __restrict__ clouds;   // per-cloud k, radius, row offset
    const int32_t* __restrict__ mcloud;    // cloud of each Morton position, null for one cloud
    const int64_t* __restrict__ chunks;    // work list: first Morton position of each 64-query chunk, or
                                           // null for positions 0, 64, 128, ...
    const uint32_t* __restrict__ qlist;    // list launch: Morton positions of the queries, 64 per wave
    int64_t nlist;                         // its length
    uint32_t* __restrict__ retry;          // queries handed back (fewer than k within the radius bound, or long)
    int64_t retry_cap;                     // entries of retry
    unsigned long long* __restrict__ queue;  // work-queue counter (persistent launch) or null
    int32_t* __restrict__ knn;             // rows, cloud after cloud, local sorted indices
    float* __restrict__ kth;               // per sorted point: k-th distance of valid rows, else -1
    unsigned long long* __restrict__ counters;  // [0] candidates, [1] invalid rows, [2] learned, [3] fallbacks,
                                                // [4] cells looked up, [5] work queue, [6] queries handed
                                                // back (retry list length), [7] keys inserted into the lists,
                                                // [8] most candidates of a query, [9] long queries handed on
    const LevelDev* __restrict__ lev;      // per-level table descriptors (device array)
    const float4* __restrict__ qpts;       // query search: queries in Morton order (x', y', z', bits(index))
    const unsigned long long* __restrict__ qvcell;  // query search: virtual cell, clamped to the grid
    int64_t nchunks;   // chunks in the work list (after sampling)
    int stride;        // 1 = every chunk; s > 1 = every s-th chunk of the list (sampled trial runs)
    int B, log2B, stages, maxc;
    int Bv;            // cells per axis at internal level 0 (B << SUB)
    int nlev;          // internal levels: SUB sub-cell levels + the reference pyramid
    int start_level;   // internal level of the first box (SUB = the query's own fine cell)
    int level_offset;  // added to every proposed start level
    uint32_t long_cap; // lane-per-query search: a query that has tested more candidates than this is handed
                       // back to a wave-per-query pass (0 = never)
    uint32_t bigmask;  // levels holding a cell with >= maxc points (any cloud)
    float a, ba;
};

// FP64 min / max as single instructions. The keys are never NaN, so the quieting the compiler adds to keep
// IEEE minNum semantics for arbitrary inputs (one more v_max_f64 per operand) is not needed.
__device__ inline double dmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ inline double dmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Inserts key into the ascending list L (the largest entry drops out): position i becomes
// max(L[i-1], min(key, L[i])), which is L[i-1] when key < L[i-1], key when L[i-1] <= key < L[i], else L[i].
// Two FP64 operations per position, against a 64-bit compare, a hazard wait and four selects as integers.
template <int K>
__device__ inline void topk_insert(double (&L)[K], uint64_t key) {
    const double kd = key_d(key);
#pragma unroll
    for (int i = K - 1; i > 0; --i) L[i] = dmax(L[i - 1], dmin(kd, L[i]));
    L[0] = dmin(kd, L[0]);
}

// Reference certificate at stage s (self search GridKNNCuda.cu:694-741, query search :383-433) for a query
// whose cell at the stage's level is ci: true = the reference keeps searching. With fewer than k listed the
// self search compares the radius with > and the query search with >=.
template <bool QUERY>
__device__ inline bool ref_continue(const KnnParams& P, int kq, float thres_n, int s, float x, float y, float z,
                                    const int (&ci)[3], int c, float kthd) {
    const int lvl = s == 0 ? 0 : s - 1;
    const int off = s == 0 ? 0 : 1;
    const int Bl = P.B >> lvl;
    const float h = P.ba / (float)Bl;
    const float v[3] = {x, y, z};
    bool cont = false;
    if (c < kq) {
        const float r = __builtin_sqrtf(thres_n);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float lo = __builtin_fmaf(-(float)(ci[ax] - off), h, v[ax] - P.a);
            const float hi = __builtin_fmaf((float)(ci[ax] + off + 1), h, P.a - v[ax]);
            if (ci[ax] - off >= 0 && (QUERY ? r >= lo : r > lo)) cont = true;
            if (ci[ax] + off + 1 <= Bl && (QUERY ? r >= hi : r > hi)) cont = true;
        }
    } else {
        const float r = __builtin_sqrtf(kthd);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (ci[ax] - off >= 0 && r >= __builtin_fmaf(-(float)(ci[ax] - off), h, v[ax] - P.a)) cont = true;
            if (ci[ax] + off + 1 <= Bl && r >= __builtin_fmaf((float)(ci[ax] + off + 1), h, P.a - v[ax])) cont = true;
        }
    }
    return cont;
}

__device__ inline uint32_t cell_count(const KnnParams& P, uint32_t cloud, int lvl, int cx, int cy, int cz) {
    const int Bl = P.B >> lvl;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= Bl || cy >= Bl || cz >= Bl) return 0;
    return level_lookup(P.htab, P.lev[lvl], cell_key(cloud, morton_code(cx, cy, cz))).y;
}

// Range of the cell c (coordinates at internal level il < SUB) inside the range [lo, lo + cnt) of an enclosing
// cell: the points whose sub code, shifted down by 3 il, equals the Morton code of c's position inside its
// fine cell. Two binary searches; probes counts their steps.
__device__ inline uint2 sub_range(const KnnParams& P, uint32_t lo, uint32_t cnt, int il, const int (&c)[3],
                                  uint32_t& probes) {
    const uint32_t m = (1u << (SUB - il)) - 1u;
    const uint32_t want = morton_code((uint32_t)c[0] & m, (uint32_t)c[1] & m, (uint32_t)c[2] & m);
    const int sh = 3 * il;
    uint32_t a = lo, b = lo + cnt;  // first position with prefix >= want
    while (a < b) {
        const uint32_t mid = (a + b) >> 1;
        ++probes;
        if ((vcell_sub(P.mvcell[mid]) >> sh) < want) a = mid + 1;
        else b = mid;
    }
    uint32_t e = lo + cnt;  // first position with prefix > want
    b = a;
    while (b < e) {
        const uint32_t mid = (b + e) >> 1;
        ++probes;
        if ((vcell_sub(P.mvcell[mid]) >> sh) <= want) b = mid + 1;
        else e = mid;
    }
    return make_uint2(a, b - a);
}

// Range of the cell c at internal level il: the level table for il >= SUB, else the fine cell's range (level 0
// table) narrowed by sub_range.
__device__ inline uint2 cell_range(const KnnParams& P, uint32_t cloud, int il, const int (&c)[3], uint32_t& probes) {
    if (il >= SUB)
        return level_lookup(P.htab, P.lev[il - SUB], cell_key(cloud, morton_code(c[0], c[1], c[2])));
    const int sh = SUB - il;
    const uint2 r = level_lookup(P.htab, P.lev[0], cell_key(cloud, morton_code(c[0] >> sh, c[1] >> sh, c[2] >> sh)));
    if (r.y == 0) return r;
    return sub_range(P, r.x, r.y, il, c, probes);
}

// The 27 cells of a 3x3x3 block (t = dx + 1 + 3 (dy + 1) + 9 (dz + 1)) nearest first: centre, faces, edges,
// corners, so the bound tightens early and the farther cells can be pruned.
__constant__ int8_t c_shell_order[27] = {13, 4, 10, 12, 14, 16, 22, 1, 3, 5, 7, 9, 11, 15, 17,
                                         19, 21, 23, 25, 0, 2, 6, 8, 18, 20, 24, 26};

// Next chunk of the wave: from the shared work queue (persistent launch) or the wave's fixed slot.
template <bool PERSIST>
__device__ inline int64_t next_chunk(const KnnParams& P, int64_t cur) {
    if (!PERSIST) return cur < 0 ? (int64_t)blockIdx.x * (KBLK / 64) + (threadIdx.x >> 6) : P.nchunks;
    unsigned long long v = 0;
    if (__lane_id() == 0) v = atomicAdd(P.queue, 1ull);
    return (int64_t)__shfl(v, 0, 64);
}

// QUERY = false: self search of the reference points (GRIDCUDAKNNSELF_COMPACT). QUERY = true: search of a
// separate query set against the reference points (GRIDCUDAKNNCompact, GridKNNCuda.cu:1645-1939, kernel
// construct_neighbors :340-440), which differs in four places: no point is excluded, candidates need
// d < thres (strict), the reference computes a query's cell at every level from its coordinates (it can lie
// outside the grid, and (int) truncates toward zero) and its short-list certificate uses >=. Queries come
// from qpts / qvcell (Morton order of their cell clamped to the grid); rows go to knn at the query's original
// index. A query outside the grid skips the early exit below: its search boxes (around the clamped cell) are
// not the reference's stage boxes, so it searches until the list is exact.
// Occupancy: the list (2 K registers) dominates the register budget; small lists are held to 128 VGPRs
// (four waves per SIMD) and K <= 40 to 168 (three), which the measured runs favour over fewer waves without
// spills. Four waves at K = 40 spill 110 VGPRs (252 bytes of scratch per lane, against 19 VGPRs at three) and
// run 2.0-2.4x slower: a 16M-point cluster mixture from the own cell 150.8 against 62.0 ms, from learned
// starts 105.8 against 52.2; a 4M-point torus at level 2 18.1 against 8.2.
// LIST: each wave takes 64 queries of a list (the queries a bounded search handed back), each lane its own
// query and cloud.
template <int K, bool PERSIST, bool QUERY = false, bool LIST = false>
__global__ void __launch_bounds__(KBLK)
    __attribute__((amdgpu_waves_per_eu(K <= 8 || (K <= 16 && !PERSIST) ? 4 : (K <= 40 && !PERSIST ? 3 : 1))))
    k_knn(KnnParams P) {
    uint64_t cand = 0;
    uint32_t invalid_row = 0, learned = 0, fallback = 0, cells = 0, inserted = 0, most = 0, nlong = 0;
    // Each wave takes chunks of 64 Morton-consecutive queries of one cloud, either one fixed chunk or, in a
    // persistent launch, chunk after chunk from a work queue until the list is exhausted.
    for (int64_t v = next_chunk<PERSIST>(P, -1); v < P.nchunks; v = next_chunk<PERSIST>(P, v)) {
    const int64_t start = LIST ? 0 : (P.chunks ? P.chunks[v * P.stride] : v * P.stride * 64);
    const int64_t li = v * 64 + __lane_id();
    const int64_t m = LIST ? (li < P.nlist ? (int64_t)P.qlist[li] : -1) : start + __lane_id();
    const uint32_t cloud = P.mcloud ? (uint32_t)P.mcloud[LIST ? (m < 0 ? 0 : m) : start] : 0u;
    const CloudDev cd = P.clouds[cloud];
    const int kq = cd.k;
    const float thres_n = cd.thres_n;
    const bool active = LIST ? m >= 0 : m < cd.base + cd.n;
    // Learned start: every lane of a wave starts at the smallest level proposed in the wave, so the wave's
    // 64 neighbouring queries scan the same cells in step; lanes whose proposal is higher expand from there.
    int own_level = P.start_level;
    int wave_level = own_level;
    if (P.qlevel) {
        own_level = active ? (int)P.qlevel[m] + SUB + P.level_offset : 127;
        wave_level = own_level;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) wave_level = min(wave_level, __shfl_xor(wave_level, d, 64));
        learned += active ? 1u : 0u;
    }
    if (active) {
        const uint64_t cand0 = cand;
        const float4 q = QUERY ? P.qpts[m] : P.mpts[m];
        const int32_t jq = __float_as_int(q.w);
        const uint64_t qvc = QUERY ? P.qvcell[m] : P.mvcell[m];
        const int vq[3] = {vcell_axis(qvc, 0), vcell_axis(qvc, 1), vcell_axis(qvc, 2)};
        bool outside = false;  // query search: the query's own (unclamped) cell lies outside the grid
        if (QUERY) {
            const int ux = cell_of(q.x, P.a, P.ba, P.B), uy = cell_of(q.y, P.a, P.ba, P.B),
                      uz = cell_of(q.z, P.a, P.ba, P.B);
            outside = ux < 0 || uy < 0 || uz < 0 || ux >= P.B || uy >= P.B || uz >= P.B;
        }
        // Bounded search (learned radius r): the list starts full of placeholder keys (r^2 bits, index
        // 0xFFFFFFFF), above every real key at distance <= r and below every key beyond it. Candidates beyond
        // r are then never buffered and cells beyond r are pruned from the start; the exactness stop below
        // holds as written (once every point within the bound is scanned, the list is exact). A list that
        // still holds a placeholder at the end has fewer than k points within r: the query is handed back
        // and searched again without the bound.
        const uint64_t init_key =
            P.qbound ? (((uint64_t)__float_as_uint(fminf(P.qbound[m], thres_n)) << 32) | 0xFFFFFFFFull) : KEY_EMPTY;
        double L[K];
#pragma unroll
        for (int i = 0; i < K; ++i) L[i] = key_d((i < K - kq) ? 0ull : init_key);
        // Keys that beat the list's bound wait in a small per-lane buffer and are inserted in batches: the
        // wave executes an insertion step only as often as its busiest lane needs one, instead of at every
        // candidate where any lane inserts. The list (and so the bound used for pruning) lags by at most
        // the buffered keys, which only makes pruning less tight, never wrong.
        constexpr int KNB = knn_buffer<K>();
        uint64_t pbuf[KNB];
        int npend = 0;
#pragma unroll
        for (int j = 0; j < KNB; ++j) pbuf[j] = KEY_EMPTY;
        auto flush = [&]() {
#pragma unroll
            for (int j = 0; j < KNB; ++j) {
                if (j < npend) topk_insert<K>(L, pbuf[j]);
                pbuf[j] = KEY_EMPTY;
            }
            inserted += (uint32_t)npend;
            npend = 0;
        };

        int level = wave_level;
        level = level < 0 ? 0 : (level >= P.nlev ? P.nlev - 1 : level);
        const int first_level = level;
        uint32_t fallback_q = 0;
        bool gave_up = false;  // long query, handed on to the wave-per-query pass
        int plo[3] = {P.Bv, P.Bv, P.Bv}, phi[3] = {-1, -1, -1};  // previous box (virtual coords), empty
        for (;;) {
            const int Bl = P.Bv >> level;
            const int cq[3] = {vq[0] >> level, vq[1] >> level, vq[2] >> level};
            // Early exit (exact): the box of reference level l = level - SUB is what the reference's stage
            // l + 1 scans. Once the box of level l - 1 is scanned, the list answers every earlier stage
            // exactly as the reference's own list would. If stage l + 1 does not exist (l >= stages) or would
            // meet a cell of >= maxc points (the reference then invalidates the query), nothing beyond the
            // scanned box can change the outcome: the replay below either stops at an earlier stage with this
            // list or blanks the row.
            if (level > first_level && level > SUB && !outside) {
                const int rl = level - SUB;
                if (rl >= P.stages) break;
                if ((P.bigmask >> rl) & 1u) {
                    bool big = false;
                    for (int t = 0; t < 27 && !big; ++t) {
                        if (t == 13) continue;
                        big = cell_count(P, cloud, rl, cq[0] + (t % 3) - 1, cq[1] + ((t / 3) % 3) - 1,
                                         cq[2] + (t / 9) - 1) >= (uint32_t)P.maxc;
                    }
                    if (big) break;
                }
            }
            const float qv[3] = {q.x, q.y, q.z};
            // Offers the points of Morton range [start, pend) (skipping those inside the previous box when the
            // range's cell overlaps it; that only happens above the fine level, where the previous box is made of
            // whole fine cells, so the point's fine cell decides). Candidates are fetched CH at a time before any
            // is used (one memory latency per chunk); the top-K under the (distance, index) order does not depend
            // on the offering order.
            auto scan = [&](uint32_t start, uint32_t pend, bool overlap) {
                constexpr int CH = knn_fetch<K>();
                for (uint32_t p0 = start; p0 < pend; p0 += CH) {
                    float4 ev[CH];
                    uint32_t cc[CH];
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        const uint32_t p = min(p0 + u, pend - 1);
                        ev[u] = P.mpts[p];
                        cc[u] = overlap ? P.mcell[p] : 0u;
                    }
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        if (p0 + u >= pend) continue;
                        const float4 e = ev[u];
                        const int32_t j = __float_as_int(e.w);
                        if (!QUERY && j == jq) continue;
                        if (overlap) {  // fine cell scaled to virtual coordinates (box bounds are multiples of 2^SUB)
                            const int ex = (int)(cc[u] & 1023u) << SUB, ey = (int)((cc[u] >> 10) & 1023u) << SUB,
                                      ez = (int)(cc[u] >> 20) << SUB;
                            if (ex >= plo[0] && ex <= phi[0] && ey >= plo[1] && ey <= phi[1] && ez >= plo[2] &&
                                ez <= phi[2])
                                continue;
                        }
                        ++cand;
                        const float d = sqd(q.x, q.y, q.z, e.x, e.y, e.z);
                        if (QUERY ? d < thres_n : d <= thres_n) {
                            const uint64_t kk = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)j;
                            if (kk < key_u(L[K - 1])) {
#pragma unroll
                                for (int b = 0; b < KNB; ++b) pbuf[b] = b == npend ? kk : pbuf[b];
                                ++npend;
                            }
                        }
                    }
                    if (__ballot(npend > KNB - CH)) flush();  // room for the next CH candidates
                }
            };
            // Cell test at level lv (cell coords c): skip when fully inside the previous box or when its box
            // is farther from the query than the current bound (the k-th key distance once the list is full,
            // else the search radius). The box distance is shrunk by 1e-6 (normalized units), far more than
            // the rounding of a point's float distance or of its cell assignment, so no point of a skipped
            // cell could have entered the list. Returns 0 skip, 1 scan, 2 scan with the previous-box filter.
            auto cell_test = [&](int lv, const int c[3]) -> int {
                bool inside = true, overlap = true;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const int lo = c[ax] << lv, hi = ((c[ax] + 1) << lv) - 1;
                    inside = inside && lo >= plo[ax] && hi <= phi[ax];
                    overlap = overlap && lo <= phi[ax] && hi >= plo[ax];
                }
                if (inside) return 0;
                const float hc = P.ba / (float)(P.Bv >> lv);
                const uint64_t lk = key_u(L[K - 1]);
                const float bnd = lk != KEY_EMPTY ? __uint_as_float((uint32_t)(lk >> 32)) : thres_n;
                float dd = 0.0f;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const float lo = P.a + (float)c[ax] * hc;
                    const float gap = fmaxf(fmaxf(lo - qv[ax], qv[ax] - (lo + hc)) - 1e-6f, 0.0f);
                    dd = __builtin_fmaf(gap, gap, dd);
                }
                if (dd > bnd) return 0;
                return overlap ? 2 : 1;
            };
            for (int oi = 0; oi < 27; ++oi) {
                const int t = c_shell_order[oi];
                const int c3[3] = {cq[0] + (t % 3) - 1, cq[1] + ((t / 3) % 3) - 1, cq[2] + (t / 9) - 1};
                if (c3[0] < 0 || c3[1] < 0 || c3[2] < 0 || c3[0] >= Bl || c3[1] >= Bl || c3[2] >= Bl) continue;
                const int ct = cell_test(level, c3);
                if (!ct) continue;
                ++cells;
                const uint2 r = cell_range(P, cloud, level, c3, cells);
                // A populous cell is tested and scanned child by child, nearest child first (the child on the
                // query's side along each axis, then across one, two and three mid-planes), so the list fills
                // with near points early and the farther children can be pruned. At or below the fine level a
                // cell that overlaps the previous box (one level finer) is split too: each child then lies wholly
                // inside the previous box (skipped) or wholly outside it, so no point needs the box filter.
                const bool split = level > 0 && (r.y > SPLIT_COUNT || (ct == 2 && level <= SUB));
                int near = 0;
                if (split) {
                    const float hc = P.ba / (float)(P.Bv >> (level - 1));
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax)
                        near |= (qv[ax] >= P.a + (float)(2 * c3[ax] + 1) * hc ? 1 : 0) << ax;
                }
                for (int oc = 0; oc < (split ? 8 : 1); ++oc) {
                    uint32_t start = r.x, count = r.y;
                    bool ov = ct == 2;
                    if (split) {
                        const int ch = near ^ ((0x76534210u >> (4 * oc)) & 7);  // xor masks 0,1,2,4,3,5,6,7
                        const int cc3[3] = {2 * c3[0] + (ch & 1), 2 * c3[1] + ((ch >> 1) & 1), 2 * c3[2] + (ch >> 2)};
                        const int cct = cell_test(level - 1, cc3);
                        if (!cct) continue;
                        ++cells;
                        const uint2 rc =
                            level - 1 >= SUB
                                ? level_lookup(P.htab, P.lev[level - 1 - SUB],
                                               cell_key(cloud, morton_code(cc3[0], cc3[1], cc3[2])))
                                : sub_range(P, r.x, r.y, level - 1, cc3, cells);
                        start = rc.x;
                        count = rc.y;
                        ov = cct == 2;
                    }
                    scan(start, start + count, ov);
                }
                if (!QUERY && !LIST && P.long_cap && cand - cand0 > P.long_cap) {
                    gave_up = true;
                    break;
                }
            }
            if (gave_up) break;
            flush();
            if (level >= P.nlev - 1) break;
            // exactness: every unscanned point is farther than the current bound
            const uint64_t lk = key_u(L[K - 1]);
            const bool full = lk != KEY_EMPTY;
            const float bound = full ? __uint_as_float((uint32_t)(lk >> 32)) : thres_n;
            const double R = __builtin_sqrt((double)bound) + 2e-6;
            const double hl = (double)P.ba / (double)Bl;
            const double qd[3] = {q.x, q.y, q.z};
            bool exact = true;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                if (cq[ax] - 1 >= 1 && !(R < qd[ax] - ((double)P.a + (double)(cq[ax] - 1) * hl))) exact = false;
                if (cq[ax] + 1 <= Bl - 2 && !(R < ((double)P.a + (double)(cq[ax] + 2) * hl) - qd[ax])) exact = false;
            }
            if (exact) break;
            if (P.qlevel && level >= own_level) fallback_q = 1;  // the proposed box did not certify the list
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                plo[ax] = max(0, (cq[ax] - 1) << level);
                phi[ax] = min(P.Bv - 1, ((cq[ax] + 2) << level) - 1);
            }
            ++level;
        }

        most = max(most, (uint32_t)min<uint64_t>(cand - cand0, 0xFFFFFFFFull));
        const bool handed_back = gave_up || (P.qbound && (uint32_t)key_u(L[K - 1]) == 0xFFFFFFFFu);
        if (handed_back) {
            const unsigned long long at = atomicAdd(&P.counters[6], 1ull);
            if (at < (unsigned long long)P.retry_cap) P.retry[at] = (uint32_t)m;
        }
        nlong += gave_up ? 1u : 0u;
        if (!handed_back) {
        // replay of the reference schedule on the exact list
        int c = 0;
        int lo[3] = {P.B, P.B, P.B}, hi[3] = {-1, -1, -1};
        const uint32_t bm = (uint32_t)P.B - 1u;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (i >= K - kq && key_u(L[i]) != KEY_EMPTY) {
                ++c;
                const uint32_t kk = (uint32_t)P.key[(uint32_t)key_u(L[i])];
                const int mx = kk & bm, my = (kk >> P.log2B) & bm, mz = kk >> (2 * P.log2B);
                lo[0] = min(lo[0], mx); hi[0] = max(hi[0], mx);
                lo[1] = min(lo[1], my); hi[1] = max(hi[1], my);
                lo[2] = min(lo[2], mz); hi[2] = max(hi[2], mz);
            }
        }
        const float kthd = c > 0 ? __uint_as_float((uint32_t)(key_u(L[K - 1]) >> 32)) : 0.0f;
        bool valid = false;
        for (int s = 0; s <= P.stages; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            const int off = s == 0 ? 0 : 1;
            int cx = vq[0] >> (SUB + lvl), cy = vq[1] >> (SUB + lvl), cz = vq[2] >> (SUB + lvl);
            if (QUERY) {  // the reference's cell of the query at this level (GridKNNCuda.cu:351-353)
                cx = cell_of(q.x, P.a, P.ba, P.B >> lvl);
                cy = cell_of(q.y, P.a, P.ba, P.B >> lvl);
                cz = cell_of(q.z, P.a, P.ba, P.B >> lvl);
            }
            if ((P.bigmask >> lvl) & 1u) {
                bool big = false;
                if (s == 0) {
                    big = cell_count(P, cloud, 0, cx, cy, cz) >= (uint32_t)P.maxc;
                } else {
                    for (int t = 0; t < 27 && !big; ++t) {
                        if (t == 13) continue;
                        big = cell_count(P, cloud, lvl, cx + (t % 3) - 1, cy + ((t / 3) % 3) - 1,
                                         cz + (t / 9) - 1) >= (uint32_t)P.maxc;
                    }
                }
                if (big) break;
            }
            const bool inside = c == 0 || ((lo[0] >> lvl) >= cx - off && (hi[0] >> lvl) <= cx + off &&
                                           (lo[1] >> lvl) >= cy - off && (hi[1] >> lvl) <= cy + off &&
                                           (lo[2] >> lvl) >= cz - off && (hi[2] >> lvl) <= cz + off);
            const int cst[3] = {cx, cy, cz};
            if (inside && !ref_continue<QUERY>(P, kq, thres_n, s, q.x, q.y, q.z, cst, c, kthd)) {
                valid = true;
                break;
            }
        }
        int32_t* row = P.knn + cd.row_off + (int64_t)(jq - cd.base) * kq;
        const int32_t base = QUERY ? 0 : (int32_t)cd.base;  // query rows hold reference sorted indices
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (i >= K - kq) {
                const int32_t val = (valid && key_u(L[i]) != KEY_EMPTY) ? (int32_t)(uint32_t)key_u(L[i]) - base : -1;
                row[i - (K - kq)] = val;
            }
        }
        if (!QUERY) P.kth[jq] = (valid && c > 0) ? kthd : -1.0f;
        invalid_row += valid ? 0u : 1u;
        fallback += fallback_q;
        }
    }
    }
    // one atomic per wave
    unsigned long long c0 = cand;
    uint32_t c1 = invalid_row, c2 = learned, c3 = fallback, c4 = cells, c7 = inserted, c8 = most, c9 = nlong;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        c0 += __shfl_xor(c0, d, 64);
        c9 += __shfl_xor(c9, d, 64);
        c4 += __shfl_xor(c4, d, 64);
        c7 += __shfl_xor(c7, d, 64);
        c8 = max(c8, (uint32_t)__shfl_xor(c8, d, 64));
        c1 += __shfl_xor(c1, d, 64);
        c2 += __shfl_xor(c2, d, 64);
        c3 += __shfl_xor(c3, d, 64);
    }
    if (__lane_id() == 0) {
        if (c0) atomicAdd(&P.counters[0], c0);
        if (c1) atomicAdd(&P.counters[1], (unsigned long long)c1);
        if (c2) atomicAdd(&P.counters[2], (unsigned long long)c2);
        if (c3) atomicAdd(&P.counters[3], (unsigned long long)c3);
        if (c4) atomicAdd(&P.counters[4], (unsigned long long)c4);
        if (c7) atomicAdd(&P.counters[7], (unsigned long long)c7);
        if (c8) atomicMax(&P.counters[8], (unsigned long long)c8);
        if (c9) atomicAdd(&P.counters[9], (unsigned long long)c9);
    }
}

// ---------------------------------------------------------------------------------------------------------
// Wave-per-query search (launch shape EGT_OPT_KNN_LANES = 64): the same exact search and replay as k_knn, with
// the 64 lanes of a wave working on one query at a time. Per box: lane t looks up the t-th of the 27 cells
// (nearest first) in parallel, populous cells are split into 8 children looked up in parallel too, and the
// ranges of the cells that pass the bound are packed end to end into batches of 64 candidates (one per lane,
// coalesced). The list is spread over the lanes, lane i holding the i-th smallest key (lanes past k hold the
// next keys, which never reach a row). A batch's survivors go in one by one as max(L[i-1], min(key, L[i]))
// with L[i-1] shifted in from the lane below, or, when many survive, by a bitonic sort of the batch and a
// bitonic merge with the list. Each wave serves 16 Morton-consecutive queries of a 64-query chunk and keeps
// the ranges of the last box it looked up for the next query with the same box.
namespace wv {

template <int CTRL, int ROWM, bool BC>
__device__ inline int dpp(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWM, 0xf, BC);
}
// Inclusive prefix sum over the wave: row_shr 1, 2, 4, 8 inside each row of 16 lanes (lanes shifted in from
// outside the row read 0), then row 0's total into row 1 and row 2's into row 3 (row_bcast:15), then rows
// 0-1's total into rows 2-3 (row_bcast:31).
__device__ inline int scan_add(int x) {
    x += dpp<0x111, 0xf, true>(x);
    x += dpp<0x112, 0xf, true>(x);
    x += dpp<0x114, 0xf, true>(x);
    x += dpp<0x118, 0xf, true>(x);
    x += dpp<0x142, 0xa, false>(x);
    x += dpp<0x143, 0xc, false>(x);
    return x;
}
__device__ inline int readlane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ inline double readlane_d(double x, int l) {
    const uint64_t u = key_u(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return key_d(((uint64_t)hi << 32) | lo);
}
// Lane i receives lane i - 1's value, lane 0 receives +0.0 (below every key). wave_shr:1.
__device__ inline double shift_up(double x) {
    const uint64_t u = key_u(x);
    const uint32_t lo = (uint32_t)dpp<0x138, 0xf, true>((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)dpp<0x138, 0xf, true>((int)(uint32_t)(u >> 32));
    return key_d(((uint64_t)hi << 32) | lo);
}
__device__ inline double shfl_xor_d(double x, int j) {
    const uint64_t u = key_u(x);
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)u, j, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(u >> 32), j, 64);
    return key_d(((uint64_t)hi << 32) | lo);
}
// Value of lane ^ J: quad_perm for 1 and 2, row_shl / row_shr inside rows of 16 for 4 and 8, else ds_bpermute.
template <int J>
__device__ inline int xor_lane(int x) {
    if constexpr (J == 1) {
        return dpp<0xB1, 0xf, true>(x);
    } else if constexpr (J == 2) {
        return dpp<0x4E, 0xf, true>(x);
    } else if constexpr (J == 4 || J == 8) {
        const int a = dpp<0x100 + J, 0xf, true>(x), b = dpp<0x110 + J, 0xf, true>(x);
        return (__lane_id() & J) ? b : a;
    } else {
        return __shfl_xor(x, J, 64);
    }
}
template <int J>
__device__ inline double xor_lane_d(double x) {
    const uint64_t u = key_u(x);
    const uint32_t lo = (uint32_t)xor_lane<J>((int)(uint32_t)u), hi = (uint32_t)xor_lane<J>((int)(uint32_t)(u >> 32));
    return key_d(((uint64_t)hi << 32) | lo);
}
// One compare-exchange stage of a bitonic network: partner lane ^ J; the lane keeps the smaller key when its
// J bit equals its K2 bit (ascending blocks of K2; K2 = 64 and above sorts the whole wave ascending).
template <int K2, int J>
__device__ inline double bstage(double x) {
    const int lane = __lane_id();
    const double p = xor_lane_d<J>(x);
    const bool up = K2 >= 64 || (lane & K2) == 0;
    const bool low = (lane & J) == 0;
    const double mn = dmin(x, p), mx = dmax(x, p);  // both, then a select: no exec-mask branch
    return (low == up) ? mn : mx;
}
// Bitonic sort of the 64 lanes' keys, ascending by lane.
__device__ inline double sort64(double x) {
    x = bstage<2, 1>(x);
    x = bstage<4, 2>(x); x = bstage<4, 1>(x);
    x = bstage<8, 4>(x); x = bstage<8, 2>(x); x = bstage<8, 1>(x);
    x = bstage<16, 8>(x); x = bstage<16, 4>(x); x = bstage<16, 2>(x); x = bstage<16, 1>(x);
    x = bstage<32, 16>(x); x = bstage<32, 8>(x); x = bstage<32, 4>(x); x = bstage<32, 2>(x); x = bstage<32, 1>(x);
    x = bstage<64, 32>(x); x = bstage<64, 16>(x); x = bstage<64, 8>(x); x = bstage<64, 4>(x); x = bstage<64, 2>(x);
    x = bstage<64, 1>(x);
    return x;
}
// The 64 smallest of the ascending list L and the ascending batch S, ascending: min(L[i], S[63 - i]) is a
// bitonic sequence holding them, then a bitonic merge.
__device__ inline double merge64(double L, double S) {
    double x = dmin(L, shfl_xor_d(S, 63));
    x = bstage<64, 32>(x); x = bstage<64, 16>(x); x = bstage<64, 8>(x); x = bstage<64, 4>(x); x = bstage<64, 2>(x);
    x = bstage<64, 1>(x);
    return x;
}
// Running maximum over the wave of non-negative values (the same DPP steps as scan_add).
__device__ inline int scan_max(int x) {
    x = max(x, dpp<0x111, 0xf, true>(x));
    x = max(x, dpp<0x112, 0xf, true>(x));
    x = max(x, dpp<0x114, 0xf, true>(x));
    x = max(x, dpp<0x118, 0xf, true>(x));
    x = max(x, dpp<0x142, 0xa, false>(x));
    x = max(x, dpp<0x143, 0xc, false>(x));
    return x;
}
// Orders a wave's LDS accesses before and after it.
__device__ inline void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Number of lanes whose (non-decreasing) value v is <= g, for g below the last lane's value: branch-free
// binary search over the lanes.
__device__ inline int count_le(int v, int g) {
    int pos = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (__shfl(v, pos + step - 1, 64) <= g) pos += step;
    return pos;
}

}  // namespace wv

// Cell test of k_knn (skip inside the previous box or beyond the bound), also returning the box distance.
__device__ inline int wave_cell_test(const KnnParams& P, int lv, const int (&c)[3], const int (&plo)[3],
                                     const int (&phi)[3], const float (&qv)[3], float bnd, float& dd) {
    bool inside = true, overlap = true;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int lo = c[ax] << lv, hi = ((c[ax] + 1) << lv) - 1;
        inside = inside && lo >= plo[ax] && hi <= phi[ax];
        overlap = overlap && lo <= phi[ax] && hi >= plo[ax];
    }
    dd = 0.0f;
    if (inside) return 0;
    const float hc = P.ba / (float)(P.Bv >> lv);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float lo = P.a + (float)c[ax] * hc;
        const float gap = fmaxf(fmaxf(lo - qv[ax], qv[ax] - (lo + hc)) - 1e-6f, 0.0f);
        dd = __builtin_fmaf(gap, gap, dd);
    }
    if (dd > bnd) return 0;
    return overlap ? 2 : 1;
}

// LIST: each wave takes one query of a list (long queries a lane-per-query search handed on), from the
// handle's start level.
// Lanes are mapped to a batch's entries at a constant cost: each entry with points left marks, in LDS, the slot
// its run starts at, and a running maximum over the lanes carries it to the slots after it. Against mapping
// the entries one at a time (four readlanes each), at 6 waves per SIMD on a 16M-point cluster mixture: k=64
// from learned starts 59.1 against 62.8 ms, from the own cell 81.7 against 89.4; k=40 from the own cell 69.5
// against 75.8, from learned starts 60.5 against 58.0 (there a lane per query is faster still); a 4M-point
// torus at level 2 9.7 against 10.2 (k=40) and 11.0 against 12.0 ms (k=64). A hybrid (one at a time up to 4
// entries) fell in between on every case.
// Occupancy: the kernel is VALU-bound, and its wave-uniform values outgrow the scalar registers. Held to 8 waves
// per SIMD it spilled 166 of them into vector lanes (778 v_readlane in the code); at 6 (75 VGPRs) 98 spill
// (387 v_readlane), and the search runs faster with the same rows: the mixture at k=64 from learned starts
// 68.7 -> 62.7 ms, from the own cell 97.2 -> 89.6; k=40 from the own cell 82.1 -> 75.8 (learned starts 55.9 ->
// 56.3); a 4M-point torus at level 2 11.2 -> 10.2 (k=40) and 13.4 -> 12.0 ms (k=64). 7 waves: in between.
template <bool LIST>
__global__ void __launch_bounds__(KBLK) __attribute__((amdgpu_waves_per_eu(6))) k_knn_wave(KnnParams P) {
    const int lane = __lane_id();
    const int64_t v = blockIdx.x;
    if (v >= P.nchunks) return;
    // wave index made uniform for the compiler, so the query's values live in scalar registers
    const int wv_id = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    __shared__ int wmark[KBLK / 64][64];  // per wave: the entry whose run starts at each slot of a batch
    int64_t m0, m1;
    if (LIST) {
        m0 = v * (KBLK / 64) + wv_id;
        m1 = min<int64_t>(m0 + 1, P.nlist);
    } else {
        const int64_t start = P.chunks ? P.chunks[v * P.stride] : v * P.stride * 64;
        const CloudDev& c0 = P.clouds[P.mcloud ? (uint32_t)P.mcloud[start] : 0u];
        m0 = start + (int64_t)wv_id * 16;
        m1 = min<int64_t>(min<int64_t>(m0 + 16, start + 64), c0.base + c0.n);
    }
    uint64_t ncand = 0;
    uint32_t ncells = 0, nins = 0, ninvalid = 0, nlearned = 0, nfallback = 0, most = 0;
    // ranges of the 27 cells of the last box looked up (lane t: the t-th cell nearest first)
    int c_level = -1, c_cq0 = 0, c_cq1 = 0, c_cq2 = 0;
    uint32_t c_cloud = 0;
    uint32_t rs = 0, rn = 0;
    const int tt = lane < 27 ? (int)c_shell_order[lane] : 13;
    const int dxo = (tt % 3) - 1, dyo = ((tt / 3) % 3) - 1, dzo = (tt / 9) - 1;
    for (int64_t mi = m0; mi < m1; ++mi) {
        const int64_t m = LIST ? (int64_t)P.qlist[mi] : mi;
        const uint32_t cloud = P.mcloud ? (uint32_t)P.mcloud[m] : 0u;
        const CloudDev cd = P.clouds[cloud];
        const int kq = cd.k;
        const float thres_n = cd.thres_n;
        const int64_t cend = cd.base + cd.n;
        const uint64_t cand0 = ncand;
        const float4 q = P.mpts[m];
        const int32_t jq = __float_as_int(q.w);
        const uint64_t qvc = P.mvcell[m];
        const int vq[3] = {vcell_axis(qvc, 0), vcell_axis(qvc, 1), vcell_axis(qvc, 2)};
        const float qv[3] = {q.x, q.y, q.z};
        int own_level = P.start_level;
        if (P.qlevel) {
            own_level = (int)P.qlevel[m] + SUB + P.level_offset;
            ++nlearned;
        }
        int level = own_level < 0 ? 0 : (own_level >= P.nlev ? P.nlev - 1 : own_level);
        const int first_level = level;
        uint32_t fallback_q = 0;
        int plo[3] = {P.Bv, P.Bv, P.Bv}, phi[3] = {-1, -1, -1};
        // Radius bound (as k_knn): every lane starts with the placeholder key at the proposed radius, so only
        // candidates inside it survive and cells beyond it are pruned; a list whose k-th entry is still the
        // placeholder holds fewer than k points within the radius and is handed back.
        double L = key_d(P.qbound ? (((uint64_t)__float_as_uint(fminf(P.qbound[m], thres_n)) << 32) | 0xFFFFFFFFull)
                                  : KEY_EMPTY);
        for (;;) {
            const int Bl = P.Bv >> level;
            const int cq[3] = {vq[0] >> level, vq[1] >> level, vq[2] >> level};
            const int c3[3] = {cq[0] + dxo, cq[1] + dyo, cq[2] + dzo};
            const bool in_grid = lane < 27 && c3[0] >= 0 && c3[1] >= 0 && c3[2] >= 0 && c3[0] < Bl &&
                                 c3[1] < Bl && c3[2] < Bl;
            // early exit of k_knn: the reference's next stage does not exist or meets a cell of >= maxc points
            if (level > first_level && level > SUB) {
                const int rl = level - SUB;
                if (rl >= P.stages) break;
                if ((P.bigmask >> rl) & 1u) {
                    const bool big = in_grid && tt != 13 && cell_count(P, cloud, rl, c3[0], c3[1], c3[2]) >= (uint32_t)P.maxc;
                    if (__ballot(big)) break;
                }
            }
            if (level != c_level || cq[0] != c_cq0 || cq[1] != c_cq1 || cq[2] != c_cq2 || cloud != c_cloud) {
                c_cloud = cloud;
                c_level = level;
                c_cq0 = cq[0];
                c_cq1 = cq[1];
                c_cq2 = cq[2];
                rs = 0;
                rn = 0;
                if (in_grid) {
                    ++ncells;
                    const uint2 r = cell_range(P, cloud, level, c3, ncells);
                    rs = r.x;
                    rn = r.y;
                }
            }
            const uint64_t kth0 = key_u(wv::readlane_d(L, kq - 1));
            const float bnd0 = kth0 != KEY_EMPTY ? __uint_as_float((uint32_t)(kth0 >> 32)) : thres_n;
            float dd = 0.0f;
            const int ct = (in_grid && rn > 0) ? wave_cell_test(P, level, c3, plo, phi, qv, bnd0, dd) : 0;
            const bool split = ct != 0 && level > 0 && (rn > WAVE_SPLIT_COUNT || (ct == 2 && level <= SUB));
            const int ent = ct ? (split ? 8 : 1) : 0;
            const int eincl = wv::scan_add(ent);
            const int total = wv::readlane(eincl, 63);
            for (int w0 = 0; w0 < total; w0 += 64) {
                // entry g of the box: a cell, or one of the 8 children of a split cell (nearest child first)
                const int g = w0 + lane;
                const bool act = g < total;
                const int t = min(wv::count_le(eincl, g), 26);
                const int excl_t = __shfl(eincl - ent, t, 64);
                const uint32_t ps = (uint32_t)__shfl((int)rs, t, 64), pn = (uint32_t)__shfl((int)rn, t, 64);
                const int pinfo = __shfl(ct | (split ? 4 : 0), t, 64);
                const float pdd = __shfl(dd, t, 64);
                const int ptt = (int)c_shell_order[t];
                const int pc[3] = {cq[0] + (ptt % 3) - 1, cq[1] + ((ptt / 3) % 3) - 1, cq[2] + (ptt / 9) - 1};
                uint32_t es = ps, en = pn;
                int ect = pinfo & 3;
                float edd = pdd;
                if (act && (pinfo & 4)) {
                    const int ch = g - excl_t;
                    const float hcc = P.ba / (float)(P.Bv >> (level - 1));
                    int near = 0;
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) near |= (qv[ax] >= P.a + (float)(2 * pc[ax] + 1) * hcc ? 1 : 0) << ax;
                    const int chn = near ^ ((0x76534210u >> (4 * ch)) & 7);
                    const int cc3[3] = {2 * pc[0] + (chn & 1), 2 * pc[1] + ((chn >> 1) & 1), 2 * pc[2] + (chn >> 2)};
                    ect = wave_cell_test(P, level - 1, cc3, plo, phi, qv, bnd0, edd);
                    en = 0;
                    if (ect) {
                        ++ncells;
                        const uint2 rc = level - 1 >= SUB
                                             ? level_lookup(P.htab, P.lev[level - 1 - SUB],
                                                            cell_key(cloud, morton_code(cc3[0], cc3[1], cc3[2])))
                                             : sub_range(P, ps, pn, level - 1, cc3, ncells);
                        es = rc.x;
                        en = rc.y;
                    }
                }
                if (!act || !ect) en = 0;
                const bool eov = ect == 2;
                // the entries' ranges, end to end, 64 candidates per batch; an entry beyond the bound is dropped
                uint32_t rem = en, pos = es;
                for (;;) {
                    const uint64_t kthb = key_u(wv::readlane_d(L, kq - 1));
                    const float bnd = kthb != KEY_EMPTY ? __uint_as_float((uint32_t)(kthb >> 32)) : thres_n;
                    if (edd > bnd) rem = 0;
                    const int rincl = wv::scan_add((int)rem);
                    const int tot = wv::readlane(rincl, 63);
                    if (tot == 0) break;
                    const int rexcl = rincl - (int)rem;
                    // lane l takes slot l of the ranges laid end to end: the entries of this batch, one at a time
                    // lane l takes slot l of the entries' ranges laid end to end
                    int64_t p = -1;
                    int eo = 0;
                    {
                        int* mk = wmark[wv_id];
                        mk[lane] = 0;
                        wv::lds_sync();
                        if (rem > 0 && rexcl < 64) mk[rexcl] = lane + 1;
                        wv::lds_sync();
                        const int e = max(wv::scan_max(mk[lane]) - 1, 0);
                        wv::lds_sync();
                        const int eb = __shfl((int)pos - rexcl, e, 64);
                        eo = __shfl((int)eov, e, 64);
                        p = lane < tot ? (int64_t)eb + lane : -1;
                    }
                    const int take = min((int)rem, max(0, 64 - rexcl));
                    if (!__ballot(take > 0)) break;  // every batch consumes; guards the loop's end
                    pos += (uint32_t)take;
                    rem -= (uint32_t)take;
                    uint64_t kk = KEY_EMPTY;
                    bool tested = false;
                    if (p >= cd.base && p < cend) {
                        const float4 e4 = P.mpts[p];
                        const int32_t j = __float_as_int(e4.w);
                        bool ok = j != jq;
                        if (ok && eo) {
                            const uint32_t cc = P.mcell[p];
                            const int ex0 = (int)(cc & 1023u) << SUB, ey0 = (int)((cc >> 10) & 1023u) << SUB,
                                      ez0 = (int)(cc >> 20) << SUB;
                            if (ex0 >= plo[0] && ex0 <= phi[0] && ey0 >= plo[1] && ey0 <= phi[1] && ez0 >= plo[2] &&
                                ez0 <= phi[2])
                                ok = false;
                        }
                        if (ok) {
                            tested = true;
                            const float d = sqd(q.x, q.y, q.z, e4.x, e4.y, e4.z);
                            if (d <= thres_n) kk = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)j;
                        }
                    }
                    ncand += (uint64_t)__builtin_popcountll(__ballot(tested));
                    const bool sv = kk < kthb;
                    uint64_t smask = __ballot(sv);
                    const int ns = __builtin_popcountll(smask);
                    nins += (uint32_t)ns;
                    if (ns > WAVE_SORT_MIN) {
                        const double S = wv::sort64(sv ? key_d(kk) : key_d(KEY_EMPTY));
                        L = key_u(wv::readlane_d(L, 0)) == KEY_EMPTY ? S : wv::merge64(L, S);  // empty list: S
                    } else {
                        const double kd = key_d(kk);
                        while (smask) {
                            const int b = __builtin_ctzll(smask);
                            smask &= smask - 1;
                            const double c = wv::readlane_d(kd, b);
                            L = dmax(wv::shift_up(L), dmin(c, L));
                        }
                    }
                }
            }
            if (level >= P.nlev - 1) break;
            const uint64_t lk = key_u(wv::readlane_d(L, kq - 1));
            const bool full = lk != KEY_EMPTY;
            const float bound = full ? __uint_as_float((uint32_t)(lk >> 32)) : thres_n;
            const double R = __builtin_sqrt((double)bound) + 2e-6;
            const double hl = (double)P.ba / (double)Bl;
            const double qd[3] = {q.x, q.y, q.z};
            bool exact = true;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                if (cq[ax] - 1 >= 1 && !(R < qd[ax] - ((double)P.a + (double)(cq[ax] - 1) * hl))) exact = false;
                if (cq[ax] + 1 <= Bl - 2 && !(R < ((double)P.a + (double)(cq[ax] + 2) * hl) - qd[ax])) exact = false;
            }
            if (exact) break;
            if (P.qlevel && level >= own_level) fallback_q = 1;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                plo[ax] = max(0, (cq[ax] - 1) << level);
                phi[ax] = min(P.Bv - 1, ((cq[ax] + 2) << level) - 1);
            }
            ++level;
        }
        most = max(most, (uint32_t)min<uint64_t>(ncand - cand0, 0xFFFFFFFFull));
        if (P.qbound && (uint32_t)key_u(wv::readlane_d(L, kq - 1)) == 0xFFFFFFFFu) {
            if (lane == 0) {
                const unsigned long long at = atomicAdd(&P.counters[6], 1ull);
                if (at < (unsigned long long)P.retry_cap) P.retry[at] = (uint32_t)m;
            }
            continue;
        }
        // replay of the reference schedule on the exact list (lanes below k hold it)
        const uint64_t mykey = key_u(L);
        const bool listed = lane < kq && mykey != KEY_EMPTY;
        const int c = __builtin_popcountll(__ballot(listed));
        int mx = 0, my = 0, mz = 0;
        if (listed) {
            const uint32_t bm = (uint32_t)P.B - 1u;
            const uint32_t kk = (uint32_t)P.key[(uint32_t)mykey];
            mx = kk & bm;
            my = (kk >> P.log2B) & bm;
            mz = kk >> (2 * P.log2B);
        }
        const float kthd = c > 0 ? __uint_as_float((uint32_t)(key_u(wv::readlane_d(L, kq - 1)) >> 32)) : 0.0f;
        bool valid = false;
        for (int s = 0; s <= P.stages; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            const int off = s == 0 ? 0 : 1;
            const int cx = vq[0] >> (SUB + lvl), cy = vq[1] >> (SUB + lvl), cz = vq[2] >> (SUB + lvl);
            if ((P.bigmask >> lvl) & 1u) {
                bool big = false;
                if (s == 0) {
                    big = lane == 0 && cell_count(P, cloud, 0, cx, cy, cz) >= (uint32_t)P.maxc;
                } else {
                    big = lane < 27 && tt != 13 && cell_count(P, cloud, lvl, cx + dxo, cy + dyo, cz + dzo) >= (uint32_t)P.maxc;
                }
                if (__ballot(big)) break;
            }
            const bool out = listed && !((mx >> lvl) >= cx - off && (mx >> lvl) <= cx + off && (my >> lvl) >= cy - off &&
                                         (my >> lvl) <= cy + off && (mz >> lvl) >= cz - off && (mz >> lvl) <= cz + off);
            const bool inside = __ballot(out) == 0;
            const int cst[3] = {cx, cy, cz};
            if (inside && !ref_continue<false>(P, kq, thres_n, s, q.x, q.y, q.z, cst, c, kthd)) {
                valid = true;
                break;
            }
        }
        int32_t* row = P.knn + cd.row_off + (int64_t)(jq - cd.base) * kq;
        if (lane < kq) row[lane] = (valid && listed) ? (int32_t)(uint32_t)mykey - (int32_t)cd.base : -1;
        if (lane == 0) P.kth[jq] = (valid && c > 0) ? kthd : -1.0f;
        ninvalid += valid ? 0u : 1u;
        nfallback += fallback_q;
    }
    uint32_t c4 = ncells;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c4 += __shfl_xor(c4, d, 64);
    if (lane == 0) {
        if (ncand) atomicAdd(&P.counters[0], (unsigned long long)ncand);
        if (ninvalid) atomicAdd(&P.counters[1], (unsigned long long)ninvalid);
        if (nlearned) atomicAdd(&P.counters[2], (unsigned long long)nlearned);
        if (nfallback) atomicAdd(&P.counters[3], (unsigned long long)nfallback);
        if (c4) atomicAdd(&P.counters[4], (unsigned long long)c4);
        if (nins) atomicAdd(&P.counters[7], (unsigned long long)nins);
        if (most) atomicMax(&P.counters[8], (unsigned long long)most);
    }
}

// Query preparation: each query normalized with the reference cloud's extent (normalize_points
// GridKNNCuda.cu:841-849: x' = (x - min) / (max - min)), its fine cell clamped to the grid and its virtual
// cell inside that (the search boxes start there), and the Morton code of the fine cell as the sort key, so
// a wave's 64 queries are spatial neighbours.
__global__ void __launch_bounds__(KBLK) k_query_prep(const float* __restrict__ qx, const float* __restrict__ qy,
                                                      const float* __restrict__ qz, int64_t nq, float mn, float mx,
                                                      float a, float ba, int B, float4* __restrict__ qn,
                                                      uint64_t* __restrict__ cellp, uint32_t* __restrict__ code,
                                                      uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * KBLK + threadIdx.x;
    if (i >= nq) return;
    const float d = mx - mn;
    const float x = (qx[i] - mn) / d, y = (qy[i] - mn) / d, z = (qz[i] - mn) / d;
    const int cx = min(max(cell_of(x, a, ba, B), 0), B - 1);
    const int cy = min(max(cell_of(y, a, ba, B), 0), B - 1);
    const int cz = min(max(cell_of(z, a, ba, B), 0), B - 1);
    qn[i] = make_float4(x, y, z, 0.0f);
    cellp[i] = vcell_pack(vcoord(x, a, ba, B, cx), vcoord(y, a, ba, B, cy), vcoord(z, a, ba, B, cz));
    code[i] = morton_code(cx, cy, cz);
    idx[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(KBLK) k_query_gather(const float4* __restrict__ qn, const uint64_t* __restrict__ cellp,
                                                        const uint32_t* __restrict__ idx, int64_t nq,
                                                        float4* __restrict__ qpts, uint64_t* __restrict__ qcell) {
    const int64_t m = (int64_t)blockIdx.x * KBLK + threadIdx.x;
    if (m >= nq) return;
    const uint32_t i = idx[m];
    const float4 v = qn[i];
    qpts[m] = make_float4(v.x, v.y, v.z, __int_as_float((int32_t)i));
    qcell[m] = cellp[i];
}

// Query rows with original reference indices (GetKNNIndexes GridKNNCuda.cu:1941-1968 after a query search).
__global__ void __launch_bounds__(KBLK) k_query_original(const int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ perm, int64_t count,
                                                          int32_t* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * KBLK + threadIdx.x;
    if (e >= count) return;
    const int32_t r = rows[e];
    out[e] = r >= 0 ? perm[r] : -1;
}

int kclass(int k) {
    return k <= 8 ? 8 : k <= 16 ? 16 : k <= 24 ? 24 : k <= 32 ? 32 : k <= 40 ? 40 : k <= 48 ? 48 : k <= 56 ? 56 : 64;
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return cus;
}

// dst[i] = src[i] (the values of a key-only sort).
__global__ void __launch_bounds__(KBLK) k_copy_u32(const uint32_t* __restrict__ src, int64_t n,
                                                     uint32_t* __restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * KBLK + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int K>
void launch_knn(const KnnParams& P, hipStream_t s, bool query, bool list) {
    int grid = ceil_div(P.nchunks, KBLK / 64);
    if (grid < 1) return;
    if (query) {
        hipLaunchKernelGGL((k_knn<K, false, true>), dim3(grid), dim3(KBLK), 0, s, P);
    } else if (list) {
        hipLaunchKernelGGL((k_knn<K, false, false, true>), dim3(grid), dim3(KBLK), 0, s, P);
    } else if (P.queue) {
        // persistent: as many workgroups as fit on the GPU at once, each wave looping over the queue
        int per_cu = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_knn<K, true>, KBLK, 0));
        grid = (int)std::min<int64_t>(grid, (int64_t)std::max(1, per_cu) * device_cus());
        hipLaunchKernelGGL((k_knn<K, true>), dim3(grid), dim3(KBLK), 0, s, P);
    } else {
        hipLaunchKernelGGL((k_knn<K, false>), dim3(grid), dim3(KBLK), 0, s, P);
    }
    LAUNCH_CHECK();
}

void launch_any(const KnnParams& P, int K, hipStream_t s, bool query = false, bool list = false) {
    switch (K) {
        case 8: launch_knn<8>(P, s, query, list); break;
        case 16: launch_knn<16>(P, s, query, list); break;
        case 24: launch_knn<24>(P, s, query, list); break;
        case 32: launch_knn<32>(P, s, query, list); break;
        case 40: launch_knn<40>(P, s, query, list); break;
        case 48: launch_knn<48>(P, s, query, list); break;
        case 56: launch_knn<56>(P, s, query, list); break;
        default: launch_knn<64>(P, s, query, list); break;
    }
}

// Launch shape of a neighbour-count class: the handle's per-class choice (from the class tuner or
// EGT_OPT_KNN_CLASS_SHAPES), else the handle-wide options. 0 = one lane per query, 1 = the same as one
// persistent launch, 2 = one wave per query.
int class_shape(const egt_ctx* h, int K) {
    const int c = h->class_shape[K / 8 - 1];
    if (c >= 0) return c;
    return h->knn_lanes == 64 ? 2 : (h->knn_persistent ? 1 : 0);
}

// One neighbour-count class: the search in the given launch shape, then the queries it handed back (fewer
// than k points within their radius bound, or more candidates than the long-query cap) searched again in
// Morton order (the list fills in the order the waves reach their atomics, so it is sorted first), without
// the bound, from the handle's start level: one wave per query when the class ran in the wave shape or the
// cap is on, else one lane per query. Returns how many were handed back.
int64_t search_class(egt_ctx* h, KnnParams P, int K, hipStream_t s, int shape) {
    if (shape == 1) {  // persistent: the work-queue counter starts at zero
        P.queue = P.counters + 5;
        HIP_CHECK(hipMemsetAsync(P.queue, 0, sizeof(uint64_t), s));
    } else {
        P.queue = nullptr;
    }
    if (shape == 2) P.long_cap = 0;
    const bool hand_back = P.qbound || P.long_cap;
    if (hand_back) HIP_CHECK(hipMemsetAsync(P.counters + 6, 0, sizeof(uint64_t), s));  // the list's length
    if (shape == 2) {
        if (P.nchunks > 0) {
            hipLaunchKernelGGL(k_knn_wave<false>, dim3((unsigned)P.nchunks), dim3(KBLK), 0, s, P);
            LAUNCH_CHECK();
        }
    } else {
        launch_any(P, K, s);
    }
    if (!hand_back) return 0;
    uint64_t nb = 0;
    HIP_CHECK(hipMemcpyAsync(&nb, P.counters + 6, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (nb == 0) return 0;
    if ((int64_t)nb > P.retry_cap) throw std::runtime_error("kNN hand-back list longer than its buffer");
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < h->n) ++bits;
    k_copy_u32<<<ceil_div((int64_t)nb, KBLK), KBLK, 0, s>>>(P.retry, (int64_t)nb, P.retry + h->n);
    LAUNCH_CHECK();
    radix_sort_pairs(P.retry, P.retry + h->n, (int64_t)nb, bits, s, h->scratch);
    KnnParams R = P;
    R.qbound = nullptr;
    R.qlevel = nullptr;
    R.chunks = nullptr;
    R.queue = nullptr;
    R.long_cap = 0;
    R.qlist = P.retry;
    R.nlist = (int64_t)nb;
    R.stride = 1;
    if (P.long_cap || shape == 2) {
        R.nchunks = ceil_div((int64_t)nb, KBLK / 64);
        hipLaunchKernelGGL(k_knn_wave<true>, dim3((unsigned)R.nchunks), dim3(KBLK), 0, s, R);
        LAUNCH_CHECK();
    } else {
        R.nchunks = ((int64_t)nb + 63) / 64;
        launch_any(R, K, s, false, true);
    }
    return (int64_t)nb;
}

// Search parameters of the handle; with scratch, the per-point k-th distance, the counters and the bounded
// search's hand-back list come from the handle's scratch (else they are left null for the caller).
KnnParams knn_params(egt_ctx* h, bool scratch = true) {
    KnnParams P{};
    P.mpts = h->mpts.p;
    P.mcell = h->mcell.p;
    P.mvcell = reinterpret_cast<const unsigned long long*>(h->mvcell.p);
    P.htab = h->htab.p;
    P.key = h->key.p;
    P.qlevel = (h->learned && (h->learned_mode & 1) && h->qlevel.p) ? h->qlevel.p : nullptr;
    P.qbound = (h->learned && (h->learned_mode & 2) && h->qbound.p) ? h->qbound.p : nullptr;
    P.qlist = nullptr;
    P.nlist = 0;
    // a lane-per-query search hands long queries to a wave-per-query pass (search_class clears the cap for a
    // class that runs in the wave shape)
    P.long_cap = (uint32_t)h->knn_long_cap;
    P.level_offset = h->learned_level_offset;
    // the hand-back list and, after it, room for the values of its sort
    P.retry = (scratch && (P.qbound || P.long_cap)) ? h->scratch.get<uint32_t>(37, 2 * (size_t)h->n) : nullptr;
    P.retry_cap = P.retry ? h->n : 0;
    P.clouds = h->dclouds.p;
    P.mcloud = h->nclouds > 1 ? h->mcloud.p : nullptr;
    P.knn = h->knn.p;
    P.kth = scratch ? h->scratch.get<float>(15, h->n) : nullptr;
    P.counters = scratch ? reinterpret_cast<unsigned long long*>(h->scratch.get<uint64_t>(14, 10)) : nullptr;
    P.queue = nullptr;
    P.bigmask = 0;
    for (int l = 0; l < h->nlev; ++l)
        if (h->lev[l].max_count >= h->maxc) P.bigmask |= 1u << l;
    P.lev = h->dlev.p;
    P.B = h->B;
    int log2B = 0;
    while ((1 << log2B) < h->B) ++log2B;
    P.log2B = log2B;
    P.Bv = h->B << SUB;
    P.nlev = h->nlev + SUB;
    P.start_level = SUB + (h->start_level == EGT_START_LEVEL_OCCUPANCY ? auto_start_level(h, h->clouds[0].k)
                                                                         : std::min(h->start_level, h->nlev - 1));
    int stages = 0;
    for (int b = h->B; b >= 8; b >>= 1) ++stages;
    P.stages = std::max(stages, 1);
    P.maxc = h->maxc;
    P.a = -h->eps;
    P.ba = (1.0f + h->eps) - P.a;
    P.stride = 1;
    return P;
}

// Work lists per neighbour-count class: chunks of 64 Morton-consecutive queries that never straddle two
// clouds, class after class.
std::vector<WorkClass> build_work_lists(egt_ctx* h, DBuf<int64_t>& lists) {
    std::vector<WorkClass> out;
    if (h->nclouds == 1) {
        out.push_back(WorkClass{kclass(h->clouds[0].k), 0, (h->n + 63) / 64});
        return out;
    }
    std::vector<int64_t> all;
    for (int K : {8, 16, 24, 32, 40, 48, 56, 64}) {
        const int64_t first = (int64_t)all.size();
        for (const CloudDev& c : h->clouds) {
            if (kclass(c.k) != K) continue;
            for (int64_t p = 0; p < c.n; p += 64) all.push_back(c.base + p);
        }
        if ((int64_t)all.size() > first) out.push_back(WorkClass{K, first, (int64_t)all.size() - first});
    }
    lists.alloc(std::max<size_t>(all.size(), 1));
    HIP_CHECK(hipMemcpyAsync(lists.p, all.data(), all.size() * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    return out;
}

// Lays out the rows (cloud after cloud, n_c x k_c) and uploads the per-cloud parameters.
void prepare_clouds(egt_ctx* h) {
    int64_t off = 0;
    for (CloudDev& c : h->clouds) {
        c.row_off = off;
        off += c.n * c.k;
    }
    h->knn.alloc((size_t)std::max<int64_t>(off, 1));
    h->dclouds.alloc(h->clouds.size());
    HIP_CHECK(hipMemcpyAsync(h->dclouds.p, h->clouds.data(), h->clouds.size() * sizeof(CloudDev),
                             hipMemcpyHostToDevice, h->stream));
}

}  // namespace

int auto_start_level(const egt_ctx* h, int k, int cloud) {
    // smallest level whose mean non-empty cell of the cloud holds at least k/4 points (a 3x3x3 box around a
    // surface point then holds ~2k candidates)
    const double n = (double)h->clouds[cloud].n;
    for (int l = 0; l < h->nlev; ++l) {
        const double mean = n / (double)std::max<int64_t>(1, h->cloud_cells[(size_t)cloud * h->nlev + l]);
        if (mean >= 0.25 * k) return l;
    }
    return h->nlev - 1;
}

// Search of every cloud of the handle with the k and radius stored in h->clouds.
void run_knn_clouds(egt_ctx* h) {
    hipStream_t s = h->stream;
    prepare_clouds(h);
    KnnParams P = knn_params(h);
    const std::vector<WorkClass> classes = build_work_lists(h, h->work_lists);
    h->work_classes = classes;
    HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    int64_t handed_back = 0;
    for (const WorkClass& cr : classes) {
        P.chunks = h->nclouds > 1 ? h->work_lists.p + cr.first : nullptr;
        P.nchunks = cr.count;
        handed_back += search_class(h, P, cr.K, s, class_shape(h, cr.K));
    }
    HIP_CHECK(hipEventRecord(e1, s));
    uint64_t hc[10];
    HIP_CHECK(hipMemcpyAsync(hc, P.counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventElapsedTime(&h->stats.knn_ms, e0, e1));
    h->stats.learned_bound_retries = handed_back - (int64_t)hc[9];
    h->stats.knn_insertions = (int64_t)hc[7];
    h->stats.knn_max_candidates = (int64_t)hc[8];
    h->stats.knn_long_queries = (int64_t)hc[9];
    h->stats.knn_candidates = (int64_t)hc[0];
    h->stats.invalid_rows = (int64_t)hc[1];
    h->stats.learned_queries = (int64_t)hc[2];
    h->stats.learned_fallbacks = (int64_t)hc[3];
    h->stats.knn_cells = (int64_t)hc[4];
    h->have_knn = true;
}

void run_knn(egt_ctx* h, int k, float thres_n) {
    h->clouds[0].k = k;
    h->clouds[0].thres_n = thres_n;
    run_knn_clouds(h);
    h->k = k;
}

// Query-vs-reference search (GRIDCUDAKNNCompact): queries dqx/dqy/dqz (device, original order, input units)
// against the handle's reference points with neighbour count k and squared normalized radius thres_n. Rows
// land in h->qknn in original query order with reference sorted indices; the self rows are untouched.
void run_knn_query(egt_ctx* h, const float* dqx, const float* dqy, const float* dqz, int64_t nq, int k,
                   float thres_n) {
    hipStream_t s = h->stream;
    h->nq = nq;
    h->kq = k;
    h->qknn.alloc((size_t)std::max<int64_t>(nq * k, 1));
    h->qpts.alloc((size_t)std::max<int64_t>(nq, 1));
    h->qvcell.alloc((size_t)std::max<int64_t>(nq, 1));
    KnnParams P = knn_params(h);
    P.long_cap = 0;
    HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    CloudDev qcd{};
    qcd.base = 0;
    qcd.n = nq;
    qcd.row_off = 0;
    qcd.k = k;
    qcd.thres_n = thres_n;
    if (nq > 0) {
        float4* qn = h->scratch.get<float4>(32, nq);
        uint64_t* cellp = h->scratch.get<uint64_t>(33, nq);
        uint32_t* code = h->scratch.get<uint32_t>(34, nq);
        uint32_t* idx = h->scratch.get<uint32_t>(35, nq);
        const int g = ceil_div(nq, KBLK);
        k_query_prep<<<g, KBLK, 0, s>>>(dqx, dqy, dqz, nq, h->min_extent, h->max_extent, P.a, P.ba, h->B, qn, cellp,
                                        code, idx);
        LAUNCH_CHECK();
        radix_sort_pairs(code, idx, nq, 30, s, h->scratch);
        k_query_gather<<<g, KBLK, 0, s>>>(qn, cellp, idx, nq, h->qpts.p, h->qvcell.p);
        LAUNCH_CHECK();
        h->dqcloud.alloc(1);
        HIP_CHECK(hipMemcpyAsync(h->dqcloud.p, &qcd, sizeof(CloudDev), hipMemcpyHostToDevice, s));
        P.qpts = h->qpts.p;
        P.qvcell = reinterpret_cast<const unsigned long long*>(h->qvcell.p);
        P.knn = h->qknn.p;
        P.kth = nullptr;
        P.clouds = h->dqcloud.p;
        P.mcloud = nullptr;
        P.qlevel = nullptr;
        P.qbound = nullptr;
        P.retry = nullptr;
        P.chunks = nullptr;
        P.queue = nullptr;
        P.stride = 1;
        P.nchunks = (nq + 63) / 64;
        launch_any(P, kclass(k), s, true);
    }
    HIP_CHECK(hipEventRecord(e1, s));
    uint64_t hc[5];
    HIP_CHECK(hipMemcpyAsync(hc, P.counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));  // also keeps qcd alive until its copy has run
    HIP_CHECK(hipEventElapsedTime(&h->stats.query_ms, e0, e1));
    h->stats.query_count = nq;
    h->stats.query_k = k;
    h->stats.query_candidates = (int64_t)hc[0];
    h->stats.query_invalid_rows = (int64_t)hc[1];
    h->have_query = true;
}

void run_knn_query_original(egt_ctx* h, int32_t* out) {
    const int64_t cnt = h->nq * h->kq;
    if (cnt <= 0) return;
    hipStream_t s = h->stream;
    int32_t* tmp = h->scratch.get<int32_t>(9, cnt);
    k_query_original<<<ceil_div(cnt, KBLK), KBLK, 0, s>>>(h->qknn.p, h->perm.p, cnt, tmp);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(out, tmp, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// Counters of the search kernel (KnnParams::counters) in the allocation-free query search's scratch.
constexpr int QNA_COUNTERS = 16;
static_assert(QNA_COUNTERS >= 9, "every counter k_knn writes");

// Clears the counters and writes the query set's cloud entry (kernel arguments, so the sequence holds no host
// copy and no memset).
__global__ void k_qna_init(unsigned long long* __restrict__ counters, CloudDev* __restrict__ cloud, CloudDev c) {
    if (threadIdx.x < QNA_COUNTERS) counters[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) *cloud = c;
}

}  // namespace egt

// Temporaries of the query search without allocation: one copy per GPU, held until the process ends. Host code
// never uses this name as an address (on the host it only names the symbol): each device's address comes
// from hipGetSymbolAddress with that device current.
static __device__ __align__(256) uint32_t scratch[egt::QNA_WORDS];

namespace egt {
namespace {

constexpr int QNA_DEVICES = 64;

// Address of this device's copy of the scratch array, looked up once per device (dev must be current).
uint32_t* qna_scratch(int dev) {
    static std::mutex mu;
    static uint32_t* ptr[QNA_DEVICES] = {};
    if (dev < 0 || dev >= QNA_DEVICES) throw ArgError{EGT_ERR_UNSUPPORTED, "device index beyond 63"};
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<uint32_t*>(p);
    }
    return ptr[dev];
}

}  // namespace

// The query-vs-reference search of run_knn_query with every temporary in the device scratch array and the rows
// written to the caller's buffer (original query order, reference sorted indices): no allocation, no memset,
// no host copy and no synchronization; the work is queued on the handle's stream. Layout in 32-bit words,
// NQ = QNA_MAX_QUERIES: normalized queries (4 NQ), their virtual cells (2 NQ), sort keys (NQ) and values
// (NQ), the queries in Morton order (4 NQ; the radix sort's workspace before they are written), their cells
// (2 NQ), then the counters and the query set's cloud entry.
void run_knn_query_noalloc(egt_ctx* h, const float* dqx, const float* dqy, const float* dqz, int64_t nq, int k,
                           float thres_n, int32_t* rows) {
    if (nq <= 0) return;
    if (nq > QNA_MAX_QUERIES) throw ArgError{EGT_ERR_UNSUPPORTED, "more queries than the scratch holds"};
    constexpr size_t NQ = (size_t)QNA_MAX_QUERIES;
    static_assert(QNA_WORDS >= 14 * NQ + 2 * QNA_COUNTERS + (sizeof(CloudDev) + 3) / 4, "scratch layout");
    if (radix_sort_workspace_words(nq) > 4 * NQ) throw ArgError{EGT_ERR_UNSUPPORTED, "sort workspace"};
    uint32_t* w = qna_scratch(h->device);
    float4* qn = reinterpret_cast<float4*>(w);
    uint64_t* cellp = reinterpret_cast<uint64_t*>(w + 4 * NQ);
    uint32_t* code = w + 6 * NQ;
    uint32_t* idx = w + 7 * NQ;
    float4* qpts = reinterpret_cast<float4*>(w + 8 * NQ);
    uint64_t* qcell = reinterpret_cast<uint64_t*>(w + 12 * NQ);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(w + 14 * NQ);
    CloudDev* qcloud = reinterpret_cast<CloudDev*>(w + 14 * NQ + 2 * QNA_COUNTERS);
    hipStream_t s = h->stream;
    KnnParams P = knn_params(h, false);
    P.long_cap = 0;
    CloudDev qcd{};
    qcd.base = 0;
    qcd.n = nq;
    qcd.row_off = 0;
    qcd.k = k;
    qcd.thres_n = thres_n;
    k_qna_init<<<1, 64, 0, s>>>(counters, qcloud, qcd);
    LAUNCH_CHECK();
    const int g = ceil_div(nq, KBLK);
    k_query_prep<<<g, KBLK, 0, s>>>(dqx, dqy, dqz, nq, h->min_extent, h->max_extent, P.a, P.ba, h->B, qn, cellp,
                                    code, idx);
    LAUNCH_CHECK();
    radix_sort_pairs_ws(code, idx, nq, 30, s, w + 8 * NQ);
    k_query_gather<<<g, KBLK, 0, s>>>(qn, cellp, idx, nq, qpts, qcell);
    LAUNCH_CHECK();
    P.qpts = qpts;
    P.qvcell = reinterpret_cast<const unsigned long long*>(qcell);
    P.knn = rows;
    P.kth = nullptr;
    P.counters = counters;
    P.clouds = qcloud;
    P.mcloud = nullptr;
    P.qlevel = nullptr;
    P.qbound = nullptr;
    P.retry = nullptr;
    P.chunks = nullptr;
    P.queue = nullptr;
    P.stride = 1;
    P.nchunks = (nq + 63) / 64;
    launch_any(P, kclass(k), s, true);
}

// Autotuner for the search's start level (the cell size of the first search box). For each candidate level
// it runs the plain search on an evenly spread sample of chunks (about 8192 chunks of 64 Morton-consecutive
// queries) and reads back the work counters. The score per sampled query is candidates + 3 x cells looked up
// (a cell lookup costs about three candidate tests on gfx950: a dependent hash probe and range fetch against a
// coalesced distance test; cells looked up include the binary-search steps of sub-cell ranges). Counters,
// not the sample's wall time, rank the levels: a sample's time is set by its slowest wave (isolated queries
// next to dense clusters), not by the bulk of the work. Rows do not depend on the choice. Levels -SUB .. top
// are tried; trial_ms / trial_cost get each level's sample time and score at index level + SUB (-1 for
// untried levels).
// Launch-shape tuner: times the handle's current search (start level, learned proposal) over all queries in
// each launch shape (one lane per query with one workgroup per 256 queries, the same as one persistent
// work-queue launch, one wave per query), each after a
// warm-up on the autotuner's sample, and keeps the fastest. A sample's time would mislead here: it is set by
// its slowest wave, and one query per lane-per-query wave can test tens of thousands of candidates (on a
// 16M-point cluster mixture the 3% sample took 29 ms in that shape against 3.8 ms with a wave per query, while
// the full searches took 65 and 76 ms). Rows do not depend on the shape.
int autotune_shape(egt_ctx* h, int k, float thres_n, float* shape_ms) {
    hipStream_t s = h->stream;
    h->clouds[0].k = k;
    h->clouds[0].thres_n = thres_n;
    prepare_clouds(h);
    KnnParams P = knn_params(h);
    HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));  // scratch holds earlier contents
    const int64_t total = (h->n + 63) / 64;
    KnnParams W = P;  // warm-up: the autotuner's sample
    W.stride = (int)std::max<int64_t>(1, total / 8192);
    W.nchunks = (total + W.stride - 1) / W.stride;
    P.stride = 1;
    P.nchunks = total;
    const int K = kclass(k);
    Event e0, e1;
    auto launch = [&](const KnnParams& B, int shape) {
        search_class(h, B, K, s, shape);  // with the hand-back pass of the lane-per-query shapes
    };
    int best = 0;
    float best_ms = 0.0f;
    for (int shape = 0; shape < 3; ++shape) {
        shape_ms[shape] = -1.0f;
        launch(W, shape);
        HIP_CHECK(hipEventRecord(e0, s));
        launch(P, shape);
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipStreamSynchronize(s));
        float t = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        shape_ms[shape] = t;
        if (shape == 0 || t < best_ms) {
            best = shape;
            best_ms = t;
        }
    }
    h->knn_persistent = best == 1 ? 1 : 0;
    h->knn_lanes = best == 2 ? 64 : 1;
    h->class_shape[K / 8 - 1] = -1;  // the handle-wide options now say it
    h->have_knn = false;  // the trial rows cover only the sample
    return best;
}

// Per-class launch-shape tuner (batches, where the neighbour-count classes differ): for each class of the
// current clouds, times the class's full search in each shape after one warm-up launch of that shape, and
// keeps the fastest in the handle's per-class choice. ms: 8 classes x 3 shapes, -1 where not run. Returns the
// choices packed as EGT_OPT_KNN_CLASS_SHAPES takes them.
int64_t autotune_class_shapes(egt_ctx* h, float* ms) {
    hipStream_t s = h->stream;
    prepare_clouds(h);
    KnnParams P = knn_params(h);
    HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));  // scratch holds earlier contents
    const std::vector<WorkClass> classes = build_work_lists(h, h->work_lists);
    for (int i = 0; i < 24; ++i) ms[i] = -1.0f;
    Event e0, e1;
    int64_t packed = 0;
    for (const WorkClass& cr : classes) {
        P.chunks = h->nclouds > 1 ? h->work_lists.p + cr.first : nullptr;
        P.nchunks = cr.count;
        const int ci = cr.K / 8 - 1;
        int best = 0;
        float best_ms = 0.0f;
        for (int shape = 0; shape < 3; ++shape) {
            search_class(h, P, cr.K, s, shape);
            HIP_CHECK(hipEventRecord(e0, s));
            search_class(h, P, cr.K, s, shape);
            HIP_CHECK(hipEventRecord(e1, s));
            HIP_CHECK(hipStreamSynchronize(s));
            float t = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
            ms[ci * 3 + shape] = t;
            if (shape == 0 || t < best_ms) {
                best = shape;
                best_ms = t;
            }
        }
        h->class_shape[ci] = best;
        packed |= (int64_t)(best + 1) << (2 * ci);
    }
    h->have_knn = false;  // as after the other tuners: the caller runs the search for its rows and statistics
    return packed;
}

// Learned-start tuner: with a proposal in place, times the full search for each level offset (0, +1: the
// proposed levels one level coarser), each use of the proposal (its level alone, or with its radius bound) and
// each of the lane-per-query and wave-per-query shapes, each after a warm-up on the autotuner's sample, and
// keeps the fastest. ms: [offset][mode][shape] = 2 x 2 x 2. Measured on a 4M-point torus: the offset +1 takes
// the learned search from 11.2 to 9.1 ms (the margin the model adds for its level is tighter than this cloud
// needs); on a 16M-point cluster mixture from 54.1 to 52.3 ms with the bound.
void autotune_learned(egt_ctx* h, int k, float thres_n, int* best_off, int* best_mode, int* best_shape, float* ms) {
    hipStream_t s = h->stream;
    h->clouds[0].k = k;
    h->clouds[0].thres_n = thres_n;
    prepare_clouds(h);
    const int K = kclass(k);
    const int64_t total = (h->n + 63) / 64;
    Event e0, e1;
    float best_ms = 0.0f;
    *best_off = 0;
    *best_mode = 1;
    *best_shape = 0;
    for (int off = 0; off < 2; ++off) {
        for (int mi = 0; mi < 2; ++mi) {
            h->learned_level_offset = off;
            h->learned_mode = mi == 0 ? 1 : 3;
            KnnParams P = knn_params(h);
            HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));
            KnnParams W = P;
            W.stride = (int)std::max<int64_t>(1, total / 8192);
            W.nchunks = (total + W.stride - 1) / W.stride;
            P.stride = 1;
            P.nchunks = total;
            for (int si = 0; si < 2; ++si) {
                const int shape = si == 0 ? 0 : 2;
                search_class(h, W, K, s, shape);
                HIP_CHECK(hipEventRecord(e0, s));
                search_class(h, P, K, s, shape);
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipStreamSynchronize(s));
                float t = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                ms[(off * 2 + mi) * 2 + si] = t;
                if ((off | mi | si) == 0 || t < best_ms) {
                    best_ms = t;
                    *best_off = off;
                    *best_mode = h->learned_mode;
                    *best_shape = shape;
                }
            }
        }
    }
    h->learned_level_offset = *best_off;
    h->learned_mode = *best_mode;
    h->knn_lanes = *best_shape == 2 ? 64 : 1;
    h->knn_persistent = 0;
    h->class_shape[K / 8 - 1] = -1;
    h->have_knn = false;  // as after the other tuners: the caller runs the search for its rows and statistics
}

int autotune_start_level(egt_ctx* h, int k, float thres_n, float* trial_ms, float* trial_cost) {
    hipStream_t s = h->stream;
    h->clouds[0].k = k;
    h->clouds[0].thres_n = thres_n;
    prepare_clouds(h);
    const int saved_learned = h->learned;
    h->learned = 0;
    KnnParams P = knn_params(h);
    h->learned = saved_learned;
    P.long_cap = 0;  // every query's whole work counts
    const int64_t total = (h->n + 63) / 64;
    P.stride = (int)std::max<int64_t>(1, total / 8192);
    P.nchunks = (total + P.stride - 1) / P.stride;
    const int K = kclass(k);
    const int top = std::min(h->nlev - 1, std::max(3, auto_start_level(h, k) + 1));
    for (int i = 0; i < EGT_AUTOTUNE_SLOTS; ++i) trial_ms[i] = trial_cost[i] = -1.0f;
    Event e0, e1;
    int best = 0;
    double best_cost = 0.0;
    for (int l = -SUB; l <= top; ++l) {
        P.start_level = l + SUB;
        HIP_CHECK(hipMemsetAsync(P.counters, 0, 10 * sizeof(uint64_t), s));
        HIP_CHECK(hipEventRecord(e0, s));
        launch_any(P, K, s);
        HIP_CHECK(hipEventRecord(e1, s));
        uint64_t hc[5];
        HIP_CHECK(hipMemcpyAsync(hc, P.counters, sizeof(hc), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double q = (double)std::min<int64_t>(h->n, P.nchunks * 64);
        const double cost = ((double)hc[0] + 3.0 * (double)hc[4]) / q;
        trial_ms[l + SUB] = ms;
        trial_cost[l + SUB] = (float)cost;
        if (l == -SUB || cost < best_cost) {
            best = l;
            best_cost = cost;
        }
    }
    h->start_level = best;
    h->have_knn = false;  // the trial rows cover only the sample
    return best;
}

}  // namespace egt
