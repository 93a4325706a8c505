// Taubin lambda/mu smoothing over the EGG adjacency (CSR), Jacobi ping-pong.
//
// Reference: taubin_step_GKNN EGT_CUDA.cu:229-276, iteration loop :545-634 (one iteration = a lambda pass
// then a mu pass; weights w = exp(-|pn - pi|^2), isRegularized = 0 :544).
//
// Points live in Morton order as float4 (x, y, z, pad) in two buffers, so a wave's 64 points are spatial
// neighbours and their neighbour gathers (one 16-byte load each) mostly hit the same cache lines. Blocks
// are remapped so each of the 8 XCDs walks one contiguous Morton range (workgroups are dealt to XCDs round
// robin), which keeps a block's neighbours in its own XCD's L2. The 2*iterations passes are recorded once
// into a hipGraph and replayed. The exp is egt_expf, so results match the host oracle bit for bit; the
// per-point arithmetic and neighbour order are those of the reference, only the memory layout differs.
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int TBLK = 256;
constexpr int NXCD = 8;

__device__ inline int64_t xcd_block(int nblocks) {
    const int b = blockIdx.x;
    const int per = (nblocks + NXCD - 1) / NXCD;
    return (int64_t)(b % NXCD) * per + b / NXCD;
}

__global__ void __launch_bounds__(TBLK) k_taubin_pass(const float4* __restrict__ in, float4* __restrict__ out,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col, int64_t n, float scale,
                                                       int nblocks) {
    const int64_t lb = xcd_block(nblocks);
    if (lb >= nblocks) return;
    const int64_t i = lb * TBLK + threadIdx.x;
    if (i >= n) return;
    const float4 p = in[i];
    float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
    const int32_t e0 = rowptr[i], e1 = rowptr[i + 1];
    // CH column indices, then CH gathers, are issued before any use so the row costs ~2 memory latencies
    // instead of 2 per neighbour; the accumulation below stays in slot order.
    constexpr int CH = 16;
    for (int32_t e = e0; e < e1; e += CH) {
        int32_t c[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) c[u] = e + u < e1 ? col[e + u] : -1;
        float4 q[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) q[u] = c[u] >= 0 ? in[c[u]] : p;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (c[u] >= 0) {
                const float dist = sqd(p.x, p.y, p.z, q[u].x, q[u].y, q[u].z);
                const float w = egt_expf_taubin(-dist);
                cgx = __builtin_fmaf(w, q[u].x - p.x, cgx);
                cgy = __builtin_fmaf(w, q[u].y - p.y, cgy);
                cgz = __builtin_fmaf(w, q[u].z - p.z, cgz);
                sum += w;
            }
        }
    }
    if (sum == 0.0f) sum = 1e-8f;
    out[i] = make_float4(p.x + scale * cgx / sum, p.y + scale * cgy / sum, p.z + scale * cgz / sum, 0.0f);
}

// ---------------------------------------------------------------------------------------------------
// LDS-tiled pass. A tile is TILE consecutive Morton positions (TILE = 1024 or 2048, 1024 threads, so one or
// two points per thread); its 1-hop halo (neighbours outside the tile, deduplicated and sorted, at most
// TILE entries) is listed once per smoothing call, and every CSR column is rewritten as a 16-bit local index
// (< TILE: inside the tile; TILE + h: halo entry h). A pass then stages tile + halo coordinates in LDS and
// reads all neighbours from LDS. Tiles whose halo exceeds TILE use global gathers. Two points per thread
// (TILE 2048) give each thread two independent gather chains; on a 4M-point torus at k=40 a pass takes
// 57 ms / 600 vs 75 ms / 600 with one point per thread.
constexpr int PBLK = 256;          // tile preparation
constexpr int TBLK_TILE = 1024;  // tiled pass: TILE / TBLK_TILE points per thread
constexpr int SORT_CAP = 8192;

__device__ inline void block_bitonic_sort(uint32_t* a, int n2) {
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n2; i += PBLK) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

template <int TILE>
__global__ void __launch_bounds__(PBLK) k_tile_prep(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                     int64_t n, uint16_t* __restrict__ col16,
                                                     int32_t* __restrict__ halo, int32_t* __restrict__ hcount) {
    __shared__ uint32_t raw[SORT_CAP];
    __shared__ uint32_t uniq[SORT_CAP];
    __shared__ uint32_t cnt, wsum[PBLK / 64];
    const int64_t t = blockIdx.x;
    const int64_t base = t * TILE, end = min<int64_t>(base + TILE, n);
    const int32_t eb = rowptr[base], ee = rowptr[end];
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int32_t e = eb + threadIdx.x; e < ee; e += PBLK) {
        const int32_t c = col[e];
        if (c < base || c >= end) {
            const uint32_t p = atomicAdd(&cnt, 1u);
            if (p < SORT_CAP) raw[p] = (uint32_t)c;
        }
    }
    __syncthreads();
    const uint32_t total = cnt;
    if (total > SORT_CAP) {
        if (threadIdx.x == 0) hcount[t] = -1;
        return;
    }
    int n2 = 1;
    while (n2 < (int)total) n2 <<= 1;
    for (int i = total + threadIdx.x; i < n2; i += PBLK) raw[i] = 0xFFFFFFFFu;
    __syncthreads();
    block_bitonic_sort(raw, n2);
    // unique: each thread owns a contiguous chunk
    const int per = (n2 + PBLK - 1) / PBLK;
    const int c0 = threadIdx.x * per, c1 = min(c0 + per, (int)total);
    uint32_t mine = 0;
    for (int i = c0; i < c1; ++i) mine += (i == 0 || raw[i] != raw[i - 1]);
    uint32_t v = mine;
    const int lane = __lane_id(), wid = threadIdx.x / 64;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    uint32_t off = v - mine, H = 0;
    for (int w = 0; w < PBLK / 64; ++w) {
        if (w < wid) off += wsum[w];
        H += wsum[w];
    }
    for (int i = c0; i < c1; ++i)
        if (i == 0 || raw[i] != raw[i - 1]) uniq[off++] = raw[i];
    __syncthreads();
    if (H > TILE) {
        if (threadIdx.x == 0) hcount[t] = -1;
        return;
    }
    for (uint32_t h = threadIdx.x; h < H; h += PBLK) halo[t * TILE + h] = (int32_t)uniq[h];
    for (int32_t e = eb + threadIdx.x; e < ee; e += PBLK) {
        const uint32_t c = (uint32_t)col[e];
        uint16_t loc;
        if (c >= (uint64_t)base && c < (uint64_t)end) {
            loc = (uint16_t)(c - base);
        } else {
            uint32_t lo = 0, hi = H;  // lower_bound in uniq
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (uniq[mid] < c) lo = mid + 1;
                else hi = mid;
            }
            loc = (uint16_t)(TILE + lo);
        }
        col16[e] = loc;
    }
    if (threadIdx.x == 0) hcount[t] = (int32_t)H;
}

// Neighbour loop over LDS coordinates: CH local columns are fetched before use so a row costs about one
// global latency; accumulation stays in slot order.
__device__ inline void taubin_point(const float4 p, const float4* __restrict__ src, const uint16_t* __restrict__ c16,
                                    int32_t e0, int32_t e1, float& cgx, float& cgy, float& cgz, float& sum) {
    constexpr int CH = 16;
    for (int32_t e = e0; e < e1; e += CH) {
        uint32_t c[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) c[u] = e + u < e1 ? (uint32_t)c16[e + u] : 0xFFFFFFFFu;
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (c[u] != 0xFFFFFFFFu) {
                const float4 q = src[c[u]];
                const float dist = sqd(p.x, p.y, p.z, q.x, q.y, q.z);
                const float w = egt_expf_taubin(-dist);
                cgx = __builtin_fmaf(w, q.x - p.x, cgx);
                cgy = __builtin_fmaf(w, q.y - p.y, cgy);
                cgz = __builtin_fmaf(w, q.z - p.z, cgz);
                sum += w;
            }
        }
    }
}

template <int TILE>
__global__ void __launch_bounds__(TBLK_TILE) k_taubin_tile(const float4* __restrict__ in, float4* __restrict__ out,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const uint16_t* __restrict__ col16,
                                                       const int32_t* __restrict__ halo,
                                                       const int32_t* __restrict__ hcount, int64_t n, float scale,
                                                       int ntiles) {
    extern __shared__ float4 lds[];
    __shared__ int32_t srow[TILE + 1];
    const int64_t t = xcd_block(ntiles);
    if (t >= ntiles) return;
    const int64_t base = t * TILE, end = min<int64_t>(base + TILE, n);
    const int32_t H = hcount[t];
    for (int i = threadIdx.x; i <= end - base; i += TBLK_TILE) srow[i] = rowptr[base + i];
    if (H >= 0) {
        for (int i = threadIdx.x; i < end - base; i += TBLK_TILE) lds[i] = in[base + i];
        for (int h = threadIdx.x; h < H; h += TBLK_TILE) lds[TILE + h] = in[halo[t * TILE + h]];
    }
    __syncthreads();
    for (int64_t i = base + threadIdx.x; i < end; i += TBLK_TILE) {
        const float4 p = H >= 0 ? lds[i - base] : in[i];
        float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
        const int32_t e0 = srow[i - base], e1 = srow[i - base + 1];
        if (H >= 0) {
            taubin_point(p, lds, col16, e0, e1, cgx, cgy, cgz, sum);
        } else {
            for (int32_t e = e0; e < e1; ++e) {
                const float4 q = in[col[e]];
                const float dist = sqd(p.x, p.y, p.z, q.x, q.y, q.z);
                const float w = egt_expf_taubin(-dist);
                cgx = __builtin_fmaf(w, q.x - p.x, cgx);
                cgy = __builtin_fmaf(w, q.y - p.y, cgy);
                cgz = __builtin_fmaf(w, q.z - p.z, cgz);
                sum += w;
            }
        }
        if (sum == 0.0f) sum = 1e-8f;
        out[i] = make_float4(p.x + scale * cgx / sum, p.y + scale * cgy / sum, p.z + scale * cgz / sum, 0.0f);
    }
}

// Writes the smoothed points back to the sorted-space SoA arrays and refreshes the Morton copy.
__global__ void __launch_bounds__(TBLK) k_writeback(const float4* __restrict__ p, int64_t n,
                                                     float4* __restrict__ mpts, float* __restrict__ x,
                                                     float* __restrict__ y, float* __restrict__ z) {
    const int64_t m = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (m >= n) return;
    const float4 v = p[m];
    const int32_t s = __float_as_int(mpts[m].w);
    mpts[m] = make_float4(v.x, v.y, v.z, __int_as_float(s));
    x[s] = v.x;
    y[s] = v.y;
    z[s] = v.z;
}

}  // namespace

namespace {

// Tile preparation for tile width TILE (once per smoothing call): halo lists and 16-bit local columns.
struct TilePlan {
    int tile = 0, ntiles = 0, maxH = 0;
    int64_t fallback_tiles = 0;
    uint16_t* col16 = nullptr;
    int32_t* halo = nullptr;
    int32_t* hcount = nullptr;
};

template <int TILE>
TilePlan prepare_tiles(egt_ctx* h) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    TilePlan tp;
    tp.tile = TILE;
    tp.ntiles = ceil_div(n, TILE);
    tp.col16 = h->scratch.get<uint16_t>(19, std::max<int64_t>(h->nnz, 1));
    tp.halo = h->scratch.get<int32_t>(20, (size_t)tp.ntiles * TILE);
    tp.hcount = h->scratch.get<int32_t>(21, tp.ntiles);
    k_tile_prep<TILE><<<tp.ntiles, PBLK, 0, s>>>(h->rowptr.p, h->col.p, n, tp.col16, tp.halo, tp.hcount);
    LAUNCH_CHECK();
    std::vector<int32_t> hc(tp.ntiles);
    HIP_CHECK(hipMemcpyAsync(hc.data(), tp.hcount, tp.ntiles * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int32_t v : hc) {
        if (v < 0) ++tp.fallback_tiles;
        else tp.maxH = std::max(tp.maxH, v);
    }
    return tp;
}

void launch_pass(egt_ctx* h, const TilePlan& tp, const float4* in, float4* out, float scale) {
    const size_t lds_bytes = (size_t)(tp.tile + tp.maxH) * sizeof(float4);
    const dim3 tgrid(ceil_div(tp.ntiles, NXCD) * NXCD), tblk(TBLK_TILE);
    if (tp.tile == 2048)
        k_taubin_tile<2048><<<tgrid, tblk, lds_bytes, h->stream>>>(in, out, h->rowptr.p, h->col.p, tp.col16, tp.halo,
                                                                   tp.hcount, h->n, scale, tp.ntiles);
    else
        k_taubin_tile<1024><<<tgrid, tblk, lds_bytes, h->stream>>>(in, out, h->rowptr.p, h->col.p, tp.col16, tp.halo,
                                                                   tp.hcount, h->n, scale, tp.ntiles);
}

TilePlan prepare(egt_ctx* h, int tile) { return tile == 2048 ? prepare_tiles<2048>(h) : prepare_tiles<1024>(h); }

// Tile-width autotuner: prepares each candidate width and times three lambda/mu pass pairs on copies of
// the points (the smoothing itself is untouched), keeping the faster width for this handle.
int tune_tile(egt_ctx* h, const float4* A, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    float4* C = h->scratch.get<float4>(22, n);
    float4* D = h->scratch.get<float4>(23, n);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    int best = 2048;
    float best_ms = 0.0f;
    for (int tile : {1024, 2048}) {
        const TilePlan tp = prepare(h, tile);
        HIP_CHECK(hipMemcpyAsync(C, A, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
        launch_pass(h, tp, C, D, lambda);  // warm-up
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < 3; ++r) {
            launch_pass(h, tp, D, C, mu);
            launch_pass(h, tp, C, D, lambda);
        }
        HIP_CHECK(hipEventRecord(e1, s));
        LAUNCH_CHECK();
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        h->stats.taubin_tile_ms[tile == 2048 ? 1 : 0] = ms / 6.0f;
        if (tile == 1024 || ms < best_ms) {
            best = tile;
            best_ms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

void run_taubin(egt_ctx* h, int iters, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    float4* A = h->scratch.get<float4>(12, n);
    float4* Bf = h->scratch.get<float4>(13, n);
    HIP_CHECK(hipMemcpyAsync(A, h->mpts.p, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
    if (h->taubin_tile == 0 && iters > 0) h->taubin_tile = tune_tile(h, A, lambda, mu);
    const int tile = h->taubin_tile == 1024 ? 1024 : 2048;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    const TilePlan tp = prepare(h, tile);
    h->taubin_fallback_tiles = tp.fallback_tiles;
    h->stats.taubin_tile = tile;
    h->stats.taubin_fallback_tiles = (int32_t)tp.fallback_tiles;
    auto pass = [&](const float4* in, float4* out, float scale) { launch_pass(h, tp, in, out, scale); };
    if (iters > 0) {
        if (h->taubin_graph) {
            hipGraph_t g;
            hipGraphExec_t ge;
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < iters; ++it) {
                pass(A, Bf, lambda);
                pass(Bf, A, mu);
            }
            HIP_CHECK(hipStreamEndCapture(s, &g));
            HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphLaunch(ge, s));
            HIP_CHECK(hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
        } else {
            for (int it = 0; it < iters; ++it) {
                pass(A, Bf, lambda);
                pass(Bf, A, mu);
            }
            LAUNCH_CHECK();
        }
    }
    HIP_CHECK(hipEventRecord(e1, s));
    k_writeback<<<ceil_div(n, TBLK), TBLK, 0, s>>>(A, n, h->mpts.p, h->x.p, h->y.p, h->z.p);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.taubin_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->search_stale = true;
}

}  // namespace egt
