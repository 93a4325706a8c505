// Taubin lambda/mu smoothing over the EGG adjacency (CSR), Jacobi ping-pong.
//
// Reference: taubin_step_GKNN EGT_CUDA.cu:229-276, iteration loop :545-634 (one iteration = a lambda pass
// then a mu pass; weights w = exp(-|pn - pi|^2), isRegularized = 0 :544).
//
// Points live in Morton order as packed x, y, z (12 bytes) in two buffers, so a wave's 64 points are spatial
// neighbours and their gathers mostly hit the same cache lines; a tile stages them in LDS as float4. Blocks
// are remapped so each of the 8 XCDs walks chunks of consecutive Morton tiles (workgroups are dealt to XCDs
// round robin), which keeps a block's neighbours in its own XCD's L2. The 2*iterations passes are recorded
// into hipGraphs and replayed. The exp is egt_expf, so results match the host oracle bit for bit; the
// per-point arithmetic and neighbour order are those of the reference, only the memory layout differs.
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int TBLK = 256;
constexpr int NXCD = 8;
// Block -> tile map of a pass. Workgroups are dealt to the 8 XCDs round robin; XCD x takes the chunks x, x + 8,
// x + 16, ... of XCHUNK consecutive tiles. Neighbouring tiles share halo points, so a chunk mostly reads its
// halo from its own XCD's L2, and every XCD gets a share of every region of the cloud: a dense region (slower
// per tile) is spread over all XCDs instead of holding back the one whose contiguous range it falls in.
// The grid is padded to a multiple of NXCD * XCHUNK blocks, over which the map is a bijection.
constexpr int XCHUNK = 16;
// The passes' ping-pong points: x, y, z packed in 12 bytes (one dwordx3 access), a quarter less traffic than
// float4 in the staging loads, the halo gathers and the stores; LDS keeps float4 for aligned 16-byte reads.
struct __attribute__((aligned(4))) Pt3 {
    float x, y, z;
};
__device__ inline float4 ld3(const Pt3* __restrict__ a, int64_t i) {
    const Pt3 v = a[i];
    return make_float4(v.x, v.y, v.z, 0.0f);
}
__device__ inline void st3(Pt3* __restrict__ a, int64_t i, const float4 v) { a[i] = Pt3{v.x, v.y, v.z}; }
__device__ inline int64_t xcd_tile() {
    const int b = blockIdx.x;
    const int x = b % NXCD, j = b / NXCD;
    return ((int64_t)(j / XCHUNK) * NXCD + x) * XCHUNK + j % XCHUNK;
}

// ---------------------------------------------------------------------------------------------------
// LDS-tiled pass. A tile is TILE consecutive Morton positions (TILE = 1024 or 2048, two rows per thread); its
// 1-hop halo (neighbours outside the tile, deduplicated and sorted, at most halo_max<TILE>() entries) is
// listed once per smoothing call, and every CSR column is rewritten as a 16-bit local index
// (< TILE: inside the tile; TILE + h: halo entry h). The columns are then laid out as sliced ELLPACK: the
// tile's rows sorted by length, 64 rows (one wave) per slice, slot-pair-major within a slice. A pass stages
// tile + halo coordinates in LDS and reads all neighbours from LDS; the lanes of a wave walk rows of nearly
// equal length and each load of a slot pair is one coalesced 256-byte block. The pass is limited by VALU issue (the weight's
// exp per neighbour) and by the latency of its dependent loads; equal-length lanes cut the issued VALU. Only
// the first halo_cap<TILE>() halo entries fit in LDS; a tile with more reads the rest from global memory
// through its halo list. Tiles with more than halo_max<TILE>() halo entries use CSR rows with global gathers.
// Tile preparation runs one 1024-thread block per tile: its sorts are barrier-bound, so four times the threads
// of a 256-thread block take the set-up of a 4M-point torus from 2.9 to 1.1 ms.
constexpr int PBLK = 1024;
// Threads of a tiled pass (two rows per thread at both widths). 1024-point tiles run as 512-thread blocks, four
// per CU instead of two, and are the default width: 62.3 vs 64.0 us per pass on a 4M-point torus (k=40),
// 261 vs 265 us on a 16M-point cluster mixture.
template <int TILE>
constexpr int tile_threads() { return TILE == 1024 ? 512 : 1024; }
// Largest halo a tile keeps in LDS: (TILE + cap + 1) float4 stay 1 KiB under a half (2048) or a quarter (1024)
// of the CU's 160 KiB, so two 1024-thread or four 512-thread blocks still fit per CU. In
// volumetric regions a 2048-point Morton range has a halo of 2000-3000 points at 16 kept neighbours a point.
template <int TILE>
constexpr int halo_cap() { return TILE == 1024 ? 1456 : 2992; }
// Longest halo list of a tile (the list's stride per tile); beyond it the tile uses CSR rows.
template <int TILE>
constexpr int halo_max() { return TILE == 1024 ? 4096 : 8192; }
static_assert((2048 + halo_cap<2048>() + 1) * 16 <= 160 * 1024 / 2 - 1024, "two tile blocks per CU");
static_assert((1024 + halo_cap<1024>() + 1) * 16 <= 160 * 1024 / 4 - 1024, "four tile blocks per CU");
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

// The same sort for n2 <= PBLK keys, one per thread, held in a register: stages whose partner lies within the
// wave (j < 64) exchange through lane shuffles without a barrier, the others through LDS with two barriers.
// For 1024 keys that is 10 LDS stages (20 barriers) instead of 55 barrier stages. Keys of threads at or past n2
// never meet the others (n2 is a power of two, so i ^ j < n2 for i, j < n2). Falls back to block_bitonic_sort
// for longer arrays. Taubin set-up (tile preparation and graphs): 16M-point mixture 3.47 -> 3.30 ms, 4M-point
// torus 0.86 -> 0.79 ms, same tiles.
__device__ inline void block_sort(uint32_t* a, int n2) {
    if (n2 > PBLK) {
        block_bitonic_sort(a, n2);
        return;
    }
    const int i = threadIdx.x;
    uint32_t x = i < n2 ? a[i] : 0xFFFFFFFFu;
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint32_t p;
            if (j >= 64) {
                __syncthreads();  // the previous LDS stage's reads are done
                if (i < n2) a[i] = x;
                __syncthreads();
                p = i < n2 ? a[i ^ j] : x;
            } else {
                p = (uint32_t)__shfl_xor((int)x, j, 64);
            }
            const bool up = (i & k) == 0, low = (i & j) == 0;
            const uint32_t lo = x < p ? x : p, hi = x < p ? p : x;  // unsigned compare
            x = (low == up) ? lo : hi;
        }
    }
    __syncthreads();
    if (i < n2) a[i] = x;
    __syncthreads();
}

// LDS hash set of a tile's external columns: open addressing, linear probing, 2 * halo_max slots. An insert
// is only started while the count is at most halo_max, so the set never holds more than halo_max + PBLK
// entries and every probe sequence reaches an empty slot.
template <int TILE>
constexpr int hset_size() { return 2 * halo_max<TILE>(); }
constexpr uint32_t HS_EMPTY = 0xFFFFFFFFu;

template <int TILE>
__global__ void __launch_bounds__(PBLK) k_tile_prep(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                     int64_t n, uint16_t* __restrict__ col16,
                                                     int32_t* __restrict__ halo, int32_t* __restrict__ hcount,
                                                     uint16_t* __restrict__ perm, uint32_t* __restrict__ width,
                                                     uint32_t* __restrict__ padded) {
    constexpr int HS = hset_size<TILE>(), HMAX = halo_max<TILE>();
    static_assert((HS & (HS - 1)) == 0 && HS >= TILE && HS >= HMAX + PBLK, "hash set size");
    __shared__ uint32_t hs[HS];      // hash set, then the row-sort keys
    __shared__ uint32_t uniq[HMAX];  // the halo, sorted
    __shared__ uint32_t cnt, selfrow, wsum[PBLK / 64];
    __shared__ uint32_t swid[TILE / 64];
    const int64_t t = blockIdx.x;
    const int64_t base = t * TILE, end = min<int64_t>(base + TILE, n);
    const int np = (int)(end - base);
    const int32_t eb = rowptr[base], ee = rowptr[end];
    if (threadIdx.x == 0) {
        cnt = 0;
        selfrow = 0;
    }
    for (int i = threadIdx.x; i < HS; i += PBLK) hs[i] = HS_EMPTY;
    __syncthreads();
    // A row that lists itself (a mesh vertex of a degenerate triangle, GridMeshCuda1.cu:115-125) keeps that
    // entry's weight exp(0) = 1 in the reference's loop over the listed slots (EGT_CUDA.cu:242-270); the
    // sliced tile pass tells padding slots by their column equal to the row, so such a tile runs as CSR rows.
    for (int i = threadIdx.x; i < np; i += PBLK)
        for (int32_t e = rowptr[base + i]; e < rowptr[base + i + 1]; ++e)
            if (col[e] == base + i) selfrow = 1;
    for (int32_t e = eb + threadIdx.x; e < ee; e += PBLK) {
        const uint32_t c = (uint32_t)col[e];
        if ((int64_t)c >= base && (int64_t)c < end) continue;
        if (__atomic_load_n(&cnt, __ATOMIC_RELAXED) > (uint32_t)HMAX) continue;  // over the cap: CSR tile
        uint32_t slot = (c * 0x9E3779B1u) & (uint32_t)(HS - 1);
        for (;;) {
            const uint32_t old = atomicCAS(&hs[slot], HS_EMPTY, c);
            if (old == HS_EMPTY) {
                atomicAdd(&cnt, 1u);
                break;
            }
            if (old == c) break;
            slot = (slot + 1) & (uint32_t)(HS - 1);
        }
    }
    __syncthreads();
    const uint32_t H = cnt;
    if (H > (uint32_t)HMAX || selfrow) {
        if (threadIdx.x == 0) {
            hcount[t] = -1;
            padded[t] = 0;
        }
        return;
    }
    // compact the set into uniq: each thread owns a contiguous chunk of slots
    constexpr int PER = HS / PBLK;
    const int c0 = threadIdx.x * PER;
    uint32_t mine = 0;
    for (int i = c0; i < c0 + PER; ++i) mine += hs[i] != HS_EMPTY;
    uint32_t v = mine;
    const int lane = __lane_id(), wid = threadIdx.x / 64;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    uint32_t off = v - mine;
    for (int w = 0; w < wid; ++w) off += wsum[w];
    for (int i = c0; i < c0 + PER; ++i)
        if (hs[i] != HS_EMPTY) uniq[off++] = hs[i];
    int n2 = 1;
    while (n2 < (int)H) n2 <<= 1;
    __syncthreads();
    for (int i = (int)H + threadIdx.x; i < n2; i += PBLK) uniq[i] = HS_EMPTY;
    __syncthreads();
    block_sort(uniq, n2);
    for (uint32_t h = threadIdx.x; h < H; h += PBLK) halo[t * halo_max<TILE>() + h] = (int32_t)uniq[h];
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
    // Lane order of the pass: rows sorted by length (ties by position), so the 64 rows of a slice have
    // nearly equal lengths; a slice's width is its longest row.
    __syncthreads();
    for (int i = threadIdx.x; i < TILE; i += PBLK) {
        uint32_t key = 0xFFFFFFFFu;
        if (i < np) key = ((uint32_t)min(rowptr[base + i + 1] - rowptr[base + i], (1 << 20) - 1) << 12) | (uint32_t)i;
        hs[i] = key;
    }
    for (int s = threadIdx.x; s < TILE / 64; s += PBLK) swid[s] = 0;
    __syncthreads();
    block_sort(hs, TILE);
    for (int j = threadIdx.x; j < np; j += PBLK) {
        const uint32_t li = hs[j] & 0xFFFu;
        perm[t * TILE + j] = (uint16_t)li;
        atomicMax(&swid[j >> 6], (uint32_t)(rowptr[base + li + 1] - rowptr[base + li]));
    }
    __syncthreads();
    for (int s = threadIdx.x; s < TILE / 64; s += PBLK) width[t * (TILE / 64) + s] = swid[s];
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int s = 0; s < TILE / 64; ++s) tot += 128u * ((swid[s] + 1u) >> 1);
        padded[t] = tot;
        hcount[t] = (int32_t)H;
    }
}

// Sliced-ELLPACK columns of one tile: slice s (sorted positions 64 s .. 64 s + 63) stores its width W_s
// slots as ceil(W_s / 2) slot pairs, pair-major: slots 2p and 2p + 1 of lane l share the 32-bit word
// 64 p + l (low and high half), so a wave's load of one pair is 256 contiguous bytes. Slots past a row's
// end (and the high half of an odd width's last pair) hold the row's own local index (a tile with a row
// that lists itself runs as CSR rows instead), lanes past the tile's last row hold 0. Slot order within a
// row is the CSR (neighbour) order.
template <int TILE>
__global__ void __launch_bounds__(PBLK) k_tile_sell(const int32_t* __restrict__ rowptr,
                                                     const uint16_t* __restrict__ col16, int64_t n,
                                                     const uint16_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ width,
                                                     const uint64_t* __restrict__ toff,
                                                     const int32_t* __restrict__ hcount, uint16_t* __restrict__ scol) {
    const int64_t t = blockIdx.x;
    if (hcount[t] < 0) return;
    const int64_t base = t * TILE;
    const int np = (int)(min<int64_t>(base + TILE, n) - base);
    uint16_t* dst = scol + toff[t];
    uint32_t sb = 0;
    for (int s = 0; s < TILE / 64; ++s) {
        const uint32_t P = (width[t * (TILE / 64) + s] + 1u) >> 1;
        for (uint32_t x = threadIdx.x; x < 128u * P; x += PBLK) {
            const uint32_t u = ((x >> 7) << 1) | (x & 1u);
            const int j = s * 64 + (int)((x >> 1) & 63u);
            uint16_t v = 0;
            if (j < np) {
                const int li = perm[t * TILE + j];
                const int32_t e0 = rowptr[base + li];
                v = (int32_t)u < rowptr[base + li + 1] - e0 ? col16[e0 + u] : (uint16_t)li;
            }
            dst[sb + x] = v;
        }
        sb += 128u * P;
    }
}

__device__ inline float4 taubin_result(const float4 p, float cgx, float cgy, float cgz, float sum, float scale) {
    if (sum == 0.0f) sum = 1e-8f;
    return make_float4(p.x + scale * cgx / sum, p.y + scale * cgy / sum, p.z + scale * cgz / sum, 0.0f);
}

// Neighbour weight: exp(-d^2), or with isRegularized (EGT_CUDA.cu:263, dormant in the reference) 1 in the
// lambda pass and 1 / (d^2 + 1e-8) in the mu pass.
template <bool REG>
__device__ inline float taubin_weight(float dist, float scale) {
    if (REG) return scale < 0.0f ? 1.0f / (dist + 1e-8f) : 1.0f;
    return egt_expf_taubin(-dist);
}

// The neighbour loop of a wave's two rows over LDS coordinates (sliced ELLPACK, W0 and W1 slots, both uniform
// over the wave; this lane's slot pair q of row r is the word sb_r[64 q + lane]). The two rows are walked as
// one sequence of chunks of CH slots: while a chunk's coordinates are read from LDS and its weights
// computed, the next chunk's columns (CH / 2 packed words, possibly the second row's first ones) are already
// in flight, so the column loads' global latency overlaps the arithmetic. Two column buffers alternate (the
// loop is unrolled by two, so no register copy waits on a load in flight) and every prefetch issues all its
// loads, a word past the row's end reading the valid word dummy[lane] instead, so the wait before a chunk
// counts exactly the next chunk's loads. Packing two columns per word keeps both buffers in the registers
// one unpacked chunk took (CH = 8; 16 slots in flight cost more in VGPRs, and so in occupancy, than they
// save). Accumulation stays in slot order. A padding slot holds the row's own index (self): its distance is
// 0, its weight is selected to 0 and it adds fma(-0, 0, acc) = acc and sum + 0 = sum, so every lane runs
// every slot without a branch on the padding. The weighted difference uses w * (q - p) = (-w) * (p - q):
// q - p is exactly -(p - q), and an accumulator that starts at +0 never becomes -0, so the sign of a zero
// difference cannot change a sum; the reuse saves three subtractions per neighbour.
// With OVF, local columns at or past lim (TILE + the LDS part of the halo) are halo entries kept in global
// memory: their coordinates are gathered from gin through the tile's halo list hal (indexed by column - TILE).
constexpr int TCH = 8;  // slots per chunk
// Columns of chunk g of a wave's two rows (row 0's chunk g, or row 1's chunk g - C0; C0 = row 0's chunks)
// into w; a word past its row's end, and every word of a chunk g >= the chunk count, reads dummy[lane].
__device__ inline void taubin_fetch(int g, int C0, int W0, int W1, const uint32_t* sb0, const uint32_t* sb1,
                                    const uint32_t* dummy, int lane, uint32_t (&w)[TCH / 2]) {
    const bool second = g >= C0;
    const uint32_t* sb = second ? sb1 : sb0;
    const int q0 = (second ? g - C0 : g) * (TCH / 2), P = ((second ? W1 : W0) + 1) >> 1;
#pragma unroll
    for (int q = 0; q < TCH / 2; ++q) {
        const uint32_t* b = q0 + q < P ? sb + (q0 + q) * 64 : dummy;  // wave-uniform choice
        w[q] = b[lane];
    }
}

// ba holds chunk 0's columns (taubin_fetch(0, ...)), fetched by the caller.
template <bool REG, bool OVF>
__device__ inline void taubin_two_rows(const float4* __restrict__ src, const uint32_t* sb0, const uint32_t* sb1,
                                       const uint32_t* dummy, int lane, int W0, int W1, int r0, int r1, float scale,
                                       const Pt3* __restrict__ gin, const int32_t* __restrict__ hal, uint32_t lim,
                                       uint32_t tile, uint32_t (&ba)[TCH / 2], float4& res0, float4& res1) {
    constexpr int CH = TCH, CP = CH / 2;
    const int C0 = (W0 + CH - 1) / CH, C = C0 + (W1 + CH - 1) / CH;  // chunks of each row (wave-uniform)
    // own index as the padding slots hold it; lanes past the last row (r = -1) match no column
    const uint32_t self0 = (uint32_t)r0, self1 = (uint32_t)r1;
    float4 p = src[r0 < 0 ? 0 : r0];
    if (C0 == 0) {
        res0 = taubin_result(p, 0.0f, 0.0f, 0.0f, 0.0f, scale);
        p = src[r1 < 0 ? 0 : r1];
    }
    float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
    auto fetch = [&](int g, uint32_t(&w)[CP]) { taubin_fetch(g, C0, W0, W1, sb0, sb1, dummy, lane, w); };
    auto chunk = [&](int g, const uint32_t(&cw)[CP]) {
        const bool second = g >= C0;
        const int u0 = (second ? g - C0 : g) * CH, W = second ? W1 : W0;
        const uint32_t self = second ? self1 : self0;
        uint32_t c[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) c[u] = (u & 1) ? cw[u >> 1] >> 16 : cw[u >> 1] & 0xFFFFu;
        // every LDS read of the chunk is issued before the first use
        float4 qs[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (u0 + u < W) {  // wave-uniform
                if (!OVF || c[u] < lim) qs[u] = src[c[u]];
                else qs[u] = ld3(gin, hal[c[u] - tile]);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            if (u0 + u < W) {  // wave-uniform
                const float4 q = qs[u];
                const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                const float dist = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                float w = taubin_weight<REG>(dist, scale);
                w = c[u] != self ? w : 0.0f;
                cgx = __builtin_fmaf(-w, dx, cgx);
                cgy = __builtin_fmaf(-w, dy, cgy);
                cgz = __builtin_fmaf(-w, dz, cgz);
                sum += w;
            }
        }
        if (g == C0 - 1) {  // row 0 done: the rest of the sequence is row 1
            res0 = taubin_result(p, cgx, cgy, cgz, sum, scale);
            cgx = cgy = cgz = sum = 0.0f;
            p = src[r1 < 0 ? 0 : r1];
        }
    };
    uint32_t bb[CP];
    for (int g = 0; g < C; g += 2) {
        fetch(g + 1, bb);
        chunk(g, ba);
        if (g + 1 >= C) break;
        fetch(g + 2, ba);
        chunk(g + 1, bb);
    }
    res1 = taubin_result(p, cgx, cgy, cgz, sum, scale);
}

// One lambda or mu pass over one tile. Wave w of the block takes slices w and S - 1 - w (S = TILE / 64
// slices in ascending width), so the waves of a block carry similar work; lane l of slice s handles the
// row perm[64 s + l]. Results go back through LDS so the global writes are coalesced.
template <int TILE, bool REG>
__global__ void __launch_bounds__(tile_threads<TILE>()) __attribute__((amdgpu_waves_per_eu(8))) k_taubin_tile(const Pt3* __restrict__ in, Pt3* __restrict__ out,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const uint16_t* __restrict__ scol,
                                                       const uint64_t* __restrict__ toff,
                                                       const uint32_t* __restrict__ width,
                                                       const uint16_t* __restrict__ perm,
                                                       const int32_t* __restrict__ halo,
                                                       const int32_t* __restrict__ hcount, int64_t n, float scale,
                                                       int ntiles) {
    constexpr int S = TILE / 64;
    constexpr int TBLK_TILE = tile_threads<TILE>();
    constexpr int NW = TBLK_TILE / 64;
    static_assert(S <= 64 && S / NW == 2 && S % NW == 0, "two slices per wave");
    extern __shared__ float4 lds[];
    const int64_t t = xcd_tile();
    if (t >= ntiles) return;
    const int64_t base = t * TILE, end = min<int64_t>(base + TILE, n);
    const int np = (int)(end - base);
    const int32_t H = hcount[t];
    if (H < 0) {
        // halo larger than the tile, or a row that lists itself: CSR rows with global gathers
        for (int64_t i = base + threadIdx.x; i < end; i += TBLK_TILE) {
            const float4 p = ld3(in, i);
            float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
            for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
                const float4 q = ld3(in, col[e]);
                const float dist = sqd(p.x, p.y, p.z, q.x, q.y, q.z);
                const float w = taubin_weight<REG>(dist, scale);
                cgx = __builtin_fmaf(w, q.x - p.x, cgx);
                cgy = __builtin_fmaf(w, q.y - p.y, cgy);
                cgz = __builtin_fmaf(w, q.z - p.z, cgz);
                sum += w;
            }
            st3(out, i, taubin_result(p, cgx, cgy, cgz, sum, scale));
        }
        return;
    }
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Every wave scans the tile's slice widths itself (offsets of the slices' column blocks), so its first
    // chunk of columns can be in flight together with the staging loads, before the barrier.
    const uint32_t wd = lane < S ? width[t * S + lane] : 0u;
    uint32_t v = 128u * ((wd + 1u) >> 1);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    v -= 128u * ((wd + 1u) >> 1);  // exclusive
    const uint16_t* tc = scol + toff[t];
    int row[2];
    const uint32_t* sb[2];
    int W[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int s = r == 0 ? wv : S - 1 - wv;
        const int j = s * 64 + lane;
        row[r] = j < np ? (int)perm[t * TILE + j] : -1;
        // slice offset and width are wave-uniform: keep them scalar so the slot loops branch on SGPRs
        sb[r] = reinterpret_cast<const uint32_t*>(tc + __builtin_amdgcn_readfirstlane(__shfl(v, s, 64)));
        W[r] = (int)__builtin_amdgcn_readfirstlane(__shfl(wd, s, 64));
    }
    // Staging of tile + halo coordinates: every load is issued before the first wait (the halo indices,
    // then the coordinates), so the stage costs two global round trips, and the wave's first chunk of
    // columns rides along. Indices are clamped so that every load reads inside its array, and every store
    // is made (no load is moved into a branch): rows past the tile's end go to their unused LDS rows, halo
    // slots past the tile's halo to entry TILE + HL, which no column of this tile reads from LDS (the
    // allocation has one entry more than the largest staged halo).
    const int32_t* hal = halo + t * halo_max<TILE>();
    const int HL = min(H, halo_cap<TILE>());  // halo entries staged in LDS; the rest are read from global
    constexpr int TP = TILE / TBLK_TILE, HP = (halo_cap<TILE>() + TBLK_TILE - 1) / TBLK_TILE;
    int32_t hidx[HP];
#pragma unroll
    for (int k = 0; k < HP; ++k) hidx[k] = hal[min((int)threadIdx.x + k * TBLK_TILE, max(HL - 1, 0))];
    Pt3 tv[TP];
#pragma unroll
    for (int k = 0; k < TP; ++k) tv[k] = in[base + min((int)threadIdx.x + k * TBLK_TILE, np - 1)];
    // the first 64 words of the column array (it is allocated with at least that many)
    const uint32_t* dummy = reinterpret_cast<const uint32_t*>(scol);
    uint32_t ba[TCH / 2];
    taubin_fetch(0, (W[0] + TCH - 1) / TCH, W[0], W[1], sb[0], sb[1], dummy, lane, ba);
    Pt3 hv[HP];
#pragma unroll
    for (int k = 0; k < HP; ++k) hv[k] = in[(int)threadIdx.x + k * TBLK_TILE < HL ? (int64_t)hidx[k] : base];
    // x, y, z into the float4 entries (their w is never read)
#pragma unroll
    for (int k = 0; k < TP; ++k) *reinterpret_cast<Pt3*>(&lds[threadIdx.x + k * TBLK_TILE]) = tv[k];
#pragma unroll
    for (int k = 0; k < HP; ++k) {
        const int h = (int)threadIdx.x + k * TBLK_TILE;
        *reinterpret_cast<Pt3*>(&lds[TILE + (h < HL ? h : HL)]) = hv[k];
    }
    __syncthreads();
    float4 res[2];
    if (H > halo_cap<TILE>())
        taubin_two_rows<REG, true>(lds, sb[0], sb[1], dummy, lane, W[0], W[1], row[0], row[1], scale, in, hal,
                                   TILE + HL, TILE, ba, res[0], res[1]);
    else
        taubin_two_rows<REG, false>(lds, sb[0], sb[1], dummy, lane, W[0], W[1], row[0], row[1], scale, in, hal,
                                    TILE + HL, TILE, ba, res[0], res[1]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; ++r)
        if (row[r] >= 0) lds[row[r]] = res[r];
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += TBLK_TILE) out[base + i] = *reinterpret_cast<const Pt3*>(&lds[i]);
}

// Writes the smoothed points back to the sorted-space SoA arrays and refreshes the Morton copy.
__global__ void __launch_bounds__(TBLK) k_writeback(const Pt3* __restrict__ p, int64_t n,
                                                     float4* __restrict__ mpts, float* __restrict__ x,
                                                     float* __restrict__ y, float* __restrict__ z) {
    const int64_t m = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (m >= n) return;
    const float4 v = ld3(p, m);
    const int32_t s = __float_as_int(mpts[m].w);
    mpts[m] = make_float4(v.x, v.y, v.z, __int_as_float(s));
    x[s] = v.x;
    y[s] = v.y;
    z[s] = v.z;
}

// The Morton-order points (float4, w = sorted index) as packed x, y, z for the passes.
__global__ void __launch_bounds__(TBLK) k_pack3(const float4* __restrict__ mpts, int64_t n, Pt3* __restrict__ p) {
    const int64_t m = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (m >= n) return;
    st3(p, m, mpts[m]);
}

}  // namespace

namespace {

// Tile preparation for tile width TILE (once per smoothing call): halo lists and 16-bit local columns.
struct TilePlan {
    int tile = 0, ntiles = 0, maxH = 0;
    int64_t fallback_tiles = 0, overflow_tiles = 0;  // CSR tiles; tiles with part of the halo in global memory
    uint16_t* col16 = nullptr;   // CSR-order local columns
    int32_t* halo = nullptr;
    int32_t* hcount = nullptr;
    uint16_t* perm = nullptr;    // per tile: sorted position -> local row
    uint32_t* width = nullptr;   // per slice: width in slots
    uint64_t* toff = nullptr;    // per tile: start of its column block in scol
    uint16_t* scol = nullptr;    // sliced-ELLPACK local columns
    int64_t scol_elems = 0;
};

template <int TILE>
TilePlan prepare_tiles(egt_ctx* h) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    TilePlan tp;
    tp.tile = TILE;
    tp.ntiles = ceil_div(n, TILE);
    tp.col16 = h->scratch.get<uint16_t>(19, std::max<int64_t>(h->nnz, 1));
    tp.halo = h->scratch.get<int32_t>(20, (size_t)tp.ntiles * halo_max<TILE>());
    tp.hcount = h->scratch.get<int32_t>(21, tp.ntiles);
    tp.perm = h->scratch.get<uint16_t>(28, (size_t)tp.ntiles * TILE);
    tp.width = h->scratch.get<uint32_t>(29, (size_t)tp.ntiles * (TILE / 64));
    uint32_t* padded = h->scratch.get<uint32_t>(30, tp.ntiles);
    tp.toff = h->scratch.get<uint64_t>(31, tp.ntiles);
    k_tile_prep<TILE><<<tp.ntiles, PBLK, 0, s>>>(h->rowptr.p, h->col.p, n, tp.col16, tp.halo, tp.hcount, tp.perm,
                                                 tp.width, padded);
    LAUNCH_CHECK();
    std::vector<int32_t> hc(tp.ntiles);
    std::vector<uint32_t> pad(tp.ntiles);
    HIP_CHECK(hipMemcpyAsync(hc.data(), tp.hcount, tp.ntiles * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(pad.data(), padded, tp.ntiles * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint64_t> off(tp.ntiles);
    uint64_t tot = 0;
    for (int t = 0; t < tp.ntiles; ++t) {
        if (hc[t] < 0) ++tp.fallback_tiles;
        else tp.maxH = std::max(tp.maxH, std::min(hc[t], halo_cap<TILE>()));
        if (hc[t] > halo_cap<TILE>()) ++tp.overflow_tiles;
        off[t] = tot;  // padded sizes are multiples of 64 slots, so every slice starts 128-byte aligned
        tot += pad[t];
    }
    tp.scol_elems = (int64_t)tot;
    tp.scol = h->scratch.get<uint16_t>(27, std::max<uint64_t>(tot, 128));  // the pass reads its first 64 words
    HIP_CHECK(hipMemcpyAsync(tp.toff, off.data(), tp.ntiles * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    k_tile_sell<TILE><<<tp.ntiles, PBLK, 0, s>>>(h->rowptr.p, tp.col16, n, tp.perm, tp.width, tp.toff, tp.hcount,
                                                 tp.scol);
    LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));  // the host offsets above must outlive the copy
    return tp;
}

template <int TILE, bool REG>
void launch_tile(egt_ctx* h, const TilePlan& tp, const Pt3* in, Pt3* out, float scale) {
    const size_t lds_bytes = (size_t)(tp.tile + tp.maxH + 1) * sizeof(float4);  // + the halo's spill entry
    const dim3 tgrid(ceil_div(tp.ntiles, NXCD * XCHUNK) * NXCD * XCHUNK), tblk(tile_threads<TILE>());
    k_taubin_tile<TILE, REG><<<tgrid, tblk, lds_bytes, h->stream>>>(in, out, h->rowptr.p, h->col.p, tp.scol, tp.toff,
                                                                    tp.width, tp.perm, tp.halo, tp.hcount, h->n, scale,
                                                                    tp.ntiles);
}

void launch_pass(egt_ctx* h, const TilePlan& tp, const Pt3* in, Pt3* out, float scale) {
    const bool reg = h->taubin_regularized != 0;
    if (tp.tile == 2048)
        reg ? launch_tile<2048, true>(h, tp, in, out, scale) : launch_tile<2048, false>(h, tp, in, out, scale);
    else
        reg ? launch_tile<1024, true>(h, tp, in, out, scale) : launch_tile<1024, false>(h, tp, in, out, scale);
}

TilePlan prepare(egt_ctx* h, int tile) { return tile == 2048 ? prepare_tiles<2048>(h) : prepare_tiles<1024>(h); }

// Tile-width autotuner: prepares each candidate width and times three lambda/mu pass pairs on copies of
// the points (the smoothing itself is untouched), keeping the faster width for this handle.
int tune_tile(egt_ctx* h, const Pt3* A, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Pt3* C = h->scratch.get<Pt3>(22, n);
    Pt3* D = h->scratch.get<Pt3>(23, n);
    Event e0, e1;
    int best = 2048;
    float best_ms = 0.0f;
    for (int tile : {1024, 2048}) {
        const TilePlan tp = prepare(h, tile);
        HIP_CHECK(hipMemcpyAsync(C, A, n * sizeof(Pt3), hipMemcpyDeviceToDevice, s));
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
    return best;
}

}  // namespace

void run_taubin(egt_ctx* h, int iters, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Pt3* A = h->scratch.get<Pt3>(12, n);
    Pt3* Bf = h->scratch.get<Pt3>(13, n);
    if (n > 0) {
        k_pack3<<<ceil_div(n, TBLK), TBLK, 0, s>>>(h->mpts.p, n, A);
        LAUNCH_CHECK();
    }
    if (h->taubin_tile == 0 && iters > 0) h->taubin_tile = tune_tile(h, A, lambda, mu);
    const int tile = h->taubin_tile == 1024 ? 1024 : 2048;
    Event e0, e1, ep;
    HIP_CHECK(hipEventRecord(e0, s));
    const TilePlan tp = prepare(h, tile);
    h->taubin_fallback_tiles = tp.fallback_tiles;
    h->stats.taubin_tile = tile;
    h->stats.taubin_fallback_tiles = (int32_t)tp.fallback_tiles;
    h->stats.taubin_overflow_tiles = (int32_t)tp.overflow_tiles;
    auto pass = [&](const Pt3* in, Pt3* out, float scale) { launch_pass(h, tp, in, out, scale); };
    if (iters > 0) {
        if (h->taubin_graph) {
            // A graph of up to GRAPH_ITERS iterations (2 GRAPH_ITERS passes; each iteration ends back in A) is
            // built once and replayed; the remaining iterations form a second, shorter graph. Building one
            // node per pass for hundreds of iterations costs more host time than the replays save.
            constexpr int GRAPH_ITERS = 25;  // set-up 2.9 ms vs 3.75 ms for one 600-node graph (4M torus)
            auto build = [&](int its, GraphExec& out) {
                Capture cap(s);
                for (int it = 0; it < its; ++it) {
                    pass(A, Bf, lambda);
                    pass(Bf, A, mu);
                }
                cap.end(out);
            };
            const int full = iters / GRAPH_ITERS, rest = iters % GRAPH_ITERS;
            GraphExec G0, G1;
            if (full > 0) build(GRAPH_ITERS, G0);
            if (rest > 0) build(rest, G1);
            HIP_CHECK(hipEventRecord(ep, s));  // the passes alone: from here to e1
            for (int r = 0; r < full; ++r) HIP_CHECK(hipGraphLaunch(G0.ge, s));
            if (rest > 0) HIP_CHECK(hipGraphLaunch(G1.ge, s));
            HIP_CHECK(hipStreamSynchronize(s));
        } else {
            HIP_CHECK(hipEventRecord(ep, s));
            for (int it = 0; it < iters; ++it) {
                pass(A, Bf, lambda);
                pass(Bf, A, mu);
            }
            LAUNCH_CHECK();
        }
    } else {
        HIP_CHECK(hipEventRecord(ep, s));
    }
    HIP_CHECK(hipEventRecord(e1, s));
    k_writeback<<<ceil_div(n, TBLK), TBLK, 0, s>>>(A, n, h->mpts.p, h->x.p, h->y.p, h->z.p);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.taubin_ms, e0, e1));
    float passes_ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&passes_ms, ep, e1));
    h->stats.taubin_pass_ms = iters > 0 ? passes_ms / (2.0f * iters) : 0.0f;
    h->stats.taubin_setup_ms = h->stats.taubin_ms - passes_ms;
    h->search_stale = true;
}

}  // namespace egt
