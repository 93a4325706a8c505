// Device primitives for gfx950: exclusive prefix sum over uint32 and a stable LSD radix sort of
// (uint32 key, uint32 value) pairs. Wave64 throughout; blocks of 256 threads (4 waves).
#include <algorithm>
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;  // 2048

constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_BLOCK * SORT_ITEMS;  // 4096
constexpr int WAVE = 64;

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across a 256-thread block; returns the block total in *total.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* lds4, uint32_t* total) {
    const int lane = __lane_id(), wid = threadIdx.x / WAVE;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == WAVE - 1) lds4[wid] = incl;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / WAVE; ++w) {
        const uint32_t s = lds4[w];
        if (w < wid) woff += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return incl - v + woff;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                             uint32_t* __restrict__ partial) {
    __shared__ uint32_t lds4[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    (void)block_excl_scan(s, lds4, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Per-block scan of SCAN_TILE items (thread t owns items [t*ITEMS, t*ITEMS+ITEMS) of the tile), adding the
// scanned block offset. total (if non-null) receives the grand total from the last block.
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                            int64_t n, const uint32_t* __restrict__ offsets,
                                                            uint32_t* __restrict__ total) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t tile[SCAN_TILE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SCAN_BLOCK + threadIdx.x;
        tile[r * SCAN_BLOCK + threadIdx.x] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        v[r] = tile[threadIdx.x * SCAN_ITEMS + r];
        s += v[r];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(s, lds4, &tot) + (offsets ? offsets[blockIdx.x] : 0u);
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        tile[threadIdx.x * SCAN_ITEMS + r] = run;
        run += v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SCAN_BLOCK + threadIdx.x;
        if (i < n) out[i] = tile[r * SCAN_BLOCK + threadIdx.x];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) *total = run;
}

// Exclusive scan through per-level partial buffers: part(depth, count) returns room for count values at that
// recursion depth (allocated from a handle's scratch, or carved out of a fixed workspace).
template <class PartFn>
void scan_rec(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, PartFn&& part, int depth,
              uint32_t* dtotal) {
    const int nb = ceil_div(n, SCAN_TILE);
    if (nb == 1) {
        k_scan_apply<<<1, SCAN_BLOCK, 0, s>>>(in, out, n, nullptr, dtotal);
        LAUNCH_CHECK();
        return;
    }
    if (depth >= 3) throw ArgError{EGT_ERR_UNSUPPORTED, "scan too large"};
    uint32_t* P = part(depth, 2 * (size_t)nb);
    k_scan_reduce<<<nb, SCAN_BLOCK, 0, s>>>(in, n, P);
    LAUNCH_CHECK();
    scan_rec(P, P + nb, nb, s, part, depth + 1, nullptr);
    k_scan_apply<<<nb, SCAN_BLOCK, 0, s>>>(in, out, n, P + nb, dtotal);
    LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------------------------------
// LSD radix sort, BITS-bit digits (8: measured faster than three passes of 10 bits for 30-bit keys, 2.52 against
// 2.63 ms for the whole grid build of a 4M-point cloud and 46.8 against 48.4 ms for 55M points in 512 clouds:
// with 1024 digit values a 4096-key tile scatters runs of about 4 keys, which costs more than the pass saved).
// Pass structure: histogram per tile (digit-major count table), exclusive scan of the table, stable scatter.
// Stability inside a tile: wave w owns the contiguous sub-tile [w*1024, (w+1)*1024) and walks it in 16 rounds
// of 64 consecutive keys; lanes with equal digits are found with BITS ballots, ranked with mbcnt, and a
// per-wave per-digit LDS counter advanced by the lowest such lane.
constexpr int RADIX_BITS = 8;

template <int BITS>
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                           uint32_t* __restrict__ counts, int nblocks) {
    constexpr int BINS = 1 << BITS;
    constexpr uint32_t MASK = BINS - 1;
    __shared__ uint32_t hist[BINS];
    for (int d = threadIdx.x; d < BINS; d += SORT_BLOCK) hist[d] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE;
#pragma unroll 4
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SORT_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & MASK], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < BINS; d += SORT_BLOCK) counts[(int64_t)d * nblocks + blockIdx.x] = hist[d];
}

template <int BITS>
__global__ void __launch_bounds__(SORT_BLOCK) k_sort_scatter(const uint32_t* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int64_t n, int shift,
                                                              const uint32_t* __restrict__ offsets, int nblocks) {
    constexpr int BINS = 1 << BITS;
    constexpr uint32_t MASK = BINS - 1;
    __shared__ uint32_t wcnt[SORT_BLOCK / WAVE][BINS];
    __shared__ uint32_t goff[BINS];
    const int lane = __lane_id(), wid = threadIdx.x / WAVE;
    for (int d = threadIdx.x; d < BINS; d += SORT_BLOCK) {
        for (int w = 0; w < SORT_BLOCK / WAVE; ++w) wcnt[w][d] = 0;
        goff[d] = offsets[(int64_t)d * nblocks + blockIdx.x];
    }
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * SORT_TILE + (int64_t)wid * (WAVE * SORT_ITEMS);
    uint32_t kk[SORT_ITEMS], vv[SORT_ITEMS], pos[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * WAVE + lane;
        const bool valid = i < n;
        kk[r] = valid ? kin[i] : 0u;
        vv[r] = valid ? vin[i] : 0u;
        const uint32_t digit = (kk[r] >> shift) & MASK;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint64_t bal = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
        const int leader = __ffsll((unsigned long long)peers) - 1;
        uint32_t old = 0;
        if (valid && lane == leader) old = atomicAdd(&wcnt[wid][digit], (uint32_t)__popcll(peers));
        old = __shfl(old, leader < 0 ? 0 : leader, WAVE);
        pos[r] = old + rank;
    }
    __syncthreads();
    for (int d = threadIdx.x; d < BINS; d += SORT_BLOCK) {  // per digit: exclusive prefix over the waves
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < SORT_BLOCK / WAVE; ++w) {
            const uint32_t c = wcnt[w][d];
            wcnt[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * WAVE + lane;
        if (i < n) {
            const uint32_t digit = (kk[r] >> shift) & MASK;
            const uint32_t dst = goff[digit] + wcnt[wid][digit] + pos[r];
            kout[dst] = kk[r];
            vout[dst] = vv[r];
        }
    }
}

// One stable LSD pass structure over explicit buffers (ak, av: n keys and values; cnt, off: one count per digit
// value and sort tile).
template <class PartFn>
void radix_passes(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s, uint32_t* ak,
                  uint32_t* av, uint32_t* cnt, uint32_t* off, PartFn&& part) {
    const int nb = ceil_div(n, SORT_TILE);
    uint32_t *ki = keys, *vi = vals, *ko = ak, *vo = av;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += RADIX_BITS, ++passes) {
        k_sort_hist<RADIX_BITS><<<nb, SORT_BLOCK, 0, s>>>(ki, n, shift, cnt, nb);
        LAUNCH_CHECK();
        scan_rec(cnt, off, ((int64_t)1 << RADIX_BITS) * nb, s, part, 0, nullptr);
        k_sort_scatter<RADIX_BITS><<<nb, SORT_BLOCK, 0, s>>>(ki, vi, ko, vo, n, shift, off, nb);
        LAUNCH_CHECK();
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    if (passes & 1) {
        HIP_CHECK(hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_sum_u64(const uint32_t* __restrict__ in, int64_t n,
                                                         unsigned long long* __restrict__ total) {
    uint64_t s = 0;
    for (int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SCAN_BLOCK)
        s += in[i];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, WAVE);
    if (__lane_id() == 0 && s) atomicAdd(total, (unsigned long long)s);
}

}  // namespace

static DBuf<uint32_t>& part_buf(Scratch& sc, int i);

uint64_t sum_u32(const uint32_t* in, int64_t n, hipStream_t s, Scratch& sc) {
    unsigned long long* t = reinterpret_cast<unsigned long long*>(sc.get<uint64_t>(36, 1));
    HIP_CHECK(hipMemsetAsync(t, 0, sizeof(uint64_t), s));
    if (n > 0) {
        k_sum_u64<<<(int)std::min<int64_t>(2048, ceil_div(n, SCAN_BLOCK)), SCAN_BLOCK, 0, s>>>(in, n, t);
        LAUNCH_CHECK();
    }
    uint64_t h = 0;
    HIP_CHECK(hipMemcpyAsync(&h, t, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return h;
}

namespace {
// Scan partials from the handle's scratch (buffers 0-2, one per recursion depth).
auto scratch_parts(Scratch& sc) {
    return [&sc](int depth, size_t count) {
        DBuf<uint32_t>& P = part_buf(sc, depth);
        P.alloc(count);
        return P.p;
    };
}
}  // namespace

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, Scratch& sc,
                        uint32_t* total_host) {
    if (n <= 0) {
        if (total_host) *total_host = 0;
        return;
    }
    DBuf<uint32_t>& tot = part_buf(sc, 3);
    tot.alloc(1);
    scan_rec(in, out, n, s, scratch_parts(sc), 0, tot.p);
    if (total_host) {
        HIP_CHECK(hipMemcpyAsync(total_host, tot.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}

void exclusive_scan_u32_dev(const uint32_t* in, uint32_t* out, int64_t n, hipStream_t s, Scratch& sc,
                            uint32_t* total_dev) {
    if (n <= 0) {
        HIP_CHECK(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), s));
        return;
    }
    scan_rec(in, out, n, s, scratch_parts(sc), 0, total_dev);
}

void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s, Scratch& sc) {
    if (n <= 1) return;
    const int nb = ceil_div(n, SORT_TILE);
    DBuf<uint32_t>& ak = part_buf(sc, 4);
    DBuf<uint32_t>& av = part_buf(sc, 5);
    DBuf<uint32_t>& cnt = part_buf(sc, 6);
    DBuf<uint32_t>& off = part_buf(sc, 7);
    const size_t bins = (size_t)1 << RADIX_BITS;
    ak.alloc(n);
    av.alloc(n);
    cnt.alloc(bins * nb);
    off.alloc(bins * nb);
    radix_passes(keys, vals, n, key_bits, s, ak.p, av.p, cnt.p, off.p, scratch_parts(sc));
}

size_t radix_sort_workspace_words(int64_t n) {
    const int64_t nb = (n + SORT_TILE - 1) / SORT_TILE;
    const int64_t c = ((int64_t)1 << RADIX_BITS) * nb;  // count and offset tables
    const int64_t p0 = (c + SCAN_TILE - 1) / SCAN_TILE, p1 = (p0 + SCAN_TILE - 1) / SCAN_TILE;
    const int64_t p2 = (p1 + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(2 * n + 2 * c + 2 * (p0 + p1 + p2));
}

void radix_sort_pairs_ws(uint32_t* keys, uint32_t* vals, int64_t n, int key_bits, hipStream_t s, uint32_t* ws) {
    if (n <= 1) return;
    const int64_t nb = (n + SORT_TILE - 1) / SORT_TILE, c = ((int64_t)1 << RADIX_BITS) * nb;
    const int64_t p0 = (c + SCAN_TILE - 1) / SCAN_TILE, p1 = (p0 + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t* ak = ws;
    uint32_t* av = ak + n;
    uint32_t* cnt = av + n;
    uint32_t* off = cnt + c;
    uint32_t* parts = off + c;  // depth 0: 2 p0 words, depth 1: 2 p1, depth 2: the rest
    auto part = [parts, p0, p1](int depth, size_t) {
        return depth == 0 ? parts : (depth == 1 ? parts + 2 * p0 : parts + 2 * (p0 + p1));
    };
    radix_passes(keys, vals, n, key_bits, s, ak, av, cnt, off, part);
}

static DBuf<uint32_t>& part_buf(Scratch& sc, int i) { return sc.bufs[i]; }

}  // namespace egt
