// PCA normals and their orientation over a minimum spanning forest of the neighbour graph.
//
// Reference: Normal_CUDA::GetNormals Normal_CUDA.cu:304-775 with calculateNormal :87-144, getEdgeNumber /
// fillEdges :61-85, :146-175, mst() CUDA_MST.cu:423-460 (Boruvka), BFS bfs.cu:63-111 (sign propagation).
//
// 1. PCA: one thread per point; centroid of the point and its neighbours, covariance of the neighbours,
//    eigenvector of the smallest eigenvalue by cyclic Jacobi (egt_device.h).
// 2. Edges: every undirected edge {u < v} of the rows once, in (u, v) order: row u owns it when it lists v,
//    otherwise row v does; per-u buckets come from a counting pass + prefix sum and are sorted by v.
//    Weight max(0, 1 - |n_u . n_v|); edge key = (weight bits << 32) | edge id, a total order equal to
//    (weight, u, v).
// 3. Boruvka with parity: each component takes its minimum-key edge; mutual pairs keep the smaller id as
//    root; hooks carry the parity bit par(u) ^ par(v) ^ (n_u . n_v < 0); pointer jumping on packed
//    (target << 1 | parity) words resolves roots and parities. The forest is the unique minimum spanning
//    forest under the key order.
// 4. Orientation: each tree's root is u of its lightest edge; a vertex flips when its parity differs from
//    the root's, which equals the reference's BFS rule "flip the child when n_parent . n_child < 0".
#include <algorithm>
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int NBLK = 256;
constexpr uint64_t KEY_EMPTY = ~0ull;

// One thread per Morton position; neighbours from the Morton-order CSR (slot order of the rows), points
// from the Morton float4 copy. Writes the normal of sorted point s = mpts[m].w. The neighbour loops fetch four
// neighbours at a time (indices clamped to the row, so every load is valid) and add them in slot order.
__global__ void __launch_bounds__(NBLK) k_pca(const float4* __restrict__ mpts, const int32_t* __restrict__ rowptr,
                                               const int32_t* __restrict__ col, int64_t n, float* __restrict__ nx,
                                               float* __restrict__ ny, float* __restrict__ nz) {
    const int64_t m = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (m >= n) return;
    const float4 p = mpts[m];
    const int32_t i = __float_as_int(p.w);
    const int32_t e0 = rowptr[m], e1 = rowptr[m + 1];
    float sx = p.x, sy = p.y, sz = p.z;
    int counter = 1;
    for (int32_t e = e0; e < e1; e += 4) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = mpts[col[min(e + u, e1 - 1)]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (e + u < e1) {
                sx += q[u].x;
                sy += q[u].y;
                sz += q[u].z;
                ++counter;
            }
        }
    }
    const float cx = sx / (float)counter, cy = sy / (float)counter, cz = sz / (float)counter;
    float a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0;
    for (int32_t e = e0; e < e1; e += 4) {
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = mpts[col[min(e + u, e1 - 1)]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (e + u < e1) {
                const float dx = q[u].x - cx, dy = q[u].y - cy, dz = q[u].z - cz;
                a00 = __builtin_fmaf(dx, dx, a00);
                a01 = __builtin_fmaf(dx, dy, a01);
                a02 = __builtin_fmaf(dx, dz, a02);
                a11 = __builtin_fmaf(dy, dy, a11);
                a12 = __builtin_fmaf(dy, dz, a12);
                a22 = __builtin_fmaf(dz, dz, a22);
            }
        }
    }
    float vx, vy, vz;
    eig3_smallest(a00, a01, a02, a11, a12, a22, vx, vy, vz);
    const float nrm = __builtin_sqrtf(__builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx)));
    nx[i] = vx / nrm;
    ny[i] = vy / nrm;
    nz[i] = vz / nrm;
}

// pass 0: count per owner u and record which CSR entries emit an edge; pass 1: write v into u's bucket.
// Walks the Morton-order CSR (the same rows as the sorted-space neighbour lists, holes removed). Entry (i, nb)
// emits the edge {i, nb} when i < nb, or when i > nb and row nb does not list i (pass 0 scans nb's compact
// CSR row for that and stores the answer in own[], which pass 1 reads). Edges with u = i go to the row's own
// bucket: one atomic per row reserves them (order within a bucket is free: k_edge_sort sorts it; pass 0 keeps
// their count per row in lcount[]); only the edges owned by another row (u = nb) take an atomic each.
template <int PASS>
__global__ void __launch_bounds__(NBLK) k_edges(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                                 const int32_t* __restrict__ morder, int64_t n,
                                                 uint32_t* __restrict__ cnt, uint8_t* __restrict__ own,
                                                 uint8_t* __restrict__ lcount, int32_t* __restrict__ ev) {
    const int64_t m = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (m >= n) return;
    const int32_t i = morder[m];
    const int32_t e0 = rowptr[m], e1 = rowptr[m + 1];
    uint32_t local = 0;  // edges of this row with u = i
    // the row is walked four entries per step, their loads issued together (indices clamped to the row)
    if (PASS == 0) {
        for (int32_t e = e0; e < e1; e += 4) {
            int32_t m2v[4], nbv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) m2v[u] = col[min(e + u, e1 - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) nbv[u] = morder[m2v[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (e + u >= e1) continue;
                const int32_t m2 = m2v[u], nb = nbv[u];
                bool emit = nb != i;
                if (emit && i > nb) {
                    bool listed = false;  // does row nb list i? then row nb owns the edge
                    const int32_t f1 = rowptr[m2 + 1];
                    for (int32_t f = rowptr[m2]; f < f1; f += 4) {
                        int32_t c4[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) c4[v] = col[min(f + v, f1 - 1)];
#pragma unroll
                        for (int v = 0; v < 4; ++v) listed = listed || c4[v] == (int32_t)m;
                    }
                    emit = !listed;
                    if (emit) atomicAdd(&cnt[nb], 1u);
                } else if (emit) {
                    ++local;
                }
                own[e + u] = emit ? 1 : 0;
            }
        }
        lcount[m] = (uint8_t)local;  // a row has at most 64 entries
        if (local) atomicAdd(&cnt[i], local);
    } else {
        local = lcount[m];
        uint32_t p = local ? atomicAdd(&cnt[i], local) : 0u;
        for (int32_t e = e0; e < e1; e += 4) {
            uint8_t ow[4];
            int32_t nbv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) ow[u] = own[min(e + u, e1 - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) nbv[u] = morder[col[min(e + u, e1 - 1)]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (e + u >= e1 || !ow[u]) continue;
                const int32_t nb = nbv[u];
                if (nb > i) ev[p++] = nb;
                else ev[atomicAdd(&cnt[nb], 1u)] = i;
            }
        }
    }
}

__global__ void __launch_bounds__(NBLK) k_edge_sort(const uint32_t* __restrict__ off, int64_t n, uint32_t m,
                                                     int32_t* __restrict__ ev, int32_t* __restrict__ eu) {
    const int64_t u = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (u >= n) return;
    const uint32_t b = off[u], e = u + 1 < n ? off[u + 1] : m;
    if (e - b <= 16) {  // the common case: sorted in registers by a 16-wide bitonic network, padded with INT_MAX
        int32_t r[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) r[i] = b + i < e ? ev[b + i] : INT32_MAX;
#pragma unroll
        for (int k = 2; k <= 16; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int l = i ^ j;
                    if (l > i) {
                        const int32_t x = r[i], y = r[l];
                        const bool up = (i & k) == 0;
                        r[i] = up ? min(x, y) : max(x, y);
                        r[l] = up ? max(x, y) : min(x, y);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (b + i < e) {
                ev[b + i] = r[i];
                eu[b + i] = (int32_t)u;
            }
        return;
    }
    for (uint32_t a = b + 1; a < e; ++a) {  // insertion sort of the bucket by v
        const int32_t t = ev[a];
        uint32_t c = a;
        while (c > b && ev[c - 1] > t) {
            ev[c] = ev[c - 1];
            --c;
        }
        ev[c] = t;
    }
    for (uint32_t a = b; a < e; ++a) eu[a] = (int32_t)u;
}

__global__ void __launch_bounds__(NBLK) k_edge_attr(const int32_t* __restrict__ eu, const int32_t* __restrict__ ev,
                                                     uint32_t m, const float* __restrict__ nx,
                                                     const float* __restrict__ ny, const float* __restrict__ nz,
                                                     uint64_t* __restrict__ ekey, uint8_t* __restrict__ erel,
                                                     float* __restrict__ ew) {
    const uint32_t e = blockIdx.x * NBLK + threadIdx.x;
    if (e >= m) return;
    const int32_t u = eu[e], v = ev[e];
    const float dot = __builtin_fmaf(nz[u], nz[v], __builtin_fmaf(ny[u], ny[v], nx[u] * nx[v]));
    const float w = fmaxf(0.0f, 1.0f - __builtin_fabsf(dot));
    ew[e] = w;
    ekey[e] = ((uint64_t)__float_as_uint(w) << 32) | e;
    erel[e] = dot < 0.0f ? 1 : 0;
}

__global__ void __launch_bounds__(NBLK) k_bv_init(int64_t n, int32_t* __restrict__ comp, uint8_t* __restrict__ par) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (v >= n) return;
    comp[v] = (int32_t)v;
    par[v] = 0;
}

__global__ void __launch_bounds__(NBLK) k_bv_reset(int64_t n, uint64_t* __restrict__ best, uint32_t* __restrict__ H) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (v >= n) return;
    best[v] = KEY_EMPTY;
    H[v] = (uint32_t)v << 1;
}

// atomicMin(base[idx], val) for the active lanes of a wave. Edges come in edge-id order, so lanes that share
// an endpoint's component often form runs: when a wave has any, a segmented inclusive min-scan over the lanes
// (log2(64) steps, a step merging the lane d below when it has the same target) leaves each run's minimum in
// its last lane, the only one of the run still to issue. Lanes that share a target without being adjacent may
// merge too, which is harmless: their values go to the same address. Late in Boruvka most edges join a few
// large components, so the remaining lanes often share a few targets without being adjacent: the target of
// the first remaining lane is then reduced over the wave and issued once, for up to 4 targets, as long as
// each is shared by at least two lanes; the rest issue their own atomics. Call from all lanes.
__device__ inline void wave_atomic_min(uint64_t* base, int32_t idx, uint64_t val, bool active) {
    const int lane = __lane_id();
    const int32_t k = active ? idx : -1 - lane;  // inactive lanes: distinct keys that match no target
    unsigned long long v = active ? (unsigned long long)val : ~0ull;
    bool issue = active;
    const int32_t kp = __shfl_up(k, 1, 64);
    if (__ballot(active && lane > 0 && kp == k) != 0) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t ok = __shfl_up(k, d, 64);
            const unsigned long long ov = __shfl_up(v, d, 64);
            if (lane >= d && ok == k && ov < v) v = ov;
        }
        const int32_t kn = __shfl_down(k, 1, 64);
        issue = active && (lane == 63 || kn != k);
    }
    for (int r = 0; r < 4; ++r) {
        const uint64_t amask = __ballot(issue);
        if (!amask) return;
        const int leader = __ffsll((unsigned long long)amask) - 1;
        const int32_t lk = __shfl(k, leader, 64);
        const bool mine = issue && k == lk;
        if (__popcll(__ballot(mine)) < 2) break;
        unsigned long long t = mine ? v : ~0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const unsigned long long o = __shfl_xor(t, d, 64);
            t = o < t ? o : t;
        }
        if (lane == leader) atomicMin((unsigned long long*)&base[lk], t);
        issue = issue && !mine;
    }
    if (issue) atomicMin((unsigned long long*)&base[idx], v);
}

// Block-level variant: when all active lanes of the block target one address, one atomic per block.
__device__ inline void block_atomic_min(uint64_t* base, int32_t idx, uint64_t val, bool active) {
    __shared__ int32_t widx[NBLK / 64];
    __shared__ unsigned long long wval[NBLK / 64];
    __shared__ int wstate[NBLK / 64];  // 0 idle, 1 uniform, 2 mixed
    const int wid = threadIdx.x / 64;
    const uint64_t amask = __ballot(active);
    int state = 0;
    int32_t lidx = 0;
    unsigned long long v = ~0ull;
    if (amask) {
        const int leader = __ffsll((unsigned long long)amask) - 1;
        lidx = __shfl(idx, leader, 64);
        state = __ballot(active && idx != lidx) == 0 ? 1 : 2;
        v = active ? (unsigned long long)val : ~0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const unsigned long long o = __shfl_xor(v, d, 64);
            v = o < v ? o : v;
        }
    }
    if (__lane_id() == 0) {
        widx[wid] = lidx;
        wval[wid] = v;
        wstate[wid] = state;
    }
    __syncthreads();
    if (state == 2 && active) atomicMin((unsigned long long*)&base[idx], (unsigned long long)val);
    if (threadIdx.x == 0) {
        bool all_same = true;
        int32_t bidx = -1;
        unsigned long long bv = ~0ull;
        for (int w = 0; w < NBLK / 64; ++w) {
            if (wstate[w] != 1) continue;
            if (bidx < 0) bidx = widx[w];
            if (widx[w] != bidx) all_same = false;
            bv = wval[w] < bv ? wval[w] : bv;
        }
        if (bidx >= 0) {
            if (all_same) {
                atomicMin((unsigned long long*)&base[bidx], bv);
            } else {
                for (int w = 0; w < NBLK / 64; ++w)
                    if (wstate[w] == 1) atomicMin((unsigned long long*)&base[widx[w]], wval[w]);
            }
        }
    }
}

// The alive-edge count is read from device memory (*mcount). The edge kernels of a round loop over it with a
// grid capped at BV_GRID blocks, so a round with few edges left costs a few microseconds, not a launch over
// the host's stale bound.
constexpr int BV_GRID = 4096;
__global__ void __launch_bounds__(NBLK) k_bv_best(const int32_t* __restrict__ eu, const int32_t* __restrict__ ev,
                                                   const uint64_t* __restrict__ ekey, const uint32_t* __restrict__ alive,
                                                   const uint32_t* __restrict__ mcount, const int32_t* __restrict__ comp,
                                                   uint64_t* __restrict__ best) {
    const uint32_t m = *mcount;
    for (uint32_t t0 = blockIdx.x * NBLK; t0 < m; t0 += gridDim.x * NBLK) {  // block-uniform bound
        const uint32_t t = t0 + threadIdx.x;
        int32_t cu = 0, cv = 0;
        uint64_t key = 0;
        bool act = t < m;
        if (act) {
            const uint32_t e = alive ? alive[t] : t;
            cu = comp[eu[e]];
            cv = comp[ev[e]];
            act = cu != cv;
            key = ekey[e];
        }
        wave_atomic_min(best, cu, key, act);
        wave_atomic_min(best, cv, key, act);
    }
}

// Edges still joining two components after this round's merge (the rest can never be chosen again). Flags
// are written for every position below the bound (0 past the device count), which the scan then covers.
__global__ void __launch_bounds__(NBLK) k_bv_alive(const int32_t* __restrict__ eu, const int32_t* __restrict__ ev,
                                                    const uint32_t* __restrict__ alive,
                                                    const uint32_t* __restrict__ mcount, uint32_t bound,
                                                    const int32_t* __restrict__ comp, uint32_t* __restrict__ flag) {
    const uint32_t m = *mcount;
    for (uint32_t t = blockIdx.x * NBLK + threadIdx.x; t < bound; t += gridDim.x * NBLK) {
        uint32_t f = 0;
        if (t < m) {
            const uint32_t e = alive ? alive[t] : t;
            f = comp[eu[e]] != comp[ev[e]] ? 1u : 0u;
        }
        flag[t] = f;
    }
}

__global__ void __launch_bounds__(NBLK) k_bv_compact(const uint32_t* __restrict__ alive,
                                                      const uint32_t* __restrict__ mcount,
                                                      const uint32_t* __restrict__ flag,
                                                      const uint32_t* __restrict__ pos, uint32_t* __restrict__ out) {
    const uint32_t m = *mcount;
    for (uint32_t t = blockIdx.x * NBLK + threadIdx.x; t < m; t += gridDim.x * NBLK)
        if (flag[t]) out[pos[t]] = alive ? alive[t] : t;
}

// Hooks every root with a lightest edge to the root at the edge's other end. No hook count is kept: a round
// hooks at least one root exactly when edges between components are alive at its start, which the alive-edge
// count already tells (a shared counter took one atomic per wave on one address, about 0.7 ms a round on a
// 4M-point cloud).
__global__ void __launch_bounds__(NBLK) k_bv_hook(int64_t n, const int32_t* __restrict__ eu,
                                                   const int32_t* __restrict__ ev, const uint8_t* __restrict__ erel,
                                                   const int32_t* __restrict__ comp, const uint8_t* __restrict__ par,
                                                   const uint64_t* __restrict__ best, uint32_t* __restrict__ H,
                                                   uint8_t* __restrict__ msf) {
    const int64_t c = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (c < n && comp[c] == c) {
        const uint64_t b = best[c];
        if (b != KEY_EMPTY) {
            const uint32_t e = (uint32_t)b;
            const int32_t u = eu[e], v = ev[e];
            const int32_t cu = comp[u], cv = comp[v];
            const int32_t other = cu == (int32_t)c ? cv : cu;
            if (!(best[other] == b && (int32_t)c < other)) {  // mutual pair: the smaller id stays root
                H[c] = ((uint32_t)other << 1) | (uint32_t)(par[u] ^ par[v] ^ erel[e]);
                msf[e] = 1;
            }
        }
    }
}

// Follows hook pointers to the root. The hook graph is a forest (edge keys are a strict total order and
// mutual pairs keep one root), so every walk ends; the walk is still capped at n steps and a cap hit is
// reported through *overflow instead of spinning.
__global__ void __launch_bounds__(NBLK) k_bv_jump(int64_t n, const int32_t* __restrict__ comp,
                                                   uint32_t* __restrict__ H, uint32_t* __restrict__ overflow) {
    const int64_t c = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (c >= n || comp[c] != c) return;
    uint32_t w = H[c];
    int64_t steps = 0;
    for (;;) {
        const uint32_t t = w >> 1;
        const uint32_t w2 = __hip_atomic_load(&H[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((w2 >> 1) == t) break;
        w = (w2 & ~1u) | ((w ^ w2) & 1u);
        if (++steps > n) {
            atomicAdd(overflow, 1u);
            break;
        }
    }
    __hip_atomic_store(&H[c], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(NBLK) k_bv_relabel(int64_t n, int32_t* __restrict__ comp, uint8_t* __restrict__ par,
                                                      const uint32_t* __restrict__ H) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (v >= n) return;
    const uint32_t w = H[comp[v]];
    comp[v] = (int32_t)(w >> 1);
    par[v] ^= (uint8_t)(w & 1u);
}

// Lightest edge of each tree = minimum over its vertices of their lightest incident edge (round-1 best).
__global__ void __launch_bounds__(NBLK) k_tree_min(int64_t n, const int32_t* __restrict__ comp,
                                                    const uint64_t* __restrict__ best1, uint64_t* __restrict__ tmin) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    const bool act = v < n && best1[v] != KEY_EMPTY;
    block_atomic_min(tmin, act ? comp[v] : 0, act ? best1[v] : 0ull, act);
}

__global__ void __launch_bounds__(NBLK) k_orient(int64_t n, const int32_t* __restrict__ comp,
                                                  const uint8_t* __restrict__ par, const uint64_t* __restrict__ tmin,
                                                  const int32_t* __restrict__ eu, float* __restrict__ nx,
                                                  float* __restrict__ ny, float* __restrict__ nz) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (v >= n) return;
    const uint64_t t = tmin[comp[v]];
    if (t == KEY_EMPTY) return;
    const int32_t root = eu[(uint32_t)t];
    if (par[v] ^ par[root]) {
        nx[v] = -nx[v];
        ny[v] = -ny[v];
        nz[v] = -nz[v];
    }
}

// deterministic forest statistics: per-block partial sums, summed on the host in block order
__global__ void __launch_bounds__(NBLK) k_msf_stats(const uint8_t* __restrict__ msf, const float* __restrict__ ew,
                                                     uint32_t m, double* __restrict__ part_w,
                                                     uint32_t* __restrict__ part_c) {
    __shared__ double sw[NBLK];
    __shared__ uint32_t scnt[NBLK];
    double w = 0.0;
    uint32_t c = 0;
    for (uint32_t e = blockIdx.x * NBLK + threadIdx.x; e < m; e += gridDim.x * NBLK)
        if (msf[e]) {
            w += (double)ew[e];
            ++c;
        }
    sw[threadIdx.x] = w;
    scnt[threadIdx.x] = c;
    __syncthreads();
    for (int s = NBLK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sw[threadIdx.x] += sw[threadIdx.x + s];
            scnt[threadIdx.x] += scnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_w[blockIdx.x] = sw[0];
        part_c[blockIdx.x] = scnt[0];
    }
}

__global__ void __launch_bounds__(NBLK) k_count_roots(int64_t n, const int32_t* __restrict__ comp,
                                                       const uint64_t* __restrict__ tmin, uint32_t* __restrict__ trees) {
    const int64_t v = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (v >= n) return;
    if (comp[v] == v && tmin[v] != KEY_EMPTY) atomicAdd(trees, 1u);
}

__global__ void __launch_bounds__(NBLK) k_scatter_normals(const float* __restrict__ nx, const float* __restrict__ ny,
                                                           const float* __restrict__ nz, const int32_t* __restrict__ perm,
                                                           int64_t n, float* __restrict__ ox, float* __restrict__ oy,
                                                           float* __restrict__ oz) {
    const int64_t s = (int64_t)blockIdx.x * NBLK + threadIdx.x;
    if (s >= n) return;
    const int32_t o = perm[s];
    ox[o] = nx[s];
    oy[o] = ny[s];
    oz[o] = nz[s];
}

template <class T>
T* dalloc(DBuf<uint8_t>& b, size_t& off, size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = reinterpret_cast<T*>(b.p + off);
    off += count * sizeof(T);
    return p;
}

}  // namespace

void run_normals(egt_ctx* h, float* nxh, float* nyh, float* nzh) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Event ev0, ev1;
    HIP_CHECK(hipEventRecord(ev0, s));
    const dim3 gn(ceil_div(n, NBLK)), blk(NBLK);

    float* nrm = h->scratch.get<float>(8, 6 * (size_t)n);
    float *nx = nrm, *ny = nrm + n, *nz = nrm + 2 * n;
    if (!h->csr_valid) build_csr_from_rows(h);
    k_pca<<<gn, blk, 0, s>>>(h->mpts.p, h->rowptr.p, h->col.p, n, nx, ny, nz);
    LAUNCH_CHECK();

    // canonical undirected edge list
    uint32_t* cnt = h->scratch.get<uint32_t>(9, n);
    uint32_t* off = h->scratch.get<uint32_t>(10, n);
    uint8_t* own = h->scratch.get<uint8_t>(11, std::max<int64_t>(h->nnz, 1) + n);
    uint8_t* lcount = own + std::max<int64_t>(h->nnz, 1);  // per Morton position: edges the row owns
    HIP_CHECK(hipMemsetAsync(cnt, 0, n * sizeof(uint32_t), s));
    k_edges<0><<<gn, blk, 0, s>>>(h->rowptr.p, h->col.p, h->morder.p, n, cnt, own, lcount, nullptr);
    LAUNCH_CHECK();
    uint32_t m = 0;
    exclusive_scan_u32(cnt, off, n, s, h->scratch, &m);
    HIP_CHECK(hipMemcpyAsync(cnt, off, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    DBuf<uint8_t> pool;
    const size_t mm = std::max<uint32_t>(m, 1);
    pool.alloc(mm * (4 + 4 + 8 + 1 + 1 + 4 + 16) + (size_t)n * (4 + 1 + 8 + 4 + 8 + 8) + 24 * 256 + 4096 * 12);
    size_t po = 0;
    int32_t* ev = dalloc<int32_t>(pool, po, mm);
    int32_t* eu = dalloc<int32_t>(pool, po, mm);
    uint64_t* ekey = dalloc<uint64_t>(pool, po, mm);
    uint8_t* erel = dalloc<uint8_t>(pool, po, mm);
    uint8_t* msf = dalloc<uint8_t>(pool, po, mm);
    float* ew = dalloc<float>(pool, po, mm);
    int32_t* comp = dalloc<int32_t>(pool, po, n);
    uint8_t* par = dalloc<uint8_t>(pool, po, n);
    uint64_t* best = dalloc<uint64_t>(pool, po, n);
    uint32_t* H = dalloc<uint32_t>(pool, po, n);
    uint64_t* tmin = dalloc<uint64_t>(pool, po, n);
    uint64_t* best1 = dalloc<uint64_t>(pool, po, n);
    uint32_t* alive_a = dalloc<uint32_t>(pool, po, mm);
    uint32_t* alive_b = dalloc<uint32_t>(pool, po, mm);
    uint32_t* aflag = dalloc<uint32_t>(pool, po, mm);
    uint32_t* apos = dalloc<uint32_t>(pool, po, mm);
    uint32_t* flags = dalloc<uint32_t>(pool, po, 4);
    uint32_t* dcount = dalloc<uint32_t>(pool, po, 2);  // alive-edge counts, current and next round
    double* part_w = dalloc<double>(pool, po, 1024);
    uint32_t* part_c = dalloc<uint32_t>(pool, po, 1024);
    if (po > pool.n) throw ArgError{EGT_ERR_STATE, "normals pool overflow"};

    k_edges<1><<<gn, blk, 0, s>>>(h->rowptr.p, h->col.p, h->morder.p, n, cnt, own, lcount, ev);
    LAUNCH_CHECK();
    k_edge_sort<<<gn, blk, 0, s>>>(off, n, m, ev, eu);
    LAUNCH_CHECK();
    const dim3 gm(std::max(1, ceil_div(m, NBLK)));
    if (m > 0) {
        k_edge_attr<<<gm, blk, 0, s>>>(eu, ev, m, nx, ny, nz, ekey, erel, ew);
        LAUNCH_CHECK();
    }
    HIP_CHECK(hipMemsetAsync(msf, 0, mm, s));

    // Boruvka with parity. The alive-edge count lives in device memory, so rounds follow one another without
    // a host round trip; every second round the host reads the count (to shrink the grids, and to stop once
    // no edge joins two components) and the hook walk's overflow flag. A round after convergence finds no
    // edge and changes nothing.
    k_bv_init<<<gn, blk, 0, s>>>(n, comp, par);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemsetAsync(flags, 0, 4 * sizeof(uint32_t), s));
    HIP_CHECK(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(dcount), m, 1, s));
    uint32_t* alive = nullptr;  // null = all edges
    uint32_t bound = m;         // host-side upper bound of the alive count
    int cur = 0;
    for (int round = 0; round < 64 && bound > 0; ++round) {
        const dim3 ga(std::min(BV_GRID, std::max(1, ceil_div(bound, NBLK))));
        k_bv_reset<<<gn, blk, 0, s>>>(n, best, H);
        k_bv_best<<<ga, blk, 0, s>>>(eu, ev, ekey, alive, dcount + cur, comp, best);
        if (round == 0) HIP_CHECK(hipMemcpyAsync(best1, best, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        k_bv_hook<<<gn, blk, 0, s>>>(n, eu, ev, erel, comp, par, best, H, msf);
        k_bv_jump<<<gn, blk, 0, s>>>(n, comp, H, flags + 1);
        k_bv_relabel<<<gn, blk, 0, s>>>(n, comp, par, H);
        k_bv_alive<<<ga, blk, 0, s>>>(eu, ev, alive, dcount + cur, bound, comp, aflag);
        LAUNCH_CHECK();
        exclusive_scan_u32_dev(aflag, apos, bound, s, h->scratch, dcount + (cur ^ 1));
        uint32_t* out = alive == alive_a ? alive_b : alive_a;
        k_bv_compact<<<ga, blk, 0, s>>>(alive, dcount + cur, aflag, apos, out);
        LAUNCH_CHECK();
        alive = out;
        cur ^= 1;
        if (round & 1) {
            uint32_t hf[2] = {0, 0};
            HIP_CHECK(hipMemcpyAsync(hf, flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(hf + 1, dcount + cur, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (hf[0]) throw ArgError{EGT_ERR_STATE, "spanning forest: hook walk exceeded n steps"};
            if (hf[1] == 0) break;  // no edge joins two components: converged
            bound = hf[1];
        }
    }
    {
        uint32_t of = 0;
        HIP_CHECK(hipMemcpyAsync(&of, flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (of) throw ArgError{EGT_ERR_STATE, "spanning forest: hook walk exceeded n steps"};
    }

    // orientation
    k_bv_reset<<<gn, blk, 0, s>>>(n, tmin, H);
    LAUNCH_CHECK();
    if (m > 0) {
        k_tree_min<<<gn, blk, 0, s>>>(n, comp, best1, tmin);
        LAUNCH_CHECK();
    }
    k_orient<<<gn, blk, 0, s>>>(n, comp, par, tmin, eu, nx, ny, nz);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventRecord(ev1, s));

    // statistics
    HIP_CHECK(hipMemsetAsync(flags, 0, 4 * sizeof(uint32_t), s));
    k_count_roots<<<gn, blk, 0, s>>>(n, comp, tmin, flags);
    const int sb = std::min(1024, std::max(1, ceil_div(m, NBLK)));
    k_msf_stats<<<sb, blk, 0, s>>>(msf, ew, m, part_w, part_c);
    LAUNCH_CHECK();
    std::vector<double> hw(sb);
    std::vector<uint32_t> hc(sb);
    uint32_t trees = 0;
    HIP_CHECK(hipMemcpyAsync(hw.data(), part_w, sb * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(hc.data(), part_c, sb * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&trees, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s));

    // original order
    float* on = nrm + 3 * n;
    k_scatter_normals<<<gn, blk, 0, s>>>(nx, ny, nz, h->perm.p, n, on, on + n, on + 2 * n);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(nxh, on, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(nyh, on + n, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(nzh, on + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipEventElapsedTime(&h->stats.normals_ms, ev0, ev1));
    double W = 0.0;
    int64_t C = 0;
    for (int b = 0; b < sb; ++b) {
        W += hw[b];
        C += hc[b];
    }
    h->stats.msf_weight = W;
    h->stats.msf_edges = C;
    h->stats.msf_trees = trees;
}

}  // namespace egt
