// Mesh one-ring adjacency in sorted space.
//
// Reference: GridMeshCuda::GetNeighborsFromTriangles GridMeshCuda1.cu:416-511, kernel getMeshNeihbors
// :99-138, reverseTriangleIndexes :224-232.
//
// The reference runs one launch per "round" until every triangle corner has been inserted, using a
// per-vertex atomicExch flag to serialize insertions. Here the 6 directed corner pairs of each valid
// triangle are bucketed per vertex with a counting pass + prefix sum, and one thread per vertex selects its
// k nearest distinct corners with the same register top-K as the kNN kernel (ties by sorted index).
// Triangles with an index >= n are skipped.
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int MBLK = 256;
constexpr uint64_t KEY_EMPTY = ~0ull;

__global__ void __launch_bounds__(MBLK) k_invert(const int32_t* __restrict__ perm, int64_t n,
                                                  int32_t* __restrict__ inv) {
    const int64_t s = (int64_t)blockIdx.x * MBLK + threadIdx.x;
    if (s < n) inv[perm[s]] = (int32_t)s;
}

__global__ void __launch_bounds__(MBLK) k_tri_count(const uint32_t* __restrict__ tri, int64_t ntri, int64_t n,
                                                     const int32_t* __restrict__ inv, uint32_t* __restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * MBLK + threadIdx.x;
    if (t >= ntri) return;
    const uint32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
    if (v0 >= (uint64_t)n || v1 >= (uint64_t)n || v2 >= (uint64_t)n) return;
    atomicAdd(&cnt[inv[v0]], 2u);
    atomicAdd(&cnt[inv[v1]], 2u);
    atomicAdd(&cnt[inv[v2]], 2u);
}

__global__ void __launch_bounds__(MBLK) k_tri_fill(const uint32_t* __restrict__ tri, int64_t ntri, int64_t n,
                                                    const int32_t* __restrict__ inv, uint32_t* __restrict__ cur,
                                                    int32_t* __restrict__ adj) {
    const int64_t t = (int64_t)blockIdx.x * MBLK + threadIdx.x;
    if (t >= ntri) return;
    const uint32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
    if (v0 >= (uint64_t)n || v1 >= (uint64_t)n || v2 >= (uint64_t)n) return;
    const int32_t s[3] = {inv[v0], inv[v1], inv[v2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const uint32_t p = atomicAdd(&cur[s[c]], 2u);
        adj[p] = s[(c + 1) % 3];
        adj[p + 1] = s[(c + 2) % 3];
    }
}

template <int K>
__global__ void __launch_bounds__(MBLK) k_mesh_select(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const uint32_t* __restrict__ off,
                                                       const int32_t* __restrict__ adj, int64_t n, int k,
                                                       uint32_t total, int32_t* __restrict__ knn) {
    const int64_t i = (int64_t)blockIdx.x * MBLK + threadIdx.x;
    if (i >= n) return;
    uint64_t L[K];
#pragma unroll
    for (int t = 0; t < K; ++t) L[t] = (t < K - k) ? 0ull : KEY_EMPTY;
    const float px = x[i], py = y[i], pz = z[i];
    const uint32_t e0 = off[i], e1 = i + 1 < n ? off[i + 1] : total;
    for (uint32_t p = e0; p < e1; ++p) {
        const int32_t j = adj[p];
        const float d = sqd(px, py, pz, x[j], y[j], z[j]);
        const uint64_t key = ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)j;
        if (!(key < L[K - 1])) continue;
        bool dup = false;
#pragma unroll
        for (int t = 0; t < K; ++t) dup = dup || (L[t] == key);
        if (dup) continue;
#pragma unroll
        for (int t = K - 1; t > 0; --t) {
            const bool lt_prev = key < L[t - 1];
            const bool lt_cur = key < L[t];
            L[t] = lt_prev ? L[t - 1] : (lt_cur ? key : L[t]);
        }
        L[0] = key < L[0] ? key : L[0];
    }
    int32_t* row = knn + i * k;
#pragma unroll
    for (int t = 0; t < K; ++t)
        if (t >= K - k) row[t - (K - k)] = L[t] != KEY_EMPTY ? (int32_t)(uint32_t)L[t] : -1;
}

}  // namespace

void run_mesh_adjacency(egt_ctx* h, int k) {
    const int64_t n = h->n, ntri = h->ntri;
    hipStream_t s = h->stream;
    int32_t* inv = h->scratch.get<int32_t>(8, n);
    uint32_t* cnt = h->scratch.get<uint32_t>(9, n);
    uint32_t* off = h->scratch.get<uint32_t>(10, n);
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    k_invert<<<ceil_div(n, MBLK), MBLK, 0, s>>>(h->perm.p, n, inv);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemsetAsync(cnt, 0, n * sizeof(uint32_t), s));
    if (ntri > 0) {
        k_tri_count<<<ceil_div(ntri, MBLK), MBLK, 0, s>>>(h->tri.p, ntri, n, inv, cnt);
        LAUNCH_CHECK();
    }
    uint32_t total = 0;
    exclusive_scan_u32(cnt, off, n, s, h->scratch, &total);
    int32_t* adj = h->scratch.get<int32_t>(11, std::max<uint32_t>(total, 1));
    HIP_CHECK(hipMemcpyAsync(cnt, off, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    if (ntri > 0) {
        k_tri_fill<<<ceil_div(ntri, MBLK), MBLK, 0, s>>>(h->tri.p, ntri, n, inv, cnt, adj);
        LAUNCH_CHECK();
    }
    h->knn.alloc((size_t)n * k);
    const dim3 grid(ceil_div(n, MBLK)), blk(MBLK);
    const int K = k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : k <= 40 ? 40 : 64;
    switch (K) {
        case 8: hipLaunchKernelGGL(k_mesh_select<8>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, off, adj, n, k, total, h->knn.p); break;
        case 16: hipLaunchKernelGGL(k_mesh_select<16>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, off, adj, n, k, total, h->knn.p); break;
        case 32: hipLaunchKernelGGL(k_mesh_select<32>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, off, adj, n, k, total, h->knn.p); break;
        case 40: hipLaunchKernelGGL(k_mesh_select<40>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, off, adj, n, k, total, h->knn.p); break;
        default: hipLaunchKernelGGL(k_mesh_select<64>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, off, adj, n, k, total, h->knn.p); break;
    }
    LAUNCH_CHECK();
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.knn_ms, e0, e1));
    h->k = k;
    h->have_knn = true;
    h->stats.invalid_rows = 0;
}

}  // namespace egt
