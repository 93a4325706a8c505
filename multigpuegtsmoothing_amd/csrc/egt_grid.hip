// Grid construction (bounding cube, normalization, cell key, stable sort into sorted space), the
// Morton-ordered search structure with per-level cell hash tables, and the output gathers.
//
// Reference: GridKNNCuda.cu:1023-1133 (InitStructure), :74-88 (cell key), :898-936 (getBuckets),
// getOutputPoints.cu:9-63, GridKNNCuda.cu:1941-1968 (GetKNNIndexes).
#include <algorithm>
#include <cmath>
#include <cstring>
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int BLK = 256;

// Per-block min/max of x, y, z (thrust reduce with minimum/maximum, GridKNNCuda.cu:1042-1059) plus a
// count of non-finite inputs.
__global__ void __launch_bounds__(BLK) k_minmax_partial(const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ z, int64_t n,
                                                         float* __restrict__ part, uint32_t* __restrict__ bad) {
    float v[6] = {1e8f, -1e8f, 1e8f, -1e8f, 1e8f, -1e8f};
    uint32_t nb = 0;
    for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
        const float a = x[i], b = y[i], c = z[i];
        nb += !(__builtin_isfinite(a) && __builtin_isfinite(b) && __builtin_isfinite(c));
        v[0] = fminf(v[0], a); v[1] = fmaxf(v[1], a);
        v[2] = fminf(v[2], b); v[3] = fmaxf(v[3], b);
        v[4] = fminf(v[4], c); v[5] = fmaxf(v[5], c);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int c = 0; c < 6; c += 2) {
            v[c] = fminf(v[c], __shfl_xor(v[c], d, 64));
            v[c + 1] = fmaxf(v[c + 1], __shfl_xor(v[c + 1], d, 64));
        }
        nb += __shfl_xor(nb, d, 64);
    }
    __shared__ float s[BLK / 64][6];
    __shared__ uint32_t sb[BLK / 64];
    const int wid = threadIdx.x / 64;
    if (__lane_id() == 0) {
        for (int c = 0; c < 6; ++c) s[wid][c] = v[c];
        sb[wid] = nb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BLK / 64; ++w)
            for (int c = 0; c < 6; c += 2) {
                s[0][c] = fminf(s[0][c], s[w][c]);
                s[0][c + 1] = fmaxf(s[0][c + 1], s[w][c + 1]);
            }
        uint32_t t = 0;
        for (int w = 0; w < BLK / 64; ++w) t += sb[w];
        for (int c = 0; c < 6; ++c) part[blockIdx.x * 6 + c] = s[0][c];
        if (t) atomicAdd(bad, t);
    }
}

// normalize_points GridKNNCuda.cu:840-849 fused with get_buckets_no_binary :74-88. With icloud (batch) each
// point uses its own cloud's cube. The normalized point and its key go out as one 16-byte record in input
// order (the gather into sorted order then reads one record per point), the key and index also as the sort's
// key and value arrays.
__global__ void __launch_bounds__(BLK) k_normalize_key(const float* __restrict__ x, const float* __restrict__ y,
                                                        const float* __restrict__ z, int64_t n, float mn, float d,
                                                        const int32_t* __restrict__ icloud,
                                                        const CloudDev* __restrict__ clouds, float a, float ba, int B,
                                                        int log2B, float4* __restrict__ tp, uint32_t* __restrict__ key,
                                                        uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    if (icloud) {
        const CloudDev& c = clouds[icloud[i]];
        mn = c.min_extent;
        d = c.extent;
    }
    const float xn = (x[i] - mn) / d, yn = (y[i] - mn) / d, zn = (z[i] - mn) / d;
    const uint32_t cx = (uint32_t)cell_of(xn, a, ba, B), cy = (uint32_t)cell_of(yn, a, ba, B),
                   cz = (uint32_t)cell_of(zn, a, ba, B);
    const uint32_t k = cx | (cy << log2B) | (cz << (2 * log2B));
    tp[i] = make_float4(xn, yn, zn, __uint_as_float(k));
    key[i] = k;
    idx[i] = (uint32_t)i;
}

// Cloud of every input point (binary search over the input offsets).
__global__ void __launch_bounds__(BLK) k_cloud_of(const int64_t* __restrict__ off, int nclouds, int64_t n,
                                                   int32_t* __restrict__ cloud) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nclouds - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    cloud[i] = lo;
}

// Order-preserving float <-> int map, so float min/max can use integer atomics.
__device__ inline int f2o(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}

// Per-cloud min/max of x, y, z (fold starting at +-1e8, as the reference's thrust reductions
// GridKNNCuda.cu:1042-1059) plus a count of non-finite inputs. A wave whose 64 points share one cloud
// reduces first and issues one atomic per value.
// Each thread folds MM_PER points (stride BLK, so loads stay coalesced) into running extremes of the cloud it
// is in, flushing them with atomics only when the cloud changes (clouds are contiguous in the input); the
// final extremes are reduced over the wave when the wave ends in one cloud. Every point of a wave used to
// hit the same six addresses of its cloud: 5.5 ms on 54.7M points against a fraction of that now.
constexpr int MM_PER = 16;
__device__ inline void minmax_flush(int* __restrict__ mm, int c, const float (&v)[6]) {
    for (int j = 0; j < 6; j += 2) {
        atomicMin(&mm[c * 6 + j], f2o(v[j]));
        atomicMax(&mm[c * 6 + j + 1], f2o(v[j + 1]));
    }
}
__global__ void __launch_bounds__(BLK) k_cloud_minmax(const float* __restrict__ x, const float* __restrict__ y,
                                                       const float* __restrict__ z, const int32_t* __restrict__ icloud,
                                                       int64_t n, int* __restrict__ mm, uint32_t* __restrict__ bad) {
    const int64_t first = (int64_t)blockIdx.x * BLK * MM_PER + threadIdx.x;
    int c = -1;
    float v[6] = {1e8f, -1e8f, 1e8f, -1e8f, 1e8f, -1e8f};
    uint32_t nb = 0;
    for (int r = 0; r < MM_PER; ++r) {
        const int64_t i = first + (int64_t)r * BLK;
        if (i >= n) break;
        const int ci = icloud[i];
        if (ci != c) {
            if (c >= 0) minmax_flush(mm, c, v);
            c = ci;
#pragma unroll
            for (int j = 0; j < 6; j += 2) {
                v[j] = 1e8f;
                v[j + 1] = -1e8f;
            }
        }
        const float p[3] = {x[i], y[i], z[i]};
        nb += !(__builtin_isfinite(p[0]) && __builtin_isfinite(p[1]) && __builtin_isfinite(p[2]));
        for (int ax = 0; ax < 3; ++ax) {
            v[2 * ax] = fminf(v[2 * ax], p[ax]);
            v[2 * ax + 1] = fmaxf(v[2 * ax + 1], p[ax]);
        }
    }
    const bool act = c >= 0;
    const int c0 = __shfl(c, 0, 64);
    // uniform with c0 < 0: a wave past the end with no point (nothing to add)
    const bool uniform = __ballot(c == c0) == ~0ull;
    if (uniform && c0 >= 0) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
            for (int j = 0; j < 6; j += 2) {
                v[j] = fminf(v[j], __shfl_xor(v[j], d, 64));
                v[j + 1] = fmaxf(v[j + 1], __shfl_xor(v[j + 1], d, 64));
            }
            nb += __shfl_xor(nb, d, 64);
        }
        if (__lane_id() == 0) {
            for (int j = 0; j < 6; j += 2) {
                atomicMin(&mm[c0 * 6 + j], f2o(v[j]));
                atomicMax(&mm[c0 * 6 + j + 1], f2o(v[j + 1]));
            }
            if (nb) atomicAdd(bad, nb);
        }
    } else if (!uniform && act) {
        minmax_flush(mm, c, v);
        if (nb) atomicAdd(bad, nb);
    }
}

__global__ void __launch_bounds__(BLK) k_minmax_init(int* __restrict__ mm, int nclouds) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < nclouds * 6) mm[i] = f2o((i & 1) ? -1e8f : 1e8f);
}

// Cloud of each element of an index list (second, stable sort pass by cloud).
__global__ void __launch_bounds__(BLK) k_gather_cloud(const int32_t* __restrict__ cloud_of,
                                                       const uint32_t* __restrict__ idx, int64_t n,
                                                       uint32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = (uint32_t)cloud_of[idx[i]];
}

// thrust::gather into sorted order, GridKNNCuda.cu:1114-1128: one 16-byte record per point.
__global__ void __launch_bounds__(BLK) k_gather_sorted(const float4* __restrict__ tp, const int32_t* __restrict__ icloud,
                                                        const uint32_t* __restrict__ sidx, int64_t n,
                                                        float* __restrict__ x, float* __restrict__ y,
                                                        float* __restrict__ z, int32_t* __restrict__ perm,
                                                        int32_t* __restrict__ key, int32_t* __restrict__ scloud) {
    const int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = sidx[s];
    const float4 v = tp[o];
    perm[s] = (int32_t)o;
    key[s] = __float_as_int(v.w);
    if (scloud) scloud[s] = icloud[o];
    x[s] = v.x;
    y[s] = v.y;
    z[s] = v.z;
}

// Sort keys of the Morton order: the fine cell's Morton code and the Morton code of the point's position
// inside the cell (sorted first, stably, so that points of a cell end up in sub-code order); also the point as
// one 16-byte record (x, y, z, key) in sorted order for the gather into Morton order.
__global__ void __launch_bounds__(BLK) k_morton_keys(const float* __restrict__ x, const float* __restrict__ y,
                                                      const float* __restrict__ z, const int32_t* __restrict__ key,
                                                      int64_t n, float a, float ba, int B, int log2B,
                                                      uint32_t* __restrict__ code, uint32_t* __restrict__ sub,
                                                      uint32_t* __restrict__ idx, float4* __restrict__ spts) {
    const int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    const uint32_t k = (uint32_t)key[s], m = (1u << log2B) - 1u;
    const uint32_t cx = k & m, cy = (k >> log2B) & m, cz = k >> (2 * log2B);
    const float px = x[s], py = y[s], pz = z[s];
    code[s] = morton_code(cx, cy, cz);
    sub[s] = vcell_sub(vcell_pack(vcoord(px, a, ba, B, (int)cx), vcoord(py, a, ba, B, (int)cy),
                                  vcoord(pz, a, ba, B, (int)cz)));
    idx[s] = (uint32_t)s;
    spts[s] = make_float4(px, py, pz, __int_as_float((int32_t)k));
}

__global__ void __launch_bounds__(BLK) k_gather_u32(const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ idx, int64_t n,
                                                     uint32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

__global__ void __launch_bounds__(BLK) k_morton_points(const float4* __restrict__ spts,
                                                        const int32_t* __restrict__ scloud,
                                                        const uint32_t* __restrict__ order, int64_t n, float a,
                                                        float ba, int B, int log2B, float4* __restrict__ mpts,
                                                        uint32_t* __restrict__ mcell, unsigned long long* __restrict__ mvcell,
                                                        uint32_t* __restrict__ mcode, int32_t* __restrict__ mcloud,
                                                        int32_t* __restrict__ morder, int32_t* __restrict__ minv) {
    const int64_t m = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (m >= n) return;
    const uint32_t s = order[m];
    morder[m] = (int32_t)s;
    minv[s] = (int32_t)m;
    const float4 v = spts[s];
    const float px = v.x, py = v.y, pz = v.z;
    mpts[m] = make_float4(px, py, pz, __int_as_float((int)s));
    const uint32_t k = (uint32_t)__float_as_int(v.w), msk = (1u << log2B) - 1u;
    const uint32_t cx = k & msk, cy = (k >> log2B) & msk, cz = k >> (2 * log2B);
    mcell[m] = cx | (cy << 10) | (cz << 20);
    mvcell[m] = vcell_pack(vcoord(px, a, ba, B, (int)cx), vcoord(py, a, ba, B, (int)cy), vcoord(pz, a, ba, B, (int)cz));
    mcode[m] = morton_code(cx, cy, cz);
    if (mcloud) mcloud[m] = scloud[s];
}

__global__ void __launch_bounds__(BLK) k_level_flags(const uint32_t* __restrict__ code,
                                                      const int32_t* __restrict__ mcloud, int64_t n, int shift,
                                                      uint32_t* __restrict__ flag) {
    const int64_t m = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (m >= n) return;
    bool first = m == 0 || (code[m] >> shift) != (code[m - 1] >> shift);
    if (mcloud && m > 0) first = first || mcloud[m] != mcloud[m - 1];
    flag[m] = first ? 1u : 0u;
}

__global__ void __launch_bounds__(BLK) k_level_starts(const uint32_t* __restrict__ flag,
                                                       const uint32_t* __restrict__ seg, int64_t n,
                                                       uint32_t* __restrict__ starts) {
    const int64_t m = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (m >= n) return;
    if (flag[m]) starts[seg[m]] = (uint32_t)m;
}

__global__ void __launch_bounds__(BLK) k_level_insert(const uint32_t* __restrict__ starts, int64_t cells,
                                                       int64_t n, const uint32_t* __restrict__ code,
                                                       const int32_t* __restrict__ mcloud, int shift,
                                                       uint4* __restrict__ tab, uint64_t offset, uint32_t log2cap,
                                                       int* __restrict__ maxc) {
    const int64_t c = (int64_t)blockIdx.x * BLK + threadIdx.x;
    int cnt = 0;
    if (c < cells) {
        const uint32_t st = starts[c];
        const uint32_t en = c + 1 < cells ? starts[c + 1] : (uint32_t)n;
        const unsigned long long cc = cell_key(mcloud ? (uint32_t)mcloud[st] : 0u, code[st] >> shift);
        const uint32_t mask = (1u << log2cap) - 1u;
        uint32_t s = hash_slot(cc, log2cap);
        // a table has at least twice as many slots as cells, so an empty slot is always found
        for (;;) {
            uint4* e = &tab[offset + s];
            const unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(e), (unsigned long long)HASH_EMPTY, cc);
            if (prev == HASH_EMPTY) {
                reinterpret_cast<uint2*>(e)[1] = make_uint2(st, en - st);
                break;
            }
            s = (s + 1u) & mask;
        }
        cnt = (int)(en - st);
    }
    // largest cell of the level: one atomic per wave (every cell used to hit the one address)
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cnt = max(cnt, __shfl_xor(cnt, d, 64));
    if (__lane_id() == 0) atomicMax(maxc, cnt);
}

// seg value at each cloud's first Morton position (running cell count), plus the total at the end.
__global__ void __launch_bounds__(BLK) k_cells_at_bases(const uint32_t* __restrict__ seg, const CloudDev* __restrict__ clouds,
                                                         int nclouds, const uint32_t* __restrict__ total,
                                                         uint32_t* __restrict__ out) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c < nclouds) out[c] = seg[clouds[c].base];
    if (c == nclouds) out[c] = *total;
}

__global__ void __launch_bounds__(BLK) k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLK) p[i] = v;
}

// produce_output_GKNN + rearrange_output, getOutputPoints.cu:9-30, into separate output buffers.
__global__ void __launch_bounds__(BLK) k_output(const float* __restrict__ x, const float* __restrict__ y,
                                                 const float* __restrict__ z, const int32_t* __restrict__ perm,
                                                 const int32_t* __restrict__ scloud,
                                                 const CloudDev* __restrict__ clouds, int64_t n, float mn, float d,
                                                 float* __restrict__ ox, float* __restrict__ oy,
                                                 float* __restrict__ oz) {
    const int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    if (scloud) {
        const CloudDev& c = clouds[scloud[s]];
        mn = c.min_extent;
        d = c.extent;
    }
    const int32_t o = perm[s];
    ox[o] = x[s] * d + mn;
    oy[o] = y[s] * d + mn;
    oz[o] = z[s] * d + mn;
}

// rearrangeReferenceNeighbors GridKNNCuda.cu:862-872 plus the host scatter :1959-1966, on the device.
__global__ void __launch_bounds__(BLK) k_knn_original(const int32_t* __restrict__ knn,
                                                       const int32_t* __restrict__ perm, int64_t n, int k,
                                                       int32_t* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= n * k) return;
    const int64_t s = t / k;
    const int j = (int)(t - s * k);
    const int32_t v = knn[t];
    out[(int64_t)perm[s] * k + j] = v >= 0 ? perm[v] : -1;
}

// Rows of one cloud of a batch, in that cloud's own original order and index space.
__global__ void __launch_bounds__(BLK) k_knn_original_cloud(const int32_t* __restrict__ knn,
                                                             const int32_t* __restrict__ perm, CloudDev c,
                                                             int64_t in_base, int32_t* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (t >= c.n * c.k) return;
    const int64_t r = t / c.k;
    const int j = (int)(t - r * c.k);
    const int32_t v = knn[c.row_off + t];
    const int64_t s = c.base + r;
    out[(perm[s] - in_base) * c.k + j] = v >= 0 ? (int32_t)(perm[c.base + v] - in_base) : -1;
}

int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

}  // namespace

namespace {

int ceil_log2(int64_t v) {
    int l = 0;
    while ((1ll << l) < v) ++l;
    return l;
}

void upload_clouds(egt_ctx* h) {
    h->dclouds.alloc(h->clouds.size());
    HIP_CHECK(hipMemcpyAsync(h->dclouds.p, h->clouds.data(), h->clouds.size() * sizeof(CloudDev),
                             hipMemcpyHostToDevice, h->stream));
}

// Bounding cube of one cloud from the per-axis min/max, GridKNNCuda.cu:1042-1077.
void cube_from_minmax(const float v[6], float& mn, float& mx) {
    const float minx = v[0] - 1e-6f, maxx = v[1] + 1e-6f, miny = v[2] - 1e-6f, maxy = v[3] + 1e-6f;
    const float minz = v[4] - 1e-6f, maxz = v[5] + 1e-6f;
    mn = std::min(std::min(minx, miny), minz);
    mx = std::max(std::max(maxx, maxy), maxz);
}

// Normalize, key, stable sort (by cloud, then key) and gather into sorted space. icloud is null for one
// cloud.
void sort_into_sorted_space(egt_ctx* h, const float* dx, const float* dy, const float* dz, const int32_t* icloud) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Scratch& sc = h->scratch;
    const int log2B = ilog2(h->B);
    const float a = -h->eps, b = 1.0f + h->eps, ba = b - a;
    float4* tp = sc.get<float4>(9, n);
    uint32_t* skey = sc.get<uint32_t>(12, n);
    uint32_t* sidx = sc.get<uint32_t>(13, n);
    k_normalize_key<<<ceil_div(n, BLK), BLK, 0, s>>>(dx, dy, dz, n, h->min_extent, h->max_extent - h->min_extent,
                                                      icloud, h->dclouds.p, a, ba, h->B, log2B, tp, skey, sidx);
    LAUNCH_CHECK();
    radix_sort_pairs(skey, sidx, n, 3 * log2B, s, sc);
    if (icloud) {
        // second stable pass by cloud: sorted by (cloud, key), ties in input order
        uint32_t* cl = sc.get<uint32_t>(22, n);
        k_gather_cloud<<<ceil_div(n, BLK), BLK, 0, s>>>(icloud, sidx, n, cl);
        LAUNCH_CHECK();
        radix_sort_pairs(cl, sidx, n, std::max(1, ceil_log2(h->nclouds)), s, sc);
        h->scloud.alloc(n);
    }
    h->x.alloc(n);
    h->y.alloc(n);
    h->z.alloc(n);
    h->perm.alloc(n);
    h->key.alloc(n);
    k_gather_sorted<<<ceil_div(n, BLK), BLK, 0, s>>>(tp, icloud, sidx, n, h->x.p, h->y.p, h->z.p, h->perm.p,
                                                      h->key.p, icloud ? h->scloud.p : nullptr);
    LAUNCH_CHECK();
}

}  // namespace

void build_grid(egt_ctx* h, const float* dx, const float* dy, const float* dz) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Scratch& sc = h->scratch;
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));

    // bounding cube
    const int nbp = std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(n, BLK)));
    float* dpart = sc.get<float>(8, (size_t)nbp * 6 + 1);
    uint32_t* dbad = reinterpret_cast<uint32_t*>(dpart + (size_t)nbp * 6);
    HIP_CHECK(hipMemsetAsync(dbad, 0, sizeof(uint32_t), s));
    k_minmax_partial<<<nbp, BLK, 0, s>>>(dx, dy, dz, n, dpart, dbad);
    LAUNCH_CHECK();
    std::vector<float> hp((size_t)nbp * 6 + 1);
    HIP_CHECK(hipMemcpyAsync(hp.data(), dpart, hp.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    uint32_t bad;
    std::memcpy(&bad, &hp[(size_t)nbp * 6], 4);
    if (bad) throw ArgError{EGT_ERR_INVALID_ARG, "input coordinates contain NaN or Inf"};
    float v[6] = {1e8f, -1e8f, 1e8f, -1e8f, 1e8f, -1e8f};
    for (int b = 0; b < nbp; ++b)
        for (int c = 0; c < 6; c += 2) {
            v[c] = std::min(v[c], hp[(size_t)b * 6 + c]);
            v[c + 1] = std::max(v[c + 1], hp[(size_t)b * 6 + c + 1]);
        }
    float mn, mx;
    cube_from_minmax(v, mn, mx);
    h->min_extent = mn;
    h->max_extent = mx;
    h->nclouds = 1;
    h->clouds.assign(1, CloudDev{});
    h->clouds[0].base = 0;
    h->clouds[0].n = n;
    h->clouds[0].min_extent = mn;
    h->clouds[0].extent = mx - mn;
    h->clouds[0].max_extent = mx;
    upload_clouds(h);
    sort_into_sorted_space(h, dx, dy, dz, nullptr);
    build_search_structure(h);
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.grid_ms, e0, e1));
}

void build_grid_batch(egt_ctx* h, const float* dx, const float* dy, const float* dz, const int64_t* offsets,
                      int nclouds) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Scratch& sc = h->scratch;
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    h->nclouds = nclouds;
    h->in_off.alloc(nclouds + 1);
    HIP_CHECK(hipMemcpyAsync(h->in_off.p, offsets, (nclouds + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    int32_t* icloud = sc.get<int32_t>(21, n);
    k_cloud_of<<<ceil_div(n, BLK), BLK, 0, s>>>(h->in_off.p, nclouds, n, icloud);
    LAUNCH_CHECK();
    int* mm = sc.get<int>(8, (size_t)nclouds * 6 + 1);
    uint32_t* dbad = reinterpret_cast<uint32_t*>(mm + (size_t)nclouds * 6);
    k_minmax_init<<<ceil_div(nclouds * 6, BLK), BLK, 0, s>>>(mm, nclouds);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemsetAsync(dbad, 0, sizeof(uint32_t), s));
    k_cloud_minmax<<<std::max(1, ceil_div(n, (int64_t)BLK * MM_PER)), BLK, 0, s>>>(dx, dy, dz, icloud, n, mm, dbad);
    LAUNCH_CHECK();
    std::vector<int> hm((size_t)nclouds * 6 + 1);
    HIP_CHECK(hipMemcpyAsync(hm.data(), mm, hm.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (hm[(size_t)nclouds * 6]) throw ArgError{EGT_ERR_INVALID_ARG, "input coordinates contain NaN or Inf"};
    h->clouds.assign(nclouds, CloudDev{});
    for (int c = 0; c < nclouds; ++c) {
        float v[6];
        for (int j = 0; j < 6; ++j) {
            int iv = hm[(size_t)c * 6 + j];
            iv = iv >= 0 ? iv : iv ^ 0x7FFFFFFF;
            std::memcpy(&v[j], &iv, 4);
        }
        float mn, mx;
        cube_from_minmax(v, mn, mx);
        CloudDev& cd = h->clouds[c];
        cd.base = offsets[c];
        cd.n = offsets[c + 1] - offsets[c];
        cd.min_extent = mn;
        cd.extent = mx - mn;
        cd.max_extent = mx;
    }
    h->min_extent = h->clouds[0].min_extent;
    h->max_extent = h->clouds[0].max_extent;
    upload_clouds(h);
    sort_into_sorted_space(h, dx, dy, dz, icloud);
    build_search_structure(h);
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.grid_ms, e0, e1));
}

void build_search_structure(egt_ctx* h) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    Scratch& sc = h->scratch;
    const int log2B = ilog2(h->B);
    const bool batch = h->nclouds > 1;
    DBuf<uint32_t>& code = sc.bufs[12];
    DBuf<uint32_t>& order = sc.bufs[13];
    code.alloc(n);
    order.alloc(n);
    const float a = -h->eps, ba = (1.0f + h->eps) - a;
    float4* spts = sc.get<float4>(11, n);  // sorted-space records (x, y, z, key)
    {
        // Morton order of the fine cells, points of a cell in sub-code order: stable sort by sub code, then
        // stable sort by cell code
        uint32_t* sub = sc.get<uint32_t>(22, n);
        uint32_t* cell = sc.get<uint32_t>(8, n);
        k_morton_keys<<<ceil_div(n, BLK), BLK, 0, s>>>(h->x.p, h->y.p, h->z.p, h->key.p, n, a, ba, h->B, log2B,
                                                       cell, sub, order.p, spts);
        LAUNCH_CHECK();
        radix_sort_pairs(sub, order.p, n, 3 * SUB, s, sc);
        k_gather_u32<<<ceil_div(n, BLK), BLK, 0, s>>>(cell, order.p, n, code.p);
        LAUNCH_CHECK();
    }
    radix_sort_pairs(code.p, order.p, n, 3 * log2B, s, sc);
    if (batch) {
        // second stable pass by cloud: Morton order within each cloud, clouds in input order
        uint32_t* cl = sc.get<uint32_t>(22, n);
        k_gather_cloud<<<ceil_div(n, BLK), BLK, 0, s>>>(h->scloud.p, order.p, n, cl);
        LAUNCH_CHECK();
        radix_sort_pairs(cl, order.p, n, std::max(1, ceil_log2(h->nclouds)), s, sc);
        h->mcloud.alloc(n);
    }
    h->mpts.alloc(n);
    h->mcell.alloc(n);
    h->mvcell.alloc(n);
    h->mcode.alloc(n);
    h->morder.alloc(n);
    h->minv.alloc(n);
    k_morton_points<<<ceil_div(n, BLK), BLK, 0, s>>>(spts, batch ? h->scloud.p : nullptr,
                                                      order.p, n, a, ba, h->B, log2B, h->mpts.p, h->mcell.p,
                                                      reinterpret_cast<unsigned long long*>(h->mvcell.p), h->mcode.p,
                                                      batch ? h->mcloud.p : nullptr, h->morder.p, h->minv.p);
    LAUNCH_CHECK();
    const int32_t* mcl = batch ? h->mcloud.p : nullptr;

    // level tables: level l groups 2^l fine cells per axis (Morton code >> 3l), per cloud. Counting pass, level
    // by level without a host wait: a flag per Morton position (first of its cell), a scan whose total stays on
    // the device, the cells' first positions (kept for the inserts) and, for a batch, the running count at
    // each cloud's first position. One read-back of every level's counts then sizes the tables.
    h->nlev = log2B + 1;
    const int nc = h->nclouds, nl = h->nlev;
    DBuf<uint32_t>& flag = sc.bufs[9];
    DBuf<uint32_t>& seg = sc.bufs[10];
    flag.alloc(n);
    seg.alloc(n);
    // room for the starts of every level: at most min(n, cells of the level's grid over all clouds) each
    std::vector<size_t> soff(nl + 1, 0);
    for (int l = 0; l < nl; ++l) {
        const double grid_cells = (double)nc * std::pow((double)(h->B >> l), 3.0);
        soff[l + 1] = soff[l] + (size_t)std::min<double>((double)n, grid_cells);
    }
    uint32_t* starts = sc.get<uint32_t>(38, soff[nl]);
    uint32_t* dtot = sc.get<uint32_t>(23, (size_t)nl * (nc + 2));  // per level: total, then the cloud bases
    for (int l = 0; l < nl; ++l) {
        k_level_flags<<<ceil_div(n, BLK), BLK, 0, s>>>(h->mcode.p, mcl, n, 3 * l, flag.p);
        LAUNCH_CHECK();
        uint32_t* tl = dtot + (size_t)l * (nc + 2);
        exclusive_scan_u32_dev(flag.p, seg.p, n, s, sc, tl);
        k_level_starts<<<ceil_div(n, BLK), BLK, 0, s>>>(flag.p, seg.p, n, starts + soff[l]);
        LAUNCH_CHECK();
        if (batch) {
            k_cells_at_bases<<<ceil_div(nc + 1, BLK), BLK, 0, s>>>(seg.p, h->dclouds.p, nc, tl, tl + 1);
            LAUNCH_CHECK();
        }
    }
    std::vector<uint32_t> ht((size_t)nl * (nc + 2));
    HIP_CHECK(hipMemcpyAsync(ht.data(), dtot, ht.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    h->cloud_cells.assign((size_t)nc * nl, 0);
    std::vector<uint32_t> cells(nl);
    uint64_t total = 0;
    LevelDev lv[EGT_MAX_LEVELS];
    for (int l = 0; l < nl; ++l) {
        const uint32_t* tl = &ht[(size_t)l * (nc + 2)];
        cells[l] = tl[0];
        if (batch) {
            for (int c = 0; c < nc; ++c) h->cloud_cells[(size_t)c * nl + l] = tl[1 + c + 1] - tl[1 + c];
        } else {
            h->cloud_cells[l] = cells[l];
        }
        uint32_t log2cap = 6;
        while ((1ull << log2cap) < 2ull * cells[l]) ++log2cap;
        h->lev[l].log2cap = log2cap;
        h->lev[l].offset = total;
        h->lev[l].cells = cells[l];
        lv[l] = LevelDev{total, log2cap, 0};
        total += 1ull << log2cap;
    }
    h->htab.alloc(total);
    k_fill_u64<<<1024, BLK, 0, s>>>(reinterpret_cast<uint64_t*>(h->htab.p), 2 * total, HASH_EMPTY);
    LAUNCH_CHECK();
    DBuf<uint32_t>& dmax = sc.bufs[14];
    dmax.alloc(nl);
    HIP_CHECK(hipMemsetAsync(dmax.p, 0, nl * sizeof(uint32_t), s));
    for (int l = 0; l < nl; ++l) {
        k_level_insert<<<std::max(1, ceil_div(cells[l], BLK)), BLK, 0, s>>>(
            starts + soff[l], cells[l], n, h->mcode.p, mcl, 3 * l, h->htab.p, h->lev[l].offset, h->lev[l].log2cap,
            reinterpret_cast<int*>(dmax.p) + l);
        LAUNCH_CHECK();
    }
    std::vector<int32_t> mx(nl);
    HIP_CHECK(hipMemcpyAsync(mx.data(), dmax.p, nl * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int l = 0; l < nl; ++l) {
        h->lev[l].max_count = mx[l];
        lv[l].max_count = mx[l];
    }
    h->dlev.alloc(EGT_MAX_LEVELS);
    HIP_CHECK(hipMemcpyAsync(h->dlev.p, lv, nl * sizeof(LevelDev), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

namespace {
__global__ void __launch_bounds__(BLK) k_rekey(const float* __restrict__ x, const float* __restrict__ y,
                                                const float* __restrict__ z, int64_t n, float a, float ba, int B,
                                                int log2B, int32_t* __restrict__ key) {
    const int64_t s = (int64_t)blockIdx.x * BLK + threadIdx.x;
    if (s >= n) return;
    const uint32_t cx = (uint32_t)cell_of(x[s], a, ba, B), cy = (uint32_t)cell_of(y[s], a, ba, B),
                   cz = (uint32_t)cell_of(z[s], a, ba, B);
    key[s] = (int32_t)(cx | (cy << log2B) | (cz << (2 * log2B)));
}
}  // namespace

// After smoothing moved the points: recompute each point's fine cell and rebuild the Morton search
// structure (the sorted-space order itself is kept).
void refresh_search_structure(egt_ctx* h) {
    const float a = -h->eps, ba = (1.0f + h->eps) - a;
    k_rekey<<<ceil_div(h->n, BLK), BLK, 0, h->stream>>>(h->x.p, h->y.p, h->z.p, h->n, a, ba, h->B, ilog2(h->B),
                                                        h->key.p);
    LAUNCH_CHECK();
    build_search_structure(h);
    h->search_stale = false;
}

void run_output(egt_ctx* h, float* x, float* y, float* z) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    float* ox = h->scratch.get<float>(9, 3 * n);
    k_output<<<ceil_div(n, BLK), BLK, 0, s>>>(h->x.p, h->y.p, h->z.p, h->perm.p,
                                              h->nclouds > 1 ? h->scloud.p : nullptr, h->dclouds.p, n,
                                              h->min_extent, h->max_extent - h->min_extent, ox, ox + n, ox + 2 * n);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(x, ox, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(y, ox + n, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(z, ox + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void run_knn_original(egt_ctx* h, int32_t* out) {
    const int64_t n = h->n, k = h->k;
    hipStream_t s = h->stream;
    int32_t* tmp = h->scratch.get<int32_t>(9, n * k);
    k_knn_original<<<ceil_div(n * k, BLK), BLK, 0, s>>>(h->knn.p, h->perm.p, n, (int)k, tmp);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(out, tmp, n * k * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void run_knn_original_cloud(egt_ctx* h, int cloud, int32_t* out) {
    const CloudDev& c = h->clouds[cloud];
    hipStream_t s = h->stream;
    const int64_t cnt = c.n * c.k;
    int32_t* tmp = h->scratch.get<int32_t>(9, std::max<int64_t>(cnt, 1));
    k_knn_original_cloud<<<std::max(1, ceil_div(cnt, BLK)), BLK, 0, s>>>(h->knn.p, h->perm.p, c, c.base, tmp);
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(out, tmp, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) { return morton_code(x, y, z); }

}  // namespace egt
