// Elliptic Gabriel graph filter (in place on the neighbour rows) and the CSR adjacency it feeds to Taubin.
//
// Reference: calculateEGG_GKNN EGT_CUDA.cu:172-189, isEllipicGabrielNeighbor_GKNN :136-170, frame helpers
// :55-134, batch driver :505-530.
//
// One thread per point; the point's neighbour coordinates are staged in LDS as [slot][lane] so every slot
// read is bank-conflict free. Slot j (ranked by distance) is dropped when a valid slot m < j lies strictly
// inside the ellipsoid with semi-axis |pq_j|/2 along pq_j and alpha*|pq_j|/2 across it. Tests for slot j only
// read the original rows of slots m < j, so the result equals the reference's in-place downward sweep.
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int EBLK = 64;

__device__ inline void normalize3(float& x, float& y, float& z) {
    const float nrm = __builtin_sqrtf(__builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x)));
    x = x / nrm;
    y = y / nrm;
    z = z / nrm;
}
__device__ inline float rowt(float a1, float a2, float a3, float t, float x, float y, float z) {
    return __builtin_fmaf(a3, z, __builtin_fmaf(a2, y, a1 * x)) + t;
}

template <int K>
__global__ void __launch_bounds__(EBLK) k_egg(const float* __restrict__ xs, const float* __restrict__ ys,
                                               const float* __restrict__ zs, int32_t* __restrict__ knn, int64_t n,
                                               int k, float alpha, uint32_t* __restrict__ kept) {
    __shared__ float sx[K][EBLK], sy[K][EBLK], sz[K][EBLK];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * EBLK + tid;
    if (i >= n) return;
    int32_t* row = knn + i * k;
    uint64_t valid = 0;
    for (int j = 0; j < k; ++j) {
        const int32_t nb = row[j];
        if (nb >= 0) {
            valid |= 1ull << j;
            sx[j][tid] = xs[nb];
            sy[j][tid] = ys[nb];
            sz[j][tid] = zs[nb];
        }
    }
    const float x = xs[i], y = ys[i], z = zs[i];
    const float aa = alpha * alpha;
    uint64_t keep = valid;
    for (int j = k - 1; j >= 0; --j) {
        if (!((valid >> j) & 1ull)) continue;
        const float xi = sx[j][tid], yi = sy[j][tid], zi = sz[j][tid];
        const float xo = 0.5f * (xi + x), yo = 0.5f * (yi + y), zo = 0.5f * (zi + z);
        const float d = __builtin_sqrtf(sqd(xi, yi, zi, x, y, z)) / 2.0f;
        if ((double)d < 1e-6) continue;  // kept
        float ax = xi - x, ay = yi - y, az = zi - z;
        normalize3(ax, ay, az);
        float mxv = __builtin_fabsf(ax);
        int ci = 0;
        if (mxv < __builtin_fabsf(ay)) {
            ci = 1;
            mxv = __builtin_fabsf(ay);
        }
        if (mxv < __builtin_fabsf(az)) ci = 2;
        float bx = 1.0f, by = 1.0f, bz = 1.0f;
        if (ci == 0) bx = (-ay - az) / ax;
        else if (ci == 1) by = (-ax - az) / ay;
        else bz = (-ax - ay) / az;
        normalize3(bx, by, bz);
        float cx = ay * bz - by * az;
        float cy = bx * az - ax * bz;
        float cz = ax * by - bx * ay;
        normalize3(cx, cy, cz);
        const float tx = -rowt(ax, ay, az, 0.0f, xo, yo, zo);
        const float ty = -rowt(bx, by, bz, 0.0f, xo, yo, zo);
        const float tz = -rowt(cx, cy, cz, 0.0f, xo, yo, zo);
        const float dd = d * d;
        uint64_t mset = valid & ((1ull << j) - 1ull);
        while (mset) {
            const int mm = __builtin_ctzll(mset);
            mset &= mset - 1ull;
            const float px = sx[mm][tid], py = sy[mm][tid], pz = sz[mm][tid];
            const float ex = rowt(ax, ay, az, tx, px, py, pz);
            const float ey = rowt(bx, by, bz, ty, px, py, pz);
            const float ez = rowt(cx, cy, cz, tz, px, py, pz);
            const float v = ex * ex + ey * ey / aa + ez * ez / aa;
            if (v < dd) {
                keep &= ~(1ull << j);
                break;
            }
        }
    }
    uint64_t dropped = valid & ~keep;
    while (dropped) {
        const int j = __builtin_ctzll(dropped);
        dropped &= dropped - 1ull;
        row[j] = -1;
    }
    kept[i] = (uint32_t)__popcll(keep);
}

__global__ void __launch_bounds__(256) k_row_counts(const int32_t* __restrict__ knn, int64_t n, int k,
                                                    uint32_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t c = 0;
    for (int j = 0; j < k; ++j) c += knn[i * k + j] >= 0;
    cnt[i] = c;
}

__global__ void __launch_bounds__(256) k_csr_fill(const int32_t* __restrict__ knn, int64_t n, int k,
                                                  const uint32_t* __restrict__ off, int32_t* __restrict__ rowptr,
                                                  int32_t* __restrict__ col, uint32_t total) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t o = off[i];
    rowptr[i] = (int32_t)o;
    if (i == n - 1) rowptr[n] = (int32_t)total;
    for (int j = 0; j < k; ++j) {
        const int32_t v = knn[i * k + j];
        if (v >= 0) col[o++] = v;
    }
}

}  // namespace

void build_csr_from_rows(egt_ctx* h) {
    const int64_t n = h->n;
    const int k = h->k;
    hipStream_t s = h->stream;
    uint32_t* cnt = h->scratch.get<uint32_t>(9, n);
    uint32_t* off = h->scratch.get<uint32_t>(10, n);
    k_row_counts<<<ceil_div(n, 256), 256, 0, s>>>(h->knn.p, n, k, cnt);
    LAUNCH_CHECK();
    uint32_t total = 0;
    exclusive_scan_u32(cnt, off, n, s, h->scratch, &total);
    h->rowptr.alloc(n + 1);
    h->col.alloc(std::max<uint32_t>(total, 1));
    k_csr_fill<<<ceil_div(n, 256), 256, 0, s>>>(h->knn.p, n, k, off, h->rowptr.p, h->col.p, total);
    LAUNCH_CHECK();
    h->nnz = total;
    h->stats.egg_kept = total;
}

void run_egg(egt_ctx* h, float alpha) {
    const int64_t n = h->n;
    const int k = h->k;
    hipStream_t s = h->stream;
    uint32_t* kept = h->scratch.get<uint32_t>(9, n);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    const dim3 grid(ceil_div(n, EBLK)), blk(EBLK);
    if (k <= 16) hipLaunchKernelGGL(k_egg<16>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, h->knn.p, n, k, alpha, kept);
    else if (k <= 32) hipLaunchKernelGGL(k_egg<32>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, h->knn.p, n, k, alpha, kept);
    else if (k <= 40) hipLaunchKernelGGL(k_egg<40>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, h->knn.p, n, k, alpha, kept);
    else hipLaunchKernelGGL(k_egg<64>, grid, blk, 0, s, h->x.p, h->y.p, h->z.p, h->knn.p, n, k, alpha, kept);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventRecord(e1, s));
    build_csr_from_rows(h);
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.egg_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace egt
