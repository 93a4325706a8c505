// Elliptic Gabriel graph filter (in place on the sorted-space neighbour rows) and the Morton-order CSR
// adjacency it feeds to Taubin and the normals.
//
// Reference: calculateEGG_GKNN EGT_CUDA.cu:172-189, isEllipicGabrielNeighbor_GKNN :136-170, frame helpers
// :55-134, batch driver :505-530.
//
// Lanes walk the points in Morton order so the neighbour coordinates a wave gathers sit in a few cache lines;
// 4, 8 or 16 lanes share a point. The point's neighbour coordinates are staged in LDS as [slot][point]
// (bank-conflict free). Slot j (ranked by distance) is dropped when a valid slot m < j lies strictly inside
// the ellipsoid with semi-axis |pq_j|/2 along pq_j and alpha*|pq_j|/2 across it. Tests for slot j only read
// the original entries of slots m < j, so the result equals the reference's in-place downward sweep. A
// conservative bounding-ball test skips candidates that cannot be inside before the exact frame test.
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

// Ellipsoid frame of slot j: rows a, b, c, translation t, squared semi-axis dd, ball pre-test radius^2 rb2
// around the midpoint o.
struct EggFrame {
    float ax, ay, az, bx, by, bz, cx, cy, cz, tx, ty, tz, xo, yo, zo, dd, rb2, lo, hi;
};

// True when one valid slot m < j (coordinates in LDS at sx[m * stride + col], ...) lies strictly inside the
// ellipsoid: v = ex^2 + ey^2 / aa + ez^2 / aa < dd. Slots are read four at a time, their LDS loads issued
// together. The two correctly rounded divisions are only evaluated near the boundary: the estimate
// vf = ex^2 + (ey^2 + ez^2) * (1 / aa) sums non-negative terms, so it is within about 7 units in the last
// place (4.2e-7 relative) of v as written; outside dd (1 -+ 1e-5) the estimate decides as v would.
__device__ inline bool egg_inside(const float* __restrict__ sx, const float* __restrict__ sy,
                                  const float* __restrict__ sz, int stride, int col, int j, int kmax,
                                  uint64_t valid, const EggFrame& F, float aa, float inv_aa) {
    for (int m0 = 0; m0 < j; m0 += 4) {
        float px[4], py[4], pz[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int mm = min(m0 + u, kmax - 1);
            px[u] = sx[mm * stride + col];
            py[u] = sy[mm * stride + col];
            pz[u] = sz[mm * stride + col];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int mm = m0 + u;
            if (mm >= j || !((valid >> mm) & 1ull)) continue;
            if (sqd(px[u], py[u], pz[u], F.xo, F.yo, F.zo) >= F.rb2) continue;
            const float ex = rowt(F.ax, F.ay, F.az, F.tx, px[u], py[u], pz[u]);
            const float ey = rowt(F.bx, F.by, F.bz, F.ty, px[u], py[u], pz[u]);
            const float ez = rowt(F.cx, F.cy, F.cz, F.tz, px[u], py[u], pz[u]);
            const float e2 = ex * ex, f2 = ey * ey, g2 = ez * ez;
            const float vf = e2 + (f2 + g2) * inv_aa;
            if (vf < F.lo) return true;
            if (vf > F.hi) continue;
            if (e2 + f2 / aa + g2 / aa < F.dd) return true;
        }
    }
    return false;
}

// Frame of slot j (coordinates xi, yi, zi) for the point (x, y, z); false when |pq|/2 < 1e-6 (always kept).
__device__ inline bool egg_frame(float x, float y, float z, float xi, float yi, float zi, float alpha,
                                 EggFrame& F) {
    F.xo = 0.5f * (xi + x);
    F.yo = 0.5f * (yi + y);
    F.zo = 0.5f * (zi + z);
    const float d = __builtin_sqrtf(sqd(xi, yi, zi, x, y, z)) / 2.0f;
    if ((double)d < 1e-6) return false;
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
    F.ax = ax; F.ay = ay; F.az = az;
    F.bx = bx; F.by = by; F.bz = bz;
    F.cx = cx; F.cy = cy; F.cz = cz;
    F.tx = -rowt(ax, ay, az, 0.0f, F.xo, F.yo, F.zo);
    F.ty = -rowt(bx, by, bz, 0.0f, F.xo, F.yo, F.zo);
    F.tz = -rowt(cx, cy, cz, 0.0f, F.xo, F.yo, F.zo);
    F.dd = d * d;
    F.lo = F.dd * (1.0f - 1e-5f);
    F.hi = F.dd * (1.0f + 1e-5f);
    // The ellipsoid lies in the ball of radius d*max(1, alpha) around the midpoint. A candidate outside that
    // ball padded by a relative 1e-4 and an absolute 4e-6 (well above the ~1e-6 rounding of the frame
    // transform on unit-cube coordinates) cannot pass the exact test, so it is skipped.
    const float rb = d * fmaxf(1.0f, alpha) * 1.0001f + 4e-6f;
    F.rb2 = rb * rb;
    return true;
}

// The filter with LPP lanes per point (LPP consecutive lanes of a wave share one point). mpts: current points
// in Morton order (w = sorted index); knn: sorted-space rows, updated in place; mrow: per Morton position, the
// k slots mapped to Morton positions (-1 for empty); keepm: kept-slot mask. The point's neighbour coordinates
// are staged once in LDS ([slot][point], 12 B per slot per point), and lane r tests the slots
// j = r, r + LPP, ... against all valid m < j. LDS per lane drops LPP-fold, so several times more waves fit on
// a CU to hide the gather latency. A workgroup of 64 lanes covers 64 / LPP points; the 64-point chunks of a
// neighbour-count class (chunks[b] = first Morton position, null: 64 b) take LPP workgroups each.
template <int K, int LPP>
__global__ void __launch_bounds__(EBLK) k_egg_shared(const float4* __restrict__ mpts, const int32_t* __restrict__ minv,
                                                      int32_t* __restrict__ knn, const CloudDev* __restrict__ clouds,
                                                      const int32_t* __restrict__ mcloud,
                                                      const int64_t* __restrict__ chunks, float alpha,
                                                      int32_t* __restrict__ mrow, uint64_t* __restrict__ keepm,
                                                      uint32_t* __restrict__ kept) {
    constexpr int PPB = EBLK / LPP;  // points per workgroup
    __shared__ float sx[K][PPB], sy[K][PPB], sz[K][PPB];
    const int tid = threadIdx.x, pi = tid / LPP, r = tid % LPP;
    const int64_t chunk = blockIdx.x / LPP, sub = blockIdx.x % LPP;
    const int64_t start = chunks ? chunks[chunk] : chunk * EBLK;
    const CloudDev cd = clouds[mcloud ? mcloud[start] : 0];
    const int k = cd.k;
    const int64_t m = start + sub * PPB + pi;
    const bool act = m < cd.base + cd.n;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    int32_t* row = nullptr;
    int32_t* mr = nullptr;
    const int32_t base = (int32_t)cd.base;
    constexpr int NJ = (K + LPP - 1) / LPP;  // slots per lane
    uint64_t valid = 0;
    if (act) {
        p = mpts[m];
        const int32_t s = __float_as_int(p.w);
        row = knn + cd.row_off + (int64_t)(s - cd.base) * k;
        mr = mrow + cd.row_off + (m - cd.base) * k;
        int32_t nbv[NJ];
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            const int j = r + u * LPP;
            nbv[u] = j < k ? row[j] : -1;
        }
#pragma unroll
        for (int u = 0; u < NJ; ++u) nbv[u] = nbv[u] >= 0 ? minv[nbv[u] + base] : -1;
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            const int j = r + u * LPP;
            if (nbv[u] >= 0) {
                valid |= 1ull << j;
                const float4 c = mpts[nbv[u]];
                sx[j][pi] = c.x;
                sy[j][pi] = c.y;
                sz[j][pi] = c.z;
            }
            if (j < k) mr[j] = nbv[u];
        }
    }
    // the valid mask of the point: OR over its LPP lanes
#pragma unroll
    for (int d = 1; d < LPP; d <<= 1) valid |= __shfl_xor(valid, d, 64);
    __syncthreads();
    uint64_t drop = 0;
    if (act) {
        const float x = p.x, y = p.y, z = p.z;
        const float aa = alpha * alpha;
        const float inv_aa = 1.0f / aa;
        for (int u = NJ - 1; u >= 0; --u) {
            const int j = r + u * LPP;
            if (j >= K || !((valid >> j) & 1ull)) continue;
            EggFrame F;
            if (!egg_frame(x, y, z, sx[j][pi], sy[j][pi], sz[j][pi], alpha, F)) continue;  // kept
            if (egg_inside(&sx[0][0], &sy[0][0], &sz[0][0], PPB, pi, j, K, valid, F, aa, inv_aa)) {
                drop |= 1ull << j;
                row[j] = -1;
            }
        }
    }
#pragma unroll
    for (int d = 1; d < LPP; d <<= 1) drop |= __shfl_xor(drop, d, 64);
    if (act && r == 0) {
        const uint64_t keep = valid & ~drop;
        keepm[m] = keep;
        kept[m] = (uint32_t)__popcll(keep);
    }
}

// One wave per row, CSR_ROWS consecutive rows per wave: lane j reads slot j of the row (one contiguous load
// of the k slots) and writes it, when kept, at the row's offset plus the number of kept slots below j, so the
// row's kept columns go out as one contiguous store. (A thread per row read and wrote its own row's slots, a
// cache line per lane and per slot.) The rows' offsets, keep masks and row starts are loaded once, lane r
// holding row m0 + r, and read back per row with readlane, so the rows' slot loads do not wait on each other.
constexpr int CSR_ROWS = 16;
__global__ void __launch_bounds__(256) k_csr_fill(const int32_t* __restrict__ mrow, const uint64_t* __restrict__ keepm,
                                                  int64_t n, const CloudDev* __restrict__ clouds,
                                                  const int32_t* __restrict__ mcloud, const uint32_t* __restrict__ off,
                                                  int32_t* __restrict__ rowptr, int32_t* __restrict__ col,
                                                  uint32_t total) {
    const int lane = __lane_id();
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t m0 = w * CSR_ROWS;
    if (m0 >= n) return;  // wave-uniform
    const int nr = (int)min<int64_t>(CSR_ROWS, n - m0);
    const int64_t mr = m0 + min(lane & (CSR_ROWS - 1), nr - 1);
    const uint32_t o_l = off[mr];
    const uint64_t keep_l = keepm[mr];
    const CloudDev& cd = clouds[mcloud ? mcloud[mr] : 0];
    const int64_t rb_l = cd.row_off + (mr - cd.base) * cd.k;
    const int k_l = cd.k;
    if (lane < nr) rowptr[m0 + lane] = (int32_t)o_l;
    if (m0 + nr == n && lane == 0) rowptr[n] = (int32_t)total;
    // The rows of consecutive Morton positions lie back to back in mrow. When the wave's rows share one k (all
    // but the chunks where a cloud ends), they are read as one block, 64 slots per load (a row per load left
    // 64 - k lanes idle: 48 at k=16); slot idx of the block is slot idx - r k of row r = floor(idx / k), and
    // the row's offset and keep mask come from lane r. On a 16M-point mixture, EGG + CSR 4.52 -> 3.90 ms at k=16
    // and 13.25 -> 12.76 ms at k=40 (k=64: unchanged, a row already fills the wave).
    const int k0 = __builtin_amdgcn_readfirstlane(k_l);
    if (!__ballot(lane < nr && k_l != k0)) {
        const int64_t rb0 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb_l, 0)) |
                                      ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)((uint64_t)rb_l >> 32), 0)
                                       << 32));
        const int T = nr * k0;  // at most 16 x 64 slots
        // (idx + 0.5) / k is at least 0.5 / k >= 1/128 away from an integer and below 1024, far more than the
        // rounding of the product, so the truncation is the exact row
        const float rk = 1.0f / (float)k0;
        for (int base = 0; base < T; base += 64) {
            const int idx = base + lane;
            const bool in = idx < T;
            const int r = min((int)(((float)idx + 0.5f) * rk), nr - 1);
            const int slot = idx - r * k0;
            const int32_t v = mrow[rb0 + (in ? idx : 0)];
            const uint32_t o = (uint32_t)__shfl((int)o_l, r, 64);
            const uint64_t keep = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)keep_l, r, 64)) |
                                  ((uint64_t)(uint32_t)__shfl((int)(keep_l >> 32), r, 64) << 32);
            if (in && ((keep >> slot) & 1ull)) col[o + (uint32_t)__popcll(keep & ((1ull << slot) - 1ull))] = v;
        }
        return;
    }
    const uint64_t lt = (1ull << lane) - 1ull;  // lanes below this one
#pragma unroll
    for (int i = 0; i < CSR_ROWS; ++i) {
        // rows past the end repeat the last row: the same values to the same places, and no branch in the loop
        const int r = min(i, nr - 1);
        const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)o_l, r);
        const uint64_t keep = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)keep_l, r)) |
                              ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(keep_l >> 32), r) << 32);
        const int64_t rb = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)rb_l, r)) |
                                     ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)((uint64_t)rb_l >> 32), r)
                                      << 32));
        const int k = __builtin_amdgcn_readlane(k_l, r);
        const int32_t v = mrow[rb + min(lane, k - 1)];
        if ((keep >> lane) & 1ull) col[o + (uint32_t)__popcll(keep & lt)] = v;
    }
}

// Rows -> Morton-order slots without filtering (normals requested straight after the neighbour search).
__global__ void __launch_bounds__(256) k_rows_to_mrow(const int32_t* __restrict__ morder,
                                                      const int32_t* __restrict__ minv,
                                                      const int32_t* __restrict__ knn, int64_t n,
                                                      const CloudDev* __restrict__ clouds,
                                                      const int32_t* __restrict__ mcloud, int32_t* __restrict__ mrow,
                                                      uint64_t* __restrict__ keepm, uint32_t* __restrict__ kept) {
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= n) return;
    const CloudDev& cd = clouds[mcloud ? mcloud[m] : 0];
    const int k = cd.k;
    const int32_t* row = knn + cd.row_off + (int64_t)(morder[m] - cd.base) * k;
    int32_t* mr = mrow + cd.row_off + (m - cd.base) * k;
    uint64_t valid = 0;
    for (int j = 0; j < k; ++j) {
        const int32_t nb = row[j];
        mr[j] = nb >= 0 ? minv[nb + cd.base] : -1;
        if (nb >= 0) valid |= 1ull << j;
    }
    keepm[m] = valid;
    kept[m] = (uint32_t)__popcll(valid);
}

int64_t row_elems(const egt_ctx* h) {
    int64_t t = 0;
    for (const CloudDev& c : h->clouds) t += c.n * c.k;
    return t;
}

void finish_csr(egt_ctx* h, const int32_t* mrow, const uint64_t* keepm, const uint32_t* kept) {
    const int64_t n = h->n;
    // The scan below and the CSR offsets are 32-bit. Only when the rows hold more slots than that can carry
    // is the exact kept count needed (a 64-bit sum; a 512-cloud batch of 54.7M points at mean k 40 holds
    // 2.19G slots, of which the EGG filter keeps far fewer).
    if (row_elems(h) > CSR_MAX_ENTRIES) {
        if (const char* err = csr_total_error(sum_u32(kept, n, h->stream, h->scratch)))
            throw ArgError{EGT_ERR_UNSUPPORTED, err};
    }
    uint32_t* off = h->scratch.get<uint32_t>(10, n);
    uint32_t total = 0;
    exclusive_scan_u32(kept, off, n, h->stream, h->scratch, &total);
    h->rowptr.alloc(n + 1);
    h->col.alloc(std::max<uint32_t>(total, 1));
    k_csr_fill<<<ceil_div(n, 4 * CSR_ROWS), 256, 0, h->stream>>>(mrow, keepm, n, h->dclouds.p,
                                                         h->nclouds > 1 ? h->mcloud.p : nullptr, off, h->rowptr.p,
                                                         h->col.p, total);
    LAUNCH_CHECK();
    h->nnz = total;
    h->csr_valid = true;
}

}  // namespace

void build_csr_from_rows(egt_ctx* h) {
    const int64_t n = h->n;
    sync_single_cloud(h);
    uint32_t* kept = h->scratch.get<uint32_t>(9, n);
    uint64_t* keepm = h->scratch.get<uint64_t>(17, n);
    int32_t* mrow = h->scratch.get<int32_t>(18, (size_t)std::max<int64_t>(row_elems(h), 1));
    k_rows_to_mrow<<<ceil_div(n, 256), 256, 0, h->stream>>>(h->morder.p, h->minv.p, h->knn.p, n, h->dclouds.p,
                                                            h->nclouds > 1 ? h->mcloud.p : nullptr, mrow, keepm,
                                                            kept);
    LAUNCH_CHECK();
    finish_csr(h, mrow, keepm, kept);
}

void run_egg(egt_ctx* h, float alpha) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    sync_single_cloud(h);
    uint32_t* kept = h->scratch.get<uint32_t>(9, n);
    uint64_t* keepm = h->scratch.get<uint64_t>(17, n);
    int32_t* mrow = h->scratch.get<int32_t>(18, (size_t)std::max<int64_t>(row_elems(h), 1));
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    // one launch per neighbour-count class over its chunk list (one cloud: every chunk, class of k)
    std::vector<WorkClass> classes = h->work_classes;
    if (h->nclouds == 1) classes.assign(1, WorkClass{h->clouds[0].k, 0, (n + EBLK - 1) / EBLK});
    const int32_t* mcl = h->nclouds > 1 ? h->mcloud.p : nullptr;
    for (const WorkClass& wc : classes) {
        const int64_t* chunks = h->nclouds > 1 ? h->work_lists.p + wc.first : nullptr;
        const dim3 grid4((unsigned)(wc.count * 4)), grid8((unsigned)(wc.count * 8)),
            grid16((unsigned)(wc.count * 16)), blk(EBLK);
#define EGG_ARGS h->mpts.p, h->minv.p, h->knn.p, h->dclouds.p, mcl, chunks, alpha, mrow, keepm, kept
        // lanes per point by list length (measured on a 4M-point torus and a 16M-point mixture: k=16 with 4
        // lanes 1.11 ms against 1.43 for one lane; k=40 with 8 lanes 3.95 ms against 4.27 with 4 and 4.84 with
        // 16; k=64 with 16 lanes 6.89 ms against 7.15 with 8 and 10.4 with 4)
        if (wc.K <= 16) hipLaunchKernelGGL((k_egg_shared<16, 4>), grid4, blk, 0, s, EGG_ARGS);
        else if (wc.K <= 32) hipLaunchKernelGGL((k_egg_shared<32, 8>), grid8, blk, 0, s, EGG_ARGS);
        else if (wc.K <= 40) hipLaunchKernelGGL((k_egg_shared<40, 8>), grid8, blk, 0, s, EGG_ARGS);
        else hipLaunchKernelGGL((k_egg_shared<64, 16>), grid16, blk, 0, s, EGG_ARGS);
#undef EGG_ARGS
        LAUNCH_CHECK();
    }
    finish_csr(h, mrow, keepm, kept);
    h->stats.egg_kept = h->nnz;
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.egg_ms, e0, e1));
}

}  // namespace egt
