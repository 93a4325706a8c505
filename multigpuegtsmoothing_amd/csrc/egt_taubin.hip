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
                const float w = egt_expf(-dist);
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

void run_taubin(egt_ctx* h, int iters, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    float4* A = h->scratch.get<float4>(12, n);
    float4* Bf = h->scratch.get<float4>(13, n);
    const int nb = ceil_div(n, TBLK);
    const dim3 grid(ceil_div(nb, NXCD) * NXCD), blk(TBLK);
    HIP_CHECK(hipMemcpyAsync(A, h->mpts.p, n * sizeof(float4), hipMemcpyDeviceToDevice, s));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, s));
    if (iters > 0) {
        if (h->taubin_graph) {
            hipGraph_t g;
            hipGraphExec_t ge;
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int it = 0; it < iters; ++it) {
                k_taubin_pass<<<grid, blk, 0, s>>>(A, Bf, h->rowptr.p, h->col.p, n, lambda, nb);
                k_taubin_pass<<<grid, blk, 0, s>>>(Bf, A, h->rowptr.p, h->col.p, n, mu, nb);
            }
            HIP_CHECK(hipStreamEndCapture(s, &g));
            HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphLaunch(ge, s));
            HIP_CHECK(hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
        } else {
            for (int it = 0; it < iters; ++it) {
                k_taubin_pass<<<grid, blk, 0, s>>>(A, Bf, h->rowptr.p, h->col.p, n, lambda, nb);
                k_taubin_pass<<<grid, blk, 0, s>>>(Bf, A, h->rowptr.p, h->col.p, n, mu, nb);
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
