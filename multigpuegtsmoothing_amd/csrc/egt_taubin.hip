// Taubin lambda/mu smoothing over the EGG adjacency (CSR), Jacobi ping-pong.
//
// Reference: taubin_step_GKNN EGT_CUDA.cu:229-276, iteration loop :545-634 (one iteration = a lambda pass
// then a mu pass; weights w = exp(-|pn - pi|^2), isRegularized = 0 :544).
//
// Points are held as float4 (x, y, z, 0) in two buffers so a neighbour gather is one 16-byte load; the
// 2*iterations passes are recorded once into a hipGraph and replayed. The exp is egt_expf (egt_device.h),
// so results match the host oracle bit for bit.
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

constexpr int TBLK = 256;

__global__ void __launch_bounds__(TBLK) k_taubin_pass(const float4* __restrict__ in, float4* __restrict__ out,
                                                       const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col, int64_t n, float scale) {
    const int64_t i = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (i >= n) return;
    const float4 p = in[i];
    float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
    const int32_t e0 = rowptr[i], e1 = rowptr[i + 1];
    for (int32_t e = e0; e < e1; ++e) {
        const float4 q = in[col[e]];
        const float dist = sqd(p.x, p.y, p.z, q.x, q.y, q.z);
        const float w = egt_expf(-dist);
        cgx = __builtin_fmaf(w, q.x - p.x, cgx);
        cgy = __builtin_fmaf(w, q.y - p.y, cgy);
        cgz = __builtin_fmaf(w, q.z - p.z, cgz);
        sum += w;
    }
    if (sum == 0.0f) sum = 1e-8f;
    out[i] = make_float4(p.x + scale * cgx / sum, p.y + scale * cgy / sum, p.z + scale * cgz / sum, 0.0f);
}

__global__ void __launch_bounds__(TBLK) k_soa_to_aos(const float* __restrict__ x, const float* __restrict__ y,
                                                      const float* __restrict__ z, int64_t n, float4* __restrict__ o) {
    const int64_t i = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (i < n) o[i] = make_float4(x[i], y[i], z[i], 0.0f);
}

__global__ void __launch_bounds__(TBLK) k_aos_to_soa(const float4* __restrict__ p, int64_t n, float* __restrict__ x,
                                                      float* __restrict__ y, float* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * TBLK + threadIdx.x;
    if (i >= n) return;
    const float4 v = p[i];
    x[i] = v.x;
    y[i] = v.y;
    z[i] = v.z;
}

}  // namespace

void run_taubin(egt_ctx* h, int iters, float lambda, float mu) {
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    float4* A = h->scratch.get<float4>(12, n);
    float4* Bf = h->scratch.get<float4>(13, n);
    const dim3 grid(ceil_div(n, TBLK)), blk(TBLK);
    k_soa_to_aos<<<grid, blk, 0, s>>>(h->x.p, h->y.p, h->z.p, n, A);
    LAUNCH_CHECK();
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
                k_taubin_pass<<<grid, blk, 0, s>>>(A, Bf, h->rowptr.p, h->col.p, n, lambda);
                k_taubin_pass<<<grid, blk, 0, s>>>(Bf, A, h->rowptr.p, h->col.p, n, mu);
            }
            hipError_t cap = hipStreamEndCapture(s, &g);
            HIP_CHECK(cap);
            HIP_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphLaunch(ge, s));
            HIP_CHECK(hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
        } else {
            for (int it = 0; it < iters; ++it) {
                k_taubin_pass<<<grid, blk, 0, s>>>(A, Bf, h->rowptr.p, h->col.p, n, lambda);
                k_taubin_pass<<<grid, blk, 0, s>>>(Bf, A, h->rowptr.p, h->col.p, n, mu);
            }
            LAUNCH_CHECK();
        }
    }
    HIP_CHECK(hipEventRecord(e1, s));
    k_aos_to_soa<<<grid, blk, 0, s>>>(A, n, h->x.p, h->y.p, h->z.p);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.taubin_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace egt
