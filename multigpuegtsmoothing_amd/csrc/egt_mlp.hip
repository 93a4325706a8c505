// Learned search-radius proposal: a 28 -> 64 -> 64 -> 1 ReLU MLP with BF16 MFMA forward and backward, FP32
// gradient reduction and an FP32 Adam step, plus the feature, target and level-selection kernels that let
// the kNN start each query at its proposed pyramid level. No counterpart in the reference, which always
// starts at the query's own cell (GridKNNCuda.cu:2409-2474); the kNN's exactness test (a bound derived from
// the reference certificate GridKNNCuda.cu:717-740) expands past any proposal that is too small, so the
// neighbour rows are the same with or without the model.
//
// Features of a query (NF = 28, padded with zeros to NIN = 32): log2(1 + count) / 16 of the 27 cells of the
// 3x3x3 block around its cell at the reference level, cells in x-fastest order (dx, dy, dz in -1..1), and
// log2(k) / 8. Target: log2(r_k / h_ref), r_k = exact k-th neighbour distance of the plain search, h_ref =
// cell edge at the reference level (the start level the plain search would use).
//
// MFMA v_mfma_f32_16x16x32_bf16 (A 16x32, B 32x16, D 16x16 FP32): lane l supplies A row (l & 15) and B
// column (l & 15) with the same eight K indices, and holds D column (l & 15), rows 4 * (l >> 4) .. +3.
// A wave handles a 32-sample tile; activations sit in LDS as bf16 so operands can be read in either
// orientation (samples as M for the forward pass, samples as K for the weight gradients).
//
// Rounding: every MFMA operand is rounded to bf16 (round to nearest even) and products accumulate in FP32.
// Biases, ReLU, the 64 -> 1 output layer (bf16 inputs, FP32 sequential fmaf), the loss, the bias gradients
// and Adam run in FP32. Gradients are reduced per wave and then summed in wave order: deterministic.
#include <algorithm>
#include <cmath>
#include <vector>
#include "egt_device.h"
#include "egt_internal.h"

namespace egt {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NF = 28, NIN = 32, NH = 64;
// FP32 parameter vector: W1[32][64] (rows 28..31 stay zero), b1[64], W2[64][64], b2[64], W3[64], b3
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, NPARAM = OFF_B3 + 1;
constexpr int MBLK = 256;  // 4 waves
constexpr int WPB = MBLK / 64;
constexpr int TS = 32;  // samples per wave tile
constexpr int MAX_GRAD_WAVES = 256;

__device__ inline __bf16 bf(float x) { return (__bf16)x; }
__device__ inline float fb(__bf16 x) { return (float)x; }

__device__ inline void wave_sync() {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

struct WaveLds {
    __bf16 X[TS][NIN];
    __bf16 H1[TS][NH];  // reused for dH1
    __bf16 H2[TS][NH];  // reused for dH2
    float y[TS];
    float dy[TS];
};

struct FwdOps {
    bf16x8 w1[4];
    bf16x8 w2[2][4];
};
struct BwdOps {
    bf16x8 w2t[2][4];
};

__device__ inline void load_fwd(const float* __restrict__ P, FwdOps& o) {
    const int l = __lane_id(), j = l & 15, kb = 8 * (l >> 4);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 8; ++e) o.w1[t][e] = bf(P[OFF_W1 + (kb + e) * NH + 16 * t + j]);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 8; ++e) o.w2[c][t][e] = bf(P[OFF_W2 + (32 * c + kb + e) * NH + 16 * t + j]);
}
__device__ inline void load_bwd(const float* __restrict__ P, BwdOps& o) {
    const int l = __lane_id(), j = l & 15, kb = 8 * (l >> 4);
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int e = 0; e < 8; ++e) o.w2t[c][t][e] = bf(P[OFF_W2 + (16 * t + j) * NH + 32 * c + kb + e]);
}

// Forward of the tile in L.X: bf16 H1, H2 and FP32 y in LDS.
__device__ inline void tile_forward(WaveLds& L, const float* __restrict__ P, const FwdOps& o) {
    const int l = __lane_id(), j = l & 15, kb = 8 * (l >> 4), rb = 4 * (l >> 4);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        bf16x8 a;
#pragma unroll
        for (int e = 0; e < 8; ++e) a[e] = L.X[16 * s + j][kb + e];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, o.w1[t], d, 0, 0, 0);
            const float bb = P[OFF_B1 + 16 * t + j];
#pragma unroll
            for (int r = 0; r < 4; ++r) L.H1[16 * s + rb + r][16 * t + j] = bf(fmaxf(d[r] + bb, 0.0f));
        }
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        bf16x8 a[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 8; ++e) a[c][e] = L.H1[16 * s + j][32 * c + kb + e];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], o.w2[0][t], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], o.w2[1][t], d, 0, 0, 0);
            const float bb = P[OFF_B2 + 16 * t + j];
#pragma unroll
            for (int r = 0; r < 4; ++r) L.H2[16 * s + rb + r][16 * t + j] = bf(fmaxf(d[r] + bb, 0.0f));
        }
    }
    wave_sync();
    if (l < TS) {
        float acc = 0.0f;
        for (int h = 0; h < NH; ++h) acc = __builtin_fmaf(fb(L.H2[l][h]), fb(bf(P[OFF_W3 + h])), acc);
        L.y[l] = acc + P[OFF_B3];
    }
    wave_sync();
}

struct Grads {
    f32x4 w1[2][4];
    f32x4 w2[4][4];
    float w3, b1, b2, b3;  // lane l owns W3[l], b1[l], b2[l]; b3 accumulates on lane 0
};

// Backward of the tile (L.dy set). Overwrites H2 with dH2 and H1 with dH1.
__device__ inline void tile_backward(WaveLds& L, const float* __restrict__ P, const BwdOps& o, Grads& g) {
    const int l = __lane_id(), j = l & 15, kb = 8 * (l >> 4), rb = 4 * (l >> 4);
    {
        float acc = 0.0f;
        for (int i = 0; i < TS; ++i) acc = __builtin_fmaf(fb(L.H2[i][l]), L.dy[i], acc);
        g.w3 += acc;
        if (l == 0) {
            float db = 0.0f;
            for (int i = 0; i < TS; ++i) db += L.dy[i];
            g.b3 += db;
        }
    }
    wave_sync();
    const float w3l = fb(bf(P[OFF_W3 + l]));
    {
        float db = 0.0f;
        for (int i = 0; i < TS; ++i) {
            const __bf16 v = bf(fb(L.H2[i][l]) > 0.0f ? L.dy[i] * w3l : 0.0f);
            L.H2[i][l] = v;
            db += fb(v);
        }
        g.b2 += db;
    }
    wave_sync();
    // dW2 += H1^T dH2, K = samples
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        bf16x8 av;
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = L.H1[kb + e][16 * a + j];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x8 bv;
#pragma unroll
            for (int e = 0; e < 8; ++e) bv[e] = L.H2[kb + e][16 * t + j];
            g.w2[a][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, g.w2[a][t], 0, 0, 0);
        }
    }
    // dH1 = (dH2 W2^T) masked by H1 > 0
    f32x4 dh[2][4];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        bf16x8 av[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 8; ++e) av[c][e] = L.H2[16 * s + j][32 * c + kb + e];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0], o.w2t[0][t], d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1], o.w2t[1][t], d, 0, 0, 0);
            dh[s][t] = d;
        }
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                __bf16& cell = L.H1[16 * s + rb + r][16 * t + j];
                cell = bf(fb(cell) > 0.0f ? dh[s][t][r] : 0.0f);
            }
    wave_sync();
    {
        float db = 0.0f;
        for (int i = 0; i < TS; ++i) db += fb(L.H1[i][l]);
        g.b1 += db;
    }
    // dW1 += X^T dH1
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        bf16x8 av;
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = L.X[kb + e][16 * a + j];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x8 bv;
#pragma unroll
            for (int e = 0; e < 8; ++e) bv[e] = L.H1[kb + e][16 * t + j];
            g.w1[a][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, g.w1[a][t], 0, 0, 0);
        }
    }
    wave_sync();
}

// Stages samples sel[first .. first + 32) (or first.. when sel is null) of the feature matrix X[.][32].
__device__ inline void stage_tile(WaveLds& L, const float* __restrict__ X, const uint32_t* __restrict__ sel,
                                  int64_t first, int64_t last) {
    const int l = __lane_id();
    for (int q = l; q < TS * NIN; q += 64) {
        const int i = q / NIN, f = q % NIN;
        const int64_t s = first + i;
        float v = 0.0f;
        if (s < last) v = X[(int64_t)(sel ? sel[s] : (uint32_t)s) * NIN + f];
        L.X[i][f] = bf(v);
    }
    wave_sync();
}

// y[i] = model(X[sel[i]]) for i < n. One wave per 32-sample tile, grid-stride.
__global__ void __launch_bounds__(MBLK) k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                       const uint32_t* __restrict__ sel, int64_t n,
                                                       float* __restrict__ y) {
    __shared__ WaveLds lds[WPB];
    WaveLds& L = lds[threadIdx.x / 64];
    FwdOps o;
    load_fwd(P, o);
    const int l = __lane_id();
    const int64_t ntile = (n + TS - 1) / TS;
    for (int64_t t = (int64_t)blockIdx.x * WPB + threadIdx.x / 64; t < ntile; t += (int64_t)gridDim.x * WPB) {
        stage_tile(L, X, sel, t * TS, n);
        tile_forward(L, P, o);
        if (l < TS && t * TS + l < n) y[t * TS + l] = L.y[l];
        wave_sync();
    }
}

// Mean-squared-error gradient over samples [b0, b0 + bn) of sel (or of the identity order). Each wave
// writes its partial gradient to part[wave][NPARAM] and its partial loss to part_loss[wave].
__global__ void __launch_bounds__(MBLK) k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                    const float* __restrict__ target, const uint32_t* __restrict__ sel,
                                                    int64_t b0, int64_t bn, float inv_batch, float* __restrict__ part,
                                                    float* __restrict__ part_loss) {
    __shared__ WaveLds lds[WPB];
    const int wv = threadIdx.x / 64, l = __lane_id();
    const int64_t gw = (int64_t)blockIdx.x * WPB + wv;
    WaveLds& L = lds[wv];
    FwdOps fo;
    load_fwd(P, fo);
    BwdOps bo;
    load_bwd(P, bo);
    Grads g;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t) g.w1[a][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t) g.w2[a][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    g.w3 = g.b1 = g.b2 = g.b3 = 0.0f;
    float loss = 0.0f;
    const int64_t ntile = (bn + TS - 1) / TS, last = b0 + bn;
    for (int64_t t = gw; t < ntile; t += (int64_t)gridDim.x * WPB) {
        const int64_t first = b0 + t * TS;
        stage_tile(L, X, sel, first, last);
        tile_forward(L, P, fo);
        if (l < TS) {
            const int64_t s = first + l;
            float d = 0.0f;
            if (s < last) {
                const float e = L.y[l] - target[sel ? sel[s] : (uint32_t)s];
                d = 2.0f * e * inv_batch;
                loss = __builtin_fmaf(e * e, inv_batch, loss);
            }
            L.dy[l] = d;
        }
        wave_sync();
        tile_backward(L, P, bo, g);
    }
    float* out = part + gw * NPARAM;
    const int j = l & 15, rb = 4 * (l >> 4);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[OFF_W1 + (16 * a + rb + r) * NH + 16 * t + j] = g.w1[a][t][r];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[OFF_W2 + (16 * a + rb + r) * NH + 16 * t + j] = g.w2[a][t][r];
    out[OFF_W3 + l] = g.w3;
    out[OFF_B1 + l] = g.b1;
    out[OFF_B2 + l] = g.b2;
    if (l == 0) out[OFF_B3] = g.b3;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) loss += __shfl_xor(loss, d, 64);
    if (l == 0) part_loss[gw] = loss;
}

// grad = sum of the wave partials (in wave order); then, when step > 0, one Adam step on P with moments M, V.
__global__ void __launch_bounds__(MBLK) k_mlp_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                                                    const float* __restrict__ part, int nparts, float* __restrict__ grad,
                                                    int step, float lr, float beta1, float beta2, float eps) {
    const int p = blockIdx.x * MBLK + threadIdx.x;
    if (p >= NPARAM) return;
    float gs = 0.0f;
    for (int q = 0; q < nparts; ++q) gs += part[(int64_t)q * NPARAM + p];
    grad[p] = gs;
    if (step <= 0) return;
    const float m = beta1 * M[p] + (1.0f - beta1) * gs;
    const float v = beta2 * V[p] + (1.0f - beta2) * (gs * gs);
    M[p] = m;
    V[p] = v;
    const float mh = m / (1.0f - powf(beta1, (float)step));
    const float vh = v / (1.0f - powf(beta2, (float)step));
    P[p] = P[p] - lr * mh / (sqrtf(vh) + eps);
}

struct FeatParams {
    const float4* mpts;
    const unsigned long long* mvcell;  // virtual cell of each Morton position (fine cell = coordinate >> SUB)
    const uint4* htab;
    const LevelDev* lev;     // level descriptors (device)
    const CloudDev* clouds;  // per cloud: reference level, neighbour count
    const int32_t* mcloud;   // cloud of each Morton position, null for one cloud
    int64_t n;
    int B;  // fine cells per axis
};

// Feature t of the query at Morton position m (its cloud's reference level and k).
__device__ inline float feature(const FeatParams& F, int64_t m, int t) {
    if (t >= NF) return 0.0f;
    const uint32_t cloud = F.mcloud ? (uint32_t)F.mcloud[m] : 0u;
    const CloudDev& cd = F.clouds[cloud];
    if (t == NF - 1) return log2f((float)cd.k) * (1.0f / 8.0f);
    const int ref = cd.ref_level;
    const uint64_t qc = F.mvcell[m];
    const int Bl = F.B >> ref;
    const int cx = vcell_axis(qc, 0) >> (SUB + ref), cy = vcell_axis(qc, 1) >> (SUB + ref),
              cz = vcell_axis(qc, 2) >> (SUB + ref);
    const int x = cx + (t % 3) - 1, y = cy + ((t / 3) % 3) - 1, z = cz + (t / 9) - 1;
    uint32_t c = 0;
    if (x >= 0 && y >= 0 && z >= 0 && x < Bl && y < Bl && z < Bl)
        c = level_lookup(F.htab, F.lev[ref], cell_key(cloud, morton_code(x, y, z))).y;
    return log2f(1.0f + (float)c) * (1.0f / 16.0f);
}

// Feature rows for every query; the row of Morton position m goes to X[m] (by_sorted = 0) or to
// X[sorted index of m].
__global__ void __launch_bounds__(256) k_features(FeatParams F, int by_sorted, float* __restrict__ X) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t m = i / NIN;
    const int t = (int)(i % NIN);
    if (m >= F.n) return;
    const int64_t row = by_sorted ? (int64_t)__float_as_int(F.mpts[m].w) : m;
    X[row * NIN + t] = feature(F, m, t);
}

// Training targets from the plain search: log2(r_k / h_ref) per Morton position; ok = 0 where the row is
// invalid or r_k == 0.
__global__ void __launch_bounds__(256) k_targets(const float4* __restrict__ mpts, const float* __restrict__ kth,
                                                  const CloudDev* __restrict__ clouds,
                                                  const int32_t* __restrict__ mcloud, int64_t n, float ba, int B,
                                                  float* __restrict__ target, uint32_t* __restrict__ ok) {
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= n) return;
    const float d = kth[__float_as_int(mpts[m].w)];
    const bool good = d > 0.0f;
    const float inv_href = (float)(B >> clouds[mcloud ? mcloud[m] : 0].ref_level) / ba;
    ok[m] = good ? 1u : 0u;
    target[m] = good ? log2f(sqrtf(d) * inv_href) : 0.0f;
}

// Proposal: features -> MLP -> per-query start level and radius bound, fused (no feature matrix in HBM). The
// bound is h_ref * 2^(y + bound_safety), squared; the search keeps only candidates within it. The level is the
// smallest (from the finest sub-cell level -SUB up) whose 3x3x3 block around the query reaches beyond
// r = h_ref * 2^(y + safety) on every side (reach along an axis = h_l * (1 + distance of the query to the
// nearer face of its cell in cell units)).
// Key of the query's cell at its cloud's reference level (the features depend on nothing else).
__device__ inline uint64_t feature_cell(const FeatParams& F, int64_t m) {
    const uint32_t cloud = F.mcloud ? (uint32_t)F.mcloud[m] : 0u;
    const int sh = SUB + F.clouds[cloud].ref_level;
    const uint64_t qc = F.mvcell[m];
    return ((uint64_t)cloud << 30) | ((uint64_t)(vcell_axis(qc, 0) >> sh) << 20) |
           ((uint64_t)(vcell_axis(qc, 1) >> sh) << 10) | (uint64_t)(vcell_axis(qc, 2) >> sh);
}

__global__ void __launch_bounds__(MBLK) k_mlp_propose(const float* __restrict__ P, FeatParams F, float a, float ba,
                                                       int nlev, float safety, float bound_safety,
                                                       int8_t* __restrict__ qlevel, float* __restrict__ qbound) {
    __shared__ WaveLds lds[WPB];
    __shared__ int lead_s[WPB][TS];
    WaveLds& L = lds[threadIdx.x / 64];
    int* lead = lead_s[threadIdx.x / 64];
    FwdOps o;
    load_fwd(P, o);
    const int l = __lane_id();
    const int64_t ntile = (F.n + TS - 1) / TS;
    for (int64_t t = (int64_t)blockIdx.x * WPB + threadIdx.x / 64; t < ntile; t += (int64_t)gridDim.x * WPB) {
        const int64_t first = t * TS;
        // Morton-consecutive queries in one reference-level cell have the same feature row: only the first
        // of each run (its leader) does the 27 cell lookups, the others copy its row.
        {
            const int64_t m = first + l;
            const uint64_t key = (l < TS && m < F.n) ? feature_cell(F, m) : ~0ull;
            const uint64_t prev = __shfl_up(key, 1, 64);
            const uint64_t leaders = __ballot(l < TS && (l == 0 || key != prev));
            const uint64_t upto = leaders & (l == 63 ? ~0ull : ((2ull << l) - 1ull));
            if (l < TS) lead[l] = 63 - __clzll(upto);
        }
        wave_sync();
        for (int q = l; q < TS * NIN; q += 64) {
            const int i = q / NIN, f = q % NIN;
            const int64_t m = first + i;
            if (lead[i] == i) L.X[i][f] = bf(m < F.n ? feature(F, m, f) : 0.0f);
        }
        wave_sync();
        for (int q = l; q < TS * NIN; q += 64) {
            const int i = q / NIN, f = q % NIN;
            if (lead[i] != i) L.X[i][f] = L.X[lead[i]][f];
        }
        wave_sync();
        tile_forward(L, P, o);
        const int64_t m = first + l;
        if (l < TS && m < F.n) {
            const float href = ba / (float)(F.B >> F.clouds[F.mcloud ? F.mcloud[m] : 0].ref_level);
            const float r = href * exp2f(L.y[l] + safety);
            const float rb = href * exp2f(L.y[l] + bound_safety);
            qbound[m] = rb * rb;
            const float4 p = F.mpts[m];
            const float v[3] = {p.x, p.y, p.z};
            int lvl = nlev - 1;
            for (int lv = -SUB; lv < nlev; ++lv) {
                const float h = ba / (float)((F.B << SUB) >> (lv + SUB));
                float reach = 3.0f * h;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    const float u = (v[ax] - a) / h;
                    const float fr = u - floorf(u);
                    reach = fminf(reach, h * (1.0f + fminf(fr, 1.0f - fr)));
                }
                if (reach > r) {
                    lvl = lv;
                    break;
                }
            }
            qlevel[m] = (int8_t)lvl;
        }
        wave_sync();
    }
}

FeatParams feat_params(const egt_ctx* h) {
    FeatParams F{};
    F.mpts = h->mpts.p;
    F.mvcell = reinterpret_cast<const unsigned long long*>(h->mvcell.p);
    F.B = h->B;
    F.htab = h->htab.p;
    F.lev = h->dlev.p;
    F.clouds = h->dclouds.p;
    F.mcloud = h->nclouds > 1 ? h->mcloud.p : nullptr;
    F.n = h->n;
    return F;
}

// Sets every cloud's neighbour count and reference level (the start level from the mean cell occupancy)
// and uploads the cloud table. ks: one k per cloud.
void set_cloud_features(egt_ctx* h, const int* ks) {
    for (int c = 0; c < h->nclouds; ++c) {
        h->clouds[c].k = ks[c];
        h->clouds[c].ref_level = auto_start_level(h, ks[c], c);
    }
    h->dclouds.alloc(h->clouds.size());
    HIP_CHECK(hipMemcpyAsync(h->dclouds.p, h->clouds.data(), h->clouds.size() * sizeof(CloudDev),
                             hipMemcpyHostToDevice, h->stream));
}

void check_fine_grid(const egt_ctx* h) {
    if (h->B != 1024)
        throw ArgError{EGT_ERR_UNSUPPORTED,
                       "the learned proposal needs bucketCount 1024 (its features live on the 1024-cell pyramid)"};
}

int grad_waves(int64_t bn) {
    const int64_t w = std::max<int64_t>(WPB, std::min<int64_t>(MAX_GRAD_WAVES, (bn + TS - 1) / TS));
    return (int)((w + WPB - 1) / WPB * WPB);
}

// One gradient step on device samples; with step > 0 also one Adam update.
void grad_step(egt_model_s* m, const float* X, const float* T, const uint32_t* sel, int64_t b0, int64_t bn, int step,
               float lr, hipStream_t s) {
    const int nw = grad_waves(bn);
    k_mlp_grad<<<nw / WPB, MBLK, 0, s>>>(m->P.p, X, T, sel, b0, bn, 1.0f / (float)bn, m->part.p, m->ploss.p);
    LAUNCH_CHECK();
    k_mlp_adam<<<ceil_div(NPARAM, MBLK), MBLK, 0, s>>>(m->P.p, m->M.p, m->V.p, m->part.p, nw, m->grad.p, step, lr,
                                                       0.9f, 0.999f, 1e-8f);
    LAUNCH_CHECK();
    m->last_parts = nw;
}

void upload_features(egt_model_s* m, const float* Xh, const float* Th, int64_t n) {
    std::vector<float> pad((size_t)n * NIN, 0.0f);
    for (int64_t i = 0; i < n; ++i)
        for (int f = 0; f < NF; ++f) pad[(size_t)i * NIN + f] = Xh[(size_t)i * NF + f];
    m->X.alloc((size_t)n * NIN);
    m->T.alloc((size_t)n);
    HIP_CHECK(hipMemcpyAsync(m->X.p, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipMemcpyAsync(m->T.p, Th, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
}

float read_loss(egt_model_s* m) {
    std::vector<float> pl(m->last_parts);
    HIP_CHECK(hipMemcpyAsync(pl.data(), m->ploss.p, pl.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    float loss = 0.0f;
    for (float v : pl) loss += v;
    return loss;
}

}  // namespace

int mlp_param_count() { return NPARAM; }
int mlp_feature_count() { return NF; }

void mlp_init(egt_model_s* m, uint64_t seed) {
    std::vector<float> w(NPARAM, 0.0f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    auto unif = [&]() {  // xorshift64, uniform in [-1, 1)
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return 2.0 * (double)(s >> 11) * (1.0 / 9007199254740992.0) - 1.0;
    };
    for (int r = 0; r < NF; ++r)
        for (int c = 0; c < NH; ++c) w[OFF_W1 + r * NH + c] = (float)(unif() * std::sqrt(6.0 / NF));
    for (int i = 0; i < NH * NH; ++i) w[OFF_W2 + i] = (float)(unif() * std::sqrt(6.0 / NH));
    for (int i = 0; i < NH; ++i) w[OFF_W3 + i] = (float)(unif() * std::sqrt(3.0 / NH));
    HIP_CHECK(hipGetDevice(&m->device));
    if (!m->stream) HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->P.alloc(NPARAM);
    m->M.alloc(NPARAM);
    m->V.alloc(NPARAM);
    m->grad.alloc(NPARAM);
    m->part.alloc((size_t)MAX_GRAD_WAVES * NPARAM);
    m->ploss.alloc(MAX_GRAD_WAVES);
    mlp_set_params(m, w.data());
}

void mlp_set_params(egt_model_s* m, const float* params) {
    std::vector<float> w(params, params + NPARAM);
    for (int r = NF; r < NIN; ++r)
        for (int c = 0; c < NH; ++c) w[OFF_W1 + r * NH + c] = 0.0f;  // padding features carry no weight
    HIP_CHECK(hipMemcpyAsync(m->P.p, w.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HIP_CHECK(hipMemsetAsync(m->M.p, 0, NPARAM * sizeof(float), m->stream));
    HIP_CHECK(hipMemsetAsync(m->V.p, 0, NPARAM * sizeof(float), m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
    m->step = 0;
}

void mlp_get_params(egt_model_s* m, float* params) {
    HIP_CHECK(hipMemcpyAsync(params, m->P.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HIP_CHECK(hipStreamSynchronize(m->stream));
}

void mlp_eval(egt_model_s* m, const float* Xh, const float* Th, int64_t n, float* y, float* grad, float* loss) {
    upload_features(m, Xh, Th, n);
    m->Y.alloc(n);
    k_mlp_forward<<<(int)std::min<int64_t>(1024, ceil_div(n, TS * WPB)), MBLK, 0, m->stream>>>(m->P.p, m->X.p, nullptr,
                                                                                                n, m->Y.p);
    LAUNCH_CHECK();
    grad_step(m, m->X.p, m->T.p, nullptr, 0, n, 0, 0.0f, m->stream);
    if (y) HIP_CHECK(hipMemcpyAsync(y, m->Y.p, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (grad) HIP_CHECK(hipMemcpyAsync(grad, m->grad.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    const float l = read_loss(m);
    if (loss) *loss = l;
}

void mlp_adam_step(egt_model_s* m, const float* Xh, const float* Th, int64_t n, float lr, float* grad, float* loss) {
    upload_features(m, Xh, Th, n);
    m->step += 1;
    grad_step(m, m->X.p, m->T.p, nullptr, 0, n, m->step, lr, m->stream);
    if (grad) HIP_CHECK(hipMemcpyAsync(grad, m->grad.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    const float l = read_loss(m);
    if (loss) *loss = l;
}

void mlp_features(egt_ctx* h, int k, float* Xh, float* Th, int* ref_level_out) {
    check_fine_grid(h);
    if (h->search_stale) refresh_search_structure(h);
    set_cloud_features(h, &k);
    const int ref = h->clouds[0].ref_level;
    const int64_t n = h->n;
    float* X = h->scratch.get<float>(24, (size_t)n * NIN);
    k_features<<<ceil_div(n * NIN, 256), 256, 0, h->stream>>>(feat_params(h), 1, X);
    LAUNCH_CHECK();
    std::vector<float> buf((size_t)n * NIN);
    HIP_CHECK(hipMemcpyAsync(buf.data(), X, buf.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    if (Xh)
        for (int64_t i = 0; i < n; ++i)
            for (int f = 0; f < NF; ++f) Xh[(size_t)i * NF + f] = buf[(size_t)i * NIN + f];
    if (Th) {
        // targets of the handle's last search with this k, sorted order; NaN where there is none
        const float inv_href = (float)(h->B >> ref) / ((1.0f + h->eps) + h->eps);
        if (h->have_knn && h->k == k && !h->smoothed) {
            std::vector<float> kth(n);
            HIP_CHECK(hipMemcpyAsync(kth.data(), h->scratch.get<float>(15, n), n * sizeof(float),
                                     hipMemcpyDeviceToHost, h->stream));
            HIP_CHECK(hipStreamSynchronize(h->stream));
            for (int64_t i = 0; i < n; ++i) Th[i] = kth[i] > 0.0f ? std::log2(std::sqrt(kth[i]) * inv_href) : NAN;
        } else {
            for (int64_t i = 0; i < n; ++i) Th[i] = NAN;
        }
    }
    if (ref_level_out) *ref_level_out = ref;
}

void mlp_train(egt_model_s* m, egt_ctx* h, const int* ks, int epochs, int batch, int64_t max_samples, float lr,
               float quantile, float* loss_out) {
    check_fine_grid(h);
    if (h->search_stale) refresh_search_structure(h);
    hipStream_t s = h->stream;
    const int64_t n = h->n;
    // exact k-th radii from the plain search (no radius limit)
    const int saved = h->learned;
    h->learned = 0;
    for (int c = 0; c < h->nclouds; ++c) {
        h->clouds[c].k = ks[c];
        h->clouds[c].thres_n = 3.4e38f;
    }
    run_knn_clouds(h);
    if (h->nclouds == 1) h->k = ks[0];
    h->learned = saved;
    set_cloud_features(h, ks);
    const float ba = (1.0f + h->eps) + h->eps;
    float* Xall = h->scratch.get<float>(24, (size_t)n * NIN);
    float* Tall = h->scratch.get<float>(25, n);
    uint32_t* ok = h->scratch.get<uint32_t>(26, n);
    k_features<<<ceil_div(n * NIN, 256), 256, 0, s>>>(feat_params(h), 0, Xall);
    LAUNCH_CHECK();
    k_targets<<<ceil_div(n, 256), 256, 0, s>>>(h->mpts.p, h->scratch.get<float>(15, n), h->dclouds.p,
                                               h->nclouds > 1 ? h->mcloud.p : nullptr, n, ba, h->B, Tall, ok);
    LAUNCH_CHECK();
    std::vector<uint32_t> okh(n);
    HIP_CHECK(hipMemcpyAsync(okh.data(), ok, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> pool;
    pool.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        if (okh[i]) pool.push_back((uint32_t)i);
    if (pool.empty()) throw ArgError{EGT_ERR_STATE, "training cloud has no valid neighbour rows"};
    uint64_t rs = m->rng;
    auto next = [&]() {
        rs ^= rs << 13;
        rs ^= rs >> 7;
        rs ^= rs << 17;
        return rs;
    };
    auto shuffle = [&](std::vector<uint32_t>& v) {
        for (size_t i = v.size(); i > 1; --i) std::swap(v[i - 1], v[next() % i]);
    };
    shuffle(pool);
    if (max_samples > 0 && (int64_t)pool.size() > max_samples) pool.resize(max_samples);
    const int64_t ns = (int64_t)pool.size();
    m->sel.alloc(ns);
    const int64_t bsz = std::max<int64_t>(TS, std::min<int64_t>(batch, ns));
    // the model's kernels run on the cloud's stream here: finish the model stream's work first
    HIP_CHECK(hipStreamSynchronize(m->stream));
    for (int e = 0; e < epochs; ++e) {
        shuffle(pool);
        HIP_CHECK(hipMemcpyAsync(m->sel.p, pool.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        const float lr_e = (epochs > 1 && e + 1 == epochs) ? 0.3f * lr : lr;  // a shorter last epoch step
        for (int64_t b0 = 0; b0 + bsz <= ns; b0 += bsz) {  // whole batches only
            m->step += 1;
            grad_step(m, Xall, Tall, m->sel.p, b0, bsz, m->step, lr_e, s);
        }
        HIP_CHECK(hipStreamSynchronize(s));  // the host order vector is rewritten next epoch
    }
    // residuals on the training samples -> safety margin = requested quantile of (target - prediction)
    HIP_CHECK(hipMemcpyAsync(m->sel.p, pool.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    m->Y.alloc(ns);
    k_mlp_forward<<<(int)std::min<int64_t>(2048, ceil_div(ns, TS * WPB)), MBLK, 0, s>>>(m->P.p, Xall, m->sel.p, ns,
                                                                                         m->Y.p);
    LAUNCH_CHECK();
    std::vector<float> yh(ns), th(n);
    HIP_CHECK(hipMemcpyAsync(yh.data(), m->Y.p, ns * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(th.data(), Tall, n * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<float> res(ns);
    double mse = 0.0;
    for (int64_t i = 0; i < ns; ++i) {
        res[i] = th[pool[i]] - yh[i];
        mse += (double)res[i] * res[i];
    }
    const float qn = std::min(1.0f, std::max(0.0f, quantile));
    const size_t qi = std::min<size_t>(ns - 1, (size_t)(qn * (float)(ns - 1)));
    std::nth_element(res.begin(), res.begin() + qi, res.end());
    m->safety = res[qi];
    // radius bound: the 0.99 quantile of the residuals plus 0.6 (log2; the radius x 1.5). A query beyond its
    // bound is searched twice; with the margin about 0.05 % of the queries are (16M-point cluster mixture: 7,272
    // at +0.6 against 34,291 at +0.3 and 239,239 at +0; 4M-point torus: 2,642, 62,822 and 681,801), while the
    // bound still keeps most far candidates out of the lists (k=40 mixture search 54.2 ms against 59.0 without
    // the bound).
    const size_t qb = std::min<size_t>(ns - 1, (size_t)(0.99 * (double)(ns - 1)));
    std::nth_element(res.begin(), res.begin() + qb, res.end());
    m->bound_safety = std::max(m->safety, res[qb]) + 0.6f;
    m->rng = rs;
    m->train_samples = ns;
    if (loss_out) *loss_out = (float)(mse / (double)ns);
}

void mlp_propose(egt_model_s* m, egt_ctx* h, const int* ks) {
    check_fine_grid(h);
    if (h->search_stale) refresh_search_structure(h);
    const int64_t n = h->n;
    hipStream_t s = h->stream;
    h->qlevel.alloc(n);
    h->qbound.alloc(n);
    set_cloud_features(h, ks);
    const float a = -h->eps, ba = (1.0f + h->eps) - a;
    HIP_CHECK(hipStreamSynchronize(m->stream));  // parameters written on the model stream are final
    Event e0, e1;
    HIP_CHECK(hipEventRecord(e0, s));
    const int64_t tiles = (n + TS - 1) / TS;
    k_mlp_propose<<<(int)std::min<int64_t>(8192, (tiles + WPB - 1) / WPB), MBLK, 0, s>>>(
        m->P.p, feat_params(h), a, ba, h->nlev, m->safety, m->bound_safety, h->qlevel.p, h->qbound.p);
    LAUNCH_CHECK();
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&h->stats.propose_ms, e0, e1));
    h->learned = 1;
}

void mlp_levels(egt_ctx* h, int8_t* out) {
    if (!h->qlevel.p || h->qlevel.n < (size_t)h->n) throw ArgError{EGT_ERR_STATE, "no proposal on this handle"};
    std::vector<int8_t> q(h->n);
    std::vector<int32_t> mo(h->n);
    HIP_CHECK(hipMemcpyAsync(q.data(), h->qlevel.p, h->n, hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(hipMemcpyAsync(mo.data(), h->morder.p, h->n * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_CHECK(hipStreamSynchronize(h->stream));
    for (int64_t m = 0; m < h->n; ++m) out[mo[m]] = q[m];
}

}  // namespace egt
