// Device helpers shared by the gfx950 kernels.
//
// Rounding contract: every float operation is rounded as written in the reference source except these
// fused forms (which nvcc also contracts by default in the reference build):
//   squared distance  sq(p, q) = fmaf(dz, dz, fmaf(dy, dy, dx * dx)), d* = p* - q*
//   EGG frame row     fmaf(a3, v3, fmaf(a2, v2, a1 * v1)) + t
//   accumulations     acc = fmaf(w, v, acc)
//   kNN face distance fmaf(-(float)c, h, x - a), fmaf((float)c, h, a - x)
// All HIP sources build with -ffp-contract=off and correctly rounded fp32 division / sqrt, so nothing
// else fuses.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace egt {

__host__ __device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ inline uint32_t morton_code(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

// Sub-cell levels of the kNN search: each fine cell of the reference grid is cut into 2^SUB cells per axis,
// so search boxes can be smaller than a cell. A point's virtual cell (per axis: fine cell << SUB | its
// position inside the fine cell) is packed with the Morton code of that position as
// x | y << 16 | z << 32 | sub code << 48. Inside a fine cell, points are in sub-code order, so every
// sub-cell at every sub level is a contiguous range.
constexpr int SUB = 3;
__host__ __device__ inline uint64_t vcell_pack(uint32_t vx, uint32_t vy, uint32_t vz) {
    constexpr uint32_t m = (1u << SUB) - 1u;
    return (uint64_t)vx | ((uint64_t)vy << 16) | ((uint64_t)vz << 32) |
           ((uint64_t)morton_code(vx & m, vy & m, vz & m) << 48);
}
__host__ __device__ inline int vcell_axis(uint64_t v, int ax) { return (int)((v >> (16 * ax)) & 0xFFFFu); }
__host__ __device__ inline uint32_t vcell_sub(uint64_t v) { return (uint32_t)(v >> 48); }

__device__ inline float sqd(float px, float py, float pz, float qx, float qy, float qz) {
    const float dx = px - qx, dy = py - qy, dz = pz - qz;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ inline int cell_of(float v, float a, float ba, int B) { return (int)(((v - a) / ba) * (float)B); }
// Virtual coordinate (resolution 2^SUB B) of normalized coordinate v whose fine cell is c. Scaling by 2^SUB is
// exact, so for a point whose cell is cell_of(v) the result lies inside cell c; the clamp keeps it there for
// cells assigned otherwise (queries clamped to the grid).
__device__ inline uint32_t vcoord(float v, float a, float ba, int B, int c) {
    const int t = (int)(((v - a) / ba) * (float)(B << SUB));
    return (uint32_t)min(max(t, c << SUB), (c << SUB) + (1 << SUB) - 1);
}

// exp(x) near 0: for |x| <= 2^-6 the cubic Taylor polynomial in Horner form (truncation error
// x^4/24 <= 2.5e-9, a small fraction of an ulp of the result).
__device__ inline bool egt_expf_small(float x) { return x >= -0.015625f && x <= 0.015625f; }
__device__ inline float egt_expf_cubic(float x) {
    const float p = __builtin_fmaf(x, 0.166666672f, 0.5f);
    return __builtin_fmaf(__builtin_fmaf(p, x, 1.0f), x, 1.0f);
}

// exp(x): the cubic near 0, else Cody-Waite reduction x = n*ln2 + r (ln2 split in a 24-bit head and a
// tail), degree-7 Taylor polynomial of exp(r) in Horner form, exact scaling by 2^n.
__device__ inline float egt_expf(float x) {
    if (egt_expf_small(x)) return egt_expf_cubic(x);
    if (!(x >= -104.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.0f) return __builtin_huge_valf();
    const float n = __builtin_rintf(x * 1.44269502163f);
    float r = __builtin_fmaf(-n, 0.693145751953125f, x);
    r = __builtin_fmaf(-n, 1.428606765330187e-06f, r);
    float p = 1.98412698e-4f;
    p = __builtin_fmaf(p, r, 1.38888889e-3f);
    p = __builtin_fmaf(p, r, 8.33333333e-3f);
    p = __builtin_fmaf(p, r, 4.16666667e-2f);
    p = __builtin_fmaf(p, r, 1.66666667e-1f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const int ni = (int)n;
    if (ni < -125) {
        const float s1 = __uint_as_float((uint32_t)(ni + 64 + 127) << 23);
        const float s2 = __uint_as_float((uint32_t)(-64 + 127) << 23);
        return (p * s1) * s2;
    }
    return p * __uint_as_float((uint32_t)(ni + 127) << 23);
}

// egt_expf for the Taubin weight argument x = -|q - p|^2 <= 0: the same cubic near 0; on [-80, 0] the
// range checks of egt_expf never fire and 2^n is a normal number, so the same reduction and polynomial
// followed by an exact ldexp give identical results; other arguments go through egt_expf. Neighbour
// distances are almost always in the cubic's range, so the cubic is computed unconditionally and the
// wide path is a branch that a wave skips when none of its lanes needs it.
__device__ inline float egt_expf_wide_nonpos(float x) {
    if (!(x >= -80.0f)) return egt_expf(x);
    const float n = __builtin_rintf(x * 1.44269502163f);
    float r = __builtin_fmaf(-n, 0.693145751953125f, x);
    r = __builtin_fmaf(-n, 1.428606765330187e-06f, r);
    float p = 1.98412698e-4f;
    p = __builtin_fmaf(p, r, 1.38888889e-3f);
    p = __builtin_fmaf(p, r, 8.33333333e-3f);
    p = __builtin_fmaf(p, r, 4.16666667e-2f);
    p = __builtin_fmaf(p, r, 1.66666667e-1f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_ldexpf(p, (int)n);
}
__device__ inline float egt_expf_taubin(float x) {
    float w = egt_expf_cubic(x);
    if (!egt_expf_small(x)) w = egt_expf_wide_nonpos(x);
    return w;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Golub & Van Loan sec. 8.5); returns the
// index of the smallest eigenvalue and its eigenvector (vx, vy, vz).
__device__ inline void eig3_smallest(float a00, float a01, float a02, float a11, float a12, float a22, float& vx,
                                     float& vy, float& vz) {
    float m[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
    float V[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    for (int sweep = 0; sweep < 8; ++sweep) {
        if (m[0][1] == 0.0f && m[0][2] == 0.0f && m[1][2] == 0.0f) break;
#pragma unroll
        for (int rot = 0; rot < 3; ++rot) {
            const int p = rot == 2 ? 1 : 0, q = rot == 0 ? 1 : 2, r = 3 - p - q;
            const float apq = m[p][q];
            if (apq == 0.0f) continue;
            const float app = m[p][p], aqq = m[q][q];
            const float theta = (aqq - app) / (2.0f * apq);
            float t = 1.0f / (__builtin_fabsf(theta) + __builtin_sqrtf(theta * theta + 1.0f));
            if (theta < 0.0f) t = -t;
            const float c = 1.0f / __builtin_sqrtf(t * t + 1.0f);
            const float s = t * c;
            m[p][p] = app - t * apq;
            m[q][q] = aqq + t * apq;
            m[p][q] = m[q][p] = 0.0f;
            const float arp = m[r][p], arq = m[r][q];
            m[r][p] = m[p][r] = c * arp - s * arq;
            m[r][q] = m[q][r] = s * arp + c * arq;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float vip = V[i][p], viq = V[i][q];
                V[i][p] = c * vip - s * viq;
                V[i][q] = s * vip + c * viq;
            }
        }
    }
    int idx = 0;
    if (m[1][1] < m[idx][idx]) idx = 1;
    if (m[2][2] < m[idx][idx]) idx = 2;
    vx = V[0][idx];
    vy = V[1][idx];
    vz = V[2][idx];
}

// Cell-table key: cloud index in the high word, level Morton code of the cell in the low word (one table
// per level holds the cells of every cloud of a batch; a single cloud is cloud 0).
__host__ __device__ inline uint64_t cell_key(uint32_t cloud, uint32_t code) {
    return ((uint64_t)cloud << 32) | code;
}

// Fibonacci hashing of a 64-bit key into a power-of-two table.
__host__ __device__ inline uint32_t hash_slot(uint64_t key, uint32_t log2cap) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64u - log2cap));
}
constexpr uint64_t HASH_EMPTY = ~0ull;

struct LevelDev {
    uint64_t offset;
    uint32_t log2cap;
    int32_t max_count;
};

// Level tables: one 16-byte entry per slot, key (low, high word) then (start, count), so a probe reads key and
// range in one load and an insert writes one cache line.
// Looks up a level cell; returns (start, count) or count 0 when empty.
__device__ inline uint2 level_lookup(const uint4* __restrict__ tab, const LevelDev& L, uint64_t key) {
    const uint32_t mask = (1u << L.log2cap) - 1u;
    uint32_t s = hash_slot(key, L.log2cap);
    for (uint32_t probe = 0; probe <= mask; ++probe) {  // tables are at most half full
        const uint4 e = tab[L.offset + s];
        const uint64_t kk = (uint64_t)e.x | ((uint64_t)e.y << 32);
        if (kk == key) return make_uint2(e.z, e.w);
        if (kk == HASH_EMPTY) break;
        s = (s + 1u) & mask;
    }
    return make_uint2(0u, 0u);
}

}  // namespace egt
