/*
 * ORACLE -- test infrastructure only.
 *
 * Host C++17 + OpenMP restatement of the grid-kNN -> elliptic-Gabriel-graph -> Taubin -> normals path of
 * agalex1974/MultiGPUEGTSmoothing. It is the checker for the HIP library and the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product library
 * (libegt.so) never links, calls or includes anything from this directory.
 *
 * Reference citations are file:line into the reference repository.
 *
 * Rounding contract (the HIP kernels implement the same sequence independently): every float operation
 * is rounded as written in the reference source, left to right, except these fused forms. The set is this
 * contract's own choice: the reference's nvcc build contracts by default and may fuse more (for example the
 * EGG cross product, EGT_CUDA.cu:97-104, and the ellipsoid value's leading term, :165, which stay unfused
 * here), so bit parity with the reference binary is unpinned:
 *   sq(p, q)         = fmaf(dz, dz, fmaf(dy, dy, dx * dx)), d* = p* - q*   (squared distances, norms)
 *   row(a, v) + t    = fmaf(a3, v3, fmaf(a2, v2, a1 * v1)) + t               (EGG frame transform)
 *   acc += w * v     -> acc = fmaf(w, v, acc)                                (Taubin / PCA sums)
 *   face distances   = fmaf(-(float)c, h, x - a) and fmaf((float)c, h, a - x) (kNN certificate)
 * exp() in Taubin is egt_expf below (cubic Taylor near 0, else Cody-Waite reduction + degree-7 Taylor), so the FP32
 * Taubin is reproducible bit for bit on any IEEE machine.
 *
 * Pinning (tests/test_oracle_*.py):
 *   orc_grid_build      -- pinned: exact integer/permutation checks against an independent numpy build
 *   orc_knn_self        -- pinned: brute-force O(N^2) kNN on small clouds, duplicate/lattice/isolated cases,
 *                          big-cell and N<=k invalidation cases built by hand
 *   orc_knn_query       -- pinned: brute-force query kNN (strict radius, no self exclusion) on small clouds,
 *                          queries inside, on the edge of and far outside the grid, duplicates of reference
 *                          points, small radii (short and empty rows), big-cell invalidation; the literal
 *                          reference schedule agrees row for row
 *   orc_egg             -- pinned: closed forms (collinear drop, unit-square alpha>1 rule, alpha->0 keeps,
 *                          zero-distance neighbour kept, EGG subset of kNN, monotone in alpha) with an FP64
 *                          predicate (also on a cloud with duplicated points)
 *   orc_taubin_f32/f64  -- pinned: lambda=mu=0 identity, symmetric-neighbour fixed point, planar stays planar,
 *                          FP32 vs FP64 bound, closed-form ring shrink; regularized weights (isRegularized,
 *                          EGT_CUDA.cu:263): ring closed form over both passes, two-neighbour closed form of
 *                          the inverse-square weights, FP32 vs FP64 bound
 *   egt_expf            -- pinned: <=2 ulp against double exp over [-90, 0]
 *   orc_eig3            -- pinned: numpy.linalg.eigh on random symmetric matrices (angle/eigenvalue bounds)
 *   orc_pca_normals     -- pinned: plane -> plane normal, sphere -> radial, a hand-built 5-point case whose
 *                          normal flips axis if the point itself enters the covariance
 *   orc_orient_normals  -- pinned: MSF weight equals scipy minimum_spanning_tree; per-component sign
 *                          consistency on sphere / two-sphere clouds
 *   orc_mesh_adjacency  -- pinned: icosphere one-ring valence 5/6 and brute-force corner union
 *   orc_output          -- pinned: round trip without smoothing within 1 ulp of the extent
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>

// parallel_for in the style of the reference's parallel.h (OpenMP branch).
#define parallel_for _Pragma("omp parallel for schedule(dynamic, 256)") for

namespace {

inline float sqf(float px, float py, float pz, float qx, float qy, float qz) {
    const float dx = px - qx, dy = py - qy, dz = pz - qz;
    return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

inline int cell_of(float v, float a, float ba, int B) { return (int)(((v - a) / ba) * (float)B); }

struct Entry {
    float d;
    int32_t j;
};
inline bool entry_less(const Entry& l, const Entry& r) { return l.d < r.d || (l.d == r.d && l.j < r.j); }

// Sorted top-k list with the reference's duplicate check (GridKNNCuda.cu:311-339), ordered by
// (distance, sorted index) instead of visit order.
struct TopK {
    int k;
    std::vector<Entry> e;
    explicit TopK(int k_) : k(k_) { e.reserve(k_ + 1); }
    void offer(float d, int32_t j) {
        Entry c{d, j};
        if ((int)e.size() == k && !entry_less(c, e.back())) return;
        auto it = std::lower_bound(e.begin(), e.end(), c, entry_less);
        if (it != e.end() && it->j == j && it->d == d) return;  // already listed
        e.insert(it, c);
        if ((int)e.size() > k) e.pop_back();
    }
};

// One level of the grid pyramid: GridKNNCuda.cu:898-955 (stable sort of sorted-space points by the
// row-major key at bucket count B, unique keys, first position of each key).
struct Level {
    int B = 0;
    std::vector<int32_t> order;
    std::vector<int64_t> ukeys;
    std::vector<int32_t> start;  // size ukeys.size()+1
    int32_t max_count = 0;
    void build(const float* xs, const float* ys, const float* zs, int64_t n, float a, float ba, int B_) {
        B = B_;
        std::vector<int64_t> key(n);
        for (int64_t i = 0; i < n; ++i) {
            const int cx = cell_of(xs[i], a, ba, B), cy = cell_of(ys[i], a, ba, B), cz = cell_of(zs[i], a, ba, B);
            key[i] = (int64_t)cx + (int64_t)cy * B + (int64_t)cz * B * B;
        }
        order.resize(n);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t l, int32_t r) { return key[l] < key[r]; });
        ukeys.clear();
        start.clear();
        for (int64_t p = 0; p < n; ++p) {
            const int64_t kk = key[order[p]];
            if (p == 0 || kk != ukeys.back()) {
                ukeys.push_back(kk);
                start.push_back((int32_t)p);
            }
        }
        start.push_back((int32_t)n);
        max_count = 0;
        for (size_t u = 0; u + 1 < start.size(); ++u) max_count = std::max(max_count, start[u + 1] - start[u]);
    }
    // binarySearchInt GridKNNCuda.cu:33-50 over the unique keys; returns count, sets [s,e)
    int32_t lookup(int cx, int cy, int cz, int32_t& s, int32_t& e) const {
        s = e = 0;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= B || cy >= B || cz >= B) return -1;
        const int64_t kk = (int64_t)cx + (int64_t)cy * B + (int64_t)cz * B * B;
        auto it = std::lower_bound(ukeys.begin(), ukeys.end(), kk);
        if (it == ukeys.end() || *it != kk) return 0;
        const size_t u = it - ukeys.begin();
        s = start[u];
        e = start[u + 1];
        return e - s;
    }
};

struct GridGeom {
    float a, b, ba;
    int B;
};

// Reference stop certificate, GridKNNCuda.cu:694-741, evaluated at stage s for a query whose exact list
// has c entries (kth = largest distance) of which all lie inside the stage box. Returns true when the
// reference would keep searching. Stage 0 is the own cell (offset 0 at level 0); stage s>=1 is ring 1
// at level s-1.
bool ref_continue(const GridGeom& g, int s, float x, float y, float z, int fx, int fy, int fz, int c, int k,
                  float kth, float thres_n) {
    const int lvl = s == 0 ? 0 : s - 1;
    const int off = s == 0 ? 0 : 1;
    const int Bl = g.B >> lvl;
    const float h = g.ba / (float)Bl;  // (max - min) / bucketsCount, GridKNNCuda.cu:2405,2468
    const int xi = fx >> lvl, yi = fy >> lvl, zi = fz >> lvl;
    const float coords[3] = {x, y, z};
    const int cells[3] = {xi, yi, zi};
    if (c < k) {
        const float r = std::sqrt(thres_n);
        for (int ax = 0; ax < 3; ++ax) {
            const float v = coords[ax];
            const int ci = cells[ax];
            if (ci - off >= 0 && r > std::fmaf(-(float)(ci - off), h, v - g.a)) return true;
            if (ci + off + 1 <= Bl && r > std::fmaf((float)(ci + off + 1), h, g.a - v)) return true;
        }
        return false;
    }
    const float r = std::sqrt(kth);
    for (int ax = 0; ax < 3; ++ax) {
        const float v = coords[ax];
        const int ci = cells[ax];
        if (ci - off >= 0 && r >= std::fmaf(-(float)(ci - off), h, v - g.a)) return true;
        if (ci + off + 1 <= Bl && r >= std::fmaf((float)(ci + off + 1), h, g.a - v)) return true;
    }
    return false;
}

int num_ring_stages(int B) {
    // ring stages run at B, B/2, ..., 8 (GridKNNCuda.cu:2464-2572); at least one ring.
    int s = 0;
    for (int b = B; b >= 8; b >>= 1) ++s;
    return std::max(s, 1);
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

int orc_num_threads() { return omp_get_max_threads(); }

/* egt_expf: exp(x) for the Taubin weights. For |x| <= 2^-6 (the usual case: neighbour distances are
 * small in the normalized cube) the cubic Taylor polynomial 1 + x + x^2/2 + x^3/6 in Horner form, whose
 * truncation error x^4/24 <= 2.5e-9 is a small fraction of an ulp of the result. Otherwise Cody-Waite
 * argument reduction x = n*ln2 + r with ln2 split into a 24-bit head and a tail, |r| <= ln2/2, degree-7
 * Taylor polynomial of exp(r) in Horner form, and exact scaling by 2^n. */
float egt_expf(float x) {
    if (x >= -0.015625f && x <= 0.015625f) {
        const float p = std::fmaf(x, 0.166666672f, 0.5f);
        return std::fmaf(std::fmaf(p, x, 1.0f), x, 1.0f);
    }
    if (!(x >= -104.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.0f) return INFINITY;
    const float n = std::rint(x * 1.44269502163f);
    float r = std::fmaf(-n, 0.693145751953125f, x);
    r = std::fmaf(-n, 1.428606765330187e-06f, r);
    float p = 1.98412698e-4f;             // 1/5040
    p = std::fmaf(p, r, 1.38888889e-3f);  // 1/720
    p = std::fmaf(p, r, 8.33333333e-3f);  // 1/120
    p = std::fmaf(p, r, 4.16666667e-2f);  // 1/24
    p = std::fmaf(p, r, 1.66666667e-1f);  // 1/6
    p = std::fmaf(p, r, 0.5f);
    p = std::fmaf(p, r, 1.0f);
    p = std::fmaf(p, r, 1.0f);
    int ni = (int)n;
    if (ni < -125) {
        uint32_t b1 = (uint32_t)(ni + 64 + 127) << 23, b2 = (uint32_t)(-64 + 127) << 23;
        float s1, s2;
        std::memcpy(&s1, &b1, 4);
        std::memcpy(&s2, &b2, 4);
        return (p * s1) * s2;
    }
    uint32_t bits = (uint32_t)(ni + 127) << 23;
    float s;
    std::memcpy(&s, &bits, 4);
    return p * s;
}

/* Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (Jacobi's method; Golub & Van Loan,
 * Matrix Computations, sec. 8.5). Replaces the reference's Eberly solver (EigenSolver.cuh:124-393): the
 * normal is the eigenvector of the smallest eigenvalue in both. a6 = {a00,a01,a02,a11,a12,a22};
 * evecs column-major (evecs[3*c + r] = component r of eigenvector c). Returns index of the smallest. */
int orc_eig3(const float* a6, float* evals, float* evecs) {
    float m[3][3] = {{a6[0], a6[1], a6[2]}, {a6[1], a6[3], a6[4]}, {a6[2], a6[4], a6[5]}};
    float V[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 8; ++sweep) {
        if (m[0][1] == 0.0f && m[0][2] == 0.0f && m[1][2] == 0.0f) break;
        for (int rot = 0; rot < 3; ++rot) {
            const int p = P[rot], q = Q[rot], r = 3 - p - q;
            const float apq = m[p][q];
            if (apq == 0.0f) continue;
            const float app = m[p][p], aqq = m[q][q];
            const float theta = (aqq - app) / (2.0f * apq);
            float t = 1.0f / (std::fabs(theta) + std::sqrt(theta * theta + 1.0f));
            if (theta < 0.0f) t = -t;
            const float c = 1.0f / std::sqrt(t * t + 1.0f);
            const float s = t * c;
            m[p][p] = app - t * apq;
            m[q][q] = aqq + t * apq;
            m[p][q] = m[q][p] = 0.0f;
            const float arp = m[r][p], arq = m[r][q];
            m[r][p] = m[p][r] = c * arp - s * arq;
            m[r][q] = m[q][r] = s * arp + c * arq;
            for (int i = 0; i < 3; ++i) {
                const float vip = V[i][p], viq = V[i][q];
                V[i][p] = c * vip - s * viq;
                V[i][q] = s * vip + c * viq;
            }
        }
    }
    evals[0] = m[0][0];
    evals[1] = m[1][1];
    evals[2] = m[2][2];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) evecs[3 * c + r] = V[r][c];
    int idx = 0;
    if (evals[1] < evals[idx]) idx = 1;
    if (evals[2] < evals[idx]) idx = 2;
    return idx;
}

/* a1-a2: bounding cube + normalization + fine cell key + stable sort.
 * GridKNNCuda.cu:1042-1077 (min/max folds from +-1e8f, padded by 1e-6f, one cube for all axes),
 * :840-849 (normalize), :74-88 (key = xi + yi*B + zi*B*B), :898-936 (stable sort of (key, index)),
 * :1107-1128 (points gathered into sorted order; perm[sorted] = original index).
 * extent2 = {minExtent, maxExtent}. Returns 0 or a negative error. */
int orc_grid_build(const float* x, const float* y, const float* z, int64_t n, float eps, int B, float* xs,
                   float* ys, float* zs, int32_t* perm, int32_t* key, float* extent2) {
    if (n < 1 || B < 1) return -1;
    float minx = 1e8f, maxx = -1e8f, miny = 1e8f, maxy = -1e8f, minz = 1e8f, maxz = -1e8f;
    for (int64_t i = 0; i < n; ++i) {
        minx = std::min(minx, x[i]);
        maxx = std::max(maxx, x[i]);
        miny = std::min(miny, y[i]);
        maxy = std::max(maxy, y[i]);
        minz = std::min(minz, z[i]);
        maxz = std::max(maxz, z[i]);
    }
    minx -= 1e-6f; miny -= 1e-6f; minz -= 1e-6f;
    maxx += 1e-6f; maxy += 1e-6f; maxz += 1e-6f;
    const float mn = std::min(std::min(minx, miny), minz);
    const float mx = std::max(std::max(maxx, maxy), maxz);
    const float d = mx - mn;
    std::vector<float> nx(n), ny(n), nz(n);
    std::vector<int32_t> k0(n);
    const float a = -eps, b = 1.0f + eps, ba = b - a;
    for (int64_t i = 0; i < n; ++i) {
        nx[i] = (x[i] - mn) / d;
        ny[i] = (y[i] - mn) / d;
        nz[i] = (z[i] - mn) / d;
        const int cx = cell_of(nx[i], a, ba, B), cy = cell_of(ny[i], a, ba, B), cz = cell_of(nz[i], a, ba, B);
        k0[i] = cx + cy * B + cz * B * B;
    }
    std::vector<int32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t l, int32_t r) { return k0[l] < k0[r]; });
    for (int64_t s = 0; s < n; ++s) {
        const int32_t o = idx[s];
        perm[s] = o;
        xs[s] = nx[o];
        ys[s] = ny[o];
        zs[s] = nz[o];
        key[s] = k0[o];
    }
    extent2[0] = mn;
    extent2[1] = mx;
    return 0;
}

/* a3-a5: exact self kNN in sorted space with the reference's invalidation.
 * Search (GridKNNCuda.cu:2322-2634): own cell at level 0, then ring 1 at levels 0,1,2,... (bucket count
 * halved per level, GridKNNCuda.cu:2465-2474), scanning each cell of the shell (construct_neighbors_self_
 * compact :647-743), inserting candidates with d <= thres, self excluded, duplicates skipped.
 * Rows are ordered by (distance, sorted index) -- the reference orders ties by visit order (FindPosition
 * :126-143). The search continues until the list is provably exact (k-th distance plus a 2e-6 margin below
 * every interior face distance, computed in double, or the box covers the grid), so rows are the exact
 * k nearest neighbours; the reference's own float certificate then decides validity:
 *   stage s = 0 .. S (S = ring stages down to bucket count 8):
 *     a visited cell with >= maxc points -> row invalid (:672, :685)
 *     stop (row valid) when all listed neighbours lie in the stage box and the reference certificate
 *     (:694-741) says stop; if still searching after stage S -> row invalid (:2575-2586)
 *   invalid rows are all -1 (negateNotValidKnnIndexes :298-309).
 * knn: n*k int32 (sorted space, -1 padding); kth: n floats (k-th distance of the exact list or -1);
 * invalid: n flags. stats[0] = invalid rows, stats[1] = rows where the literal reference schedule (list at
 * its own stop stage) differs from this result, stats[2] = total candidate distance evaluations. */
int orc_knn_self_subset(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, float eps,
                        int B, int k, float thres_n, int maxc, const int32_t* qsel, int64_t nsel, int32_t* knn,
                        float* kth_out, int32_t* invalid, int64_t* stats);

int orc_knn_self(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, float eps, int B,
                 int k, float thres_n, int maxc, int32_t* knn, float* kth_out, int32_t* invalid, int64_t* stats) {
    return orc_knn_self_subset(xs, ys, zs, key, n, eps, B, k, thres_n, maxc, nullptr, n, knn, kth_out, invalid,
                               stats);
}

/* The same search for the queries qsel[0..nsel) only (sorted indices; null = every point, nsel = n): output
 * row t belongs to query qsel[t] (timing samples of large clouds). */
int orc_knn_self_subset(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, float eps,
                        int B, int k, float thres_n, int maxc, const int32_t* qsel, int64_t nsel, int32_t* knn,
                        float* kth_out, int32_t* invalid, int64_t* stats) {
    if (n < 1 || k < 1 || B < 1 || nsel < 0 || (!qsel && nsel != n)) return -1;
    GridGeom g;
    g.a = -eps;
    g.b = 1.0f + eps;
    g.ba = g.b - g.a;
    g.B = B;
    int nlev = 0;
    for (int b = B; b >= 1; b >>= 1) ++nlev;
    std::vector<Level> lev(nlev);
    for (int l = 0; l < nlev; ++l) lev[l].build(xs, ys, zs, n, g.a, g.ba, B >> l);
    const int S = num_ring_stages(B);
    int64_t n_invalid = 0, n_literal_diff = 0, n_evals = 0;
    const double margin = 2e-6;

#pragma omp parallel for schedule(dynamic, 256) reduction(+ : n_invalid, n_literal_diff, n_evals)
    for (int64_t t = 0; t < nsel; ++t) {
        const int64_t q = qsel ? (int64_t)qsel[t] : t;
        const float x = xs[q], y = ys[q], z = zs[q];
        const int fx = key[q] % B, fy = (key[q] / B) % B, fz = key[q] / (B * B);
        TopK L(k);
        // literal reference schedule bookkeeping
        bool lit_done = false, lit_invalid = false;
        std::vector<Entry> lit_list;
        for (int s = 0;; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            if (lvl >= nlev) break;
            const Level& Lv = lev[lvl];
            const int off = s == 0 ? 0 : 1;
            const int cx = fx >> lvl, cy = fy >> lvl, cz = fz >> lvl;
            bool big = false;
            for (int dz = -off; dz <= off; ++dz)
                for (int dy = -off; dy <= off; ++dy)
                    for (int dx = -off; dx <= off; ++dx) {
                        if (s > 0 && dx == 0 && dy == 0 && dz == 0) continue;
                        int32_t st, en;
                        const int32_t cnt = Lv.lookup(cx + dx, cy + dy, cz + dz, st, en);
                        if (cnt <= 0) continue;
                        if (cnt >= maxc) big = true;
                        for (int32_t p = st; p < en; ++p) {
                            const int32_t j = Lv.order[p];
                            if (j == q) continue;
                            const float d = sqf(x, y, z, xs[j], ys[j], zs[j]);
                            ++n_evals;
                            if (d <= thres_n) L.offer(d, j);
                        }
                    }
            if (!lit_done && s <= S) {
                if (big) {
                    lit_done = true;
                    lit_invalid = true;
                } else {
                    const int c = (int)L.e.size();
                    const float kth = c > 0 ? L.e.back().d : 0.0f;
                    if (!ref_continue(g, s, x, y, z, fx, fy, fz, c, k, kth, thres_n)) {
                        lit_done = true;
                        lit_list = L.e;
                    } else if (s == S) {
                        lit_done = true;
                        lit_invalid = true;
                    }
                }
            }
            // exactness (double, with margin)
            const int Bl = B >> lvl;
            const double hl = (double)g.ba / (double)Bl;
            const double bound = (int)L.e.size() == k ? (double)L.e.back().d : (double)thres_n;
            const double R = std::sqrt(bound) + margin;
            bool exact = true;
            const double qc[3] = {x, y, z};
            const int qi[3] = {cx, cy, cz};
            for (int ax = 0; ax < 3 && exact; ++ax) {
                if (qi[ax] - off >= 1) {
                    const double face = (double)g.a + (double)(qi[ax] - off) * hl;
                    if (!(R < qc[ax] - face)) exact = false;
                }
                if (qi[ax] + off <= Bl - 2) {
                    const double face = (double)g.a + (double)(qi[ax] + off + 1) * hl;
                    if (!(R < face - qc[ax])) exact = false;
                }
            }
            if (exact && lit_done) break;
        }
        // reference validity from the exact list
        const int c = (int)L.e.size();
        const float kth = c > 0 ? L.e.back().d : 0.0f;
        int lox = 0, hix = 0, loy = 0, hiy = 0, loz = 0, hiz = 0;
        if (c > 0) {
            lox = loy = loz = B;
            hix = hiy = hiz = -1;
            for (const Entry& e : L.e) {
                const int32_t kk = key[e.j];
                const int mx_ = kk % B, my_ = (kk / B) % B, mz_ = kk / (B * B);
                lox = std::min(lox, mx_); hix = std::max(hix, mx_);
                loy = std::min(loy, my_); hiy = std::max(hiy, my_);
                loz = std::min(loz, mz_); hiz = std::max(hiz, mz_);
            }
        }
        bool valid = false, decided = false;
        for (int s = 0; s <= S && !decided; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            const int off = s == 0 ? 0 : 1;
            const Level& Lv = lev[lvl];
            const int cx = fx >> lvl, cy = fy >> lvl, cz = fz >> lvl;
            for (int dz = -off; dz <= off && !decided; ++dz)
                for (int dy = -off; dy <= off && !decided; ++dy)
                    for (int dx = -off; dx <= off; ++dx) {
                        if (s > 0 && dx == 0 && dy == 0 && dz == 0) continue;
                        int32_t st, en;
                        if (Lv.lookup(cx + dx, cy + dy, cz + dz, st, en) >= maxc) {
                            decided = true;
                            valid = false;
                            break;
                        }
                    }
            if (decided) break;
            bool inside = true;
            if (c > 0) {
                inside = (lox >> lvl) >= cx - off && (hix >> lvl) <= cx + off && (loy >> lvl) >= cy - off &&
                         (hiy >> lvl) <= cy + off && (loz >> lvl) >= cz - off && (hiz >> lvl) <= cz + off;
            }
            if (inside && !ref_continue(g, s, x, y, z, fx, fy, fz, c, k, kth, thres_n)) {
                decided = true;
                valid = true;
            }
        }
        int32_t* row = knn + (size_t)t * k;
        if (valid) {
            for (int i = 0; i < k; ++i) row[i] = i < c ? L.e[i].j : -1;
            kth_out[t] = c > 0 ? kth : -1.0f;
            invalid[t] = 0;
        } else {
            for (int i = 0; i < k; ++i) row[i] = -1;
            kth_out[t] = -1.0f;
            invalid[t] = 1;
            ++n_invalid;
        }
        // literal schedule comparison
        bool same = (lit_invalid == !valid);
        if (same && valid) {
            same = lit_list.size() == L.e.size();
            for (size_t i = 0; same && i < L.e.size(); ++i) same = lit_list[i].j == L.e[i].j;
        }
        if (!same) ++n_literal_diff;
    }
    if (stats) {
        stats[0] = n_invalid;
        stats[1] = n_literal_diff;
        stats[2] = n_evals;
    }
    return 0;
}

/* f1: query-vs-reference kNN, GridStructure::GRIDCUDAKNNCompact GridKNNCuda.cu:1645-1939 with the kernel
 * construct_neighbors :340-440. Queries (raw coordinates, original order) are normalized with the
 * reference cloud's extent (normalize_points :841-849, called at :1700-1703) and searched against the
 * sorted reference points with the self schedule (own cell at level 0, then ring 1 at bucket counts
 * B, B/2, ..., 8). It differs from the self search in four places:
 *   - candidates need d < thres (strict, :372; the self search has <=);
 *   - no point is excluded (the query is not a reference point);
 *   - the query's cell is computed at each level from its coordinates (:351-353); a query may lie outside
 *     the grid, and (int) truncates toward zero, so below the grid the cell is not a shift of the finer one;
 *   - the certificate's short-list branch (counter < k) compares with >= (:389-405; the self search has >).
 * The query order used inside the reference (rearrangeQueryPoints :1208-1267) only permutes rows and is
 * undone by GetKNNIndexes (:1941-1968), so rows here are in original query order. Entries are reference
 * sorted indices, ordered by (distance, sorted index). As in orc_knn_self the exact list decides validity
 * through the reference certificate. For a query inside the grid the ring stages cover the full 3x3x3 box of
 * each level (a level's own cell lies in the previous level's box), so the scanned list is exact. Outside
 * the grid that fails: the cell is truncated toward zero, so below the grid the own cell of the next level
 * can reach past the previous box, and the rings skip it. The reference then can stop with a list that
 * misses nearer points. Here such a query gets its exact list by brute force. stats[1] counts the rows
 * where the literal reference schedule differs. stats otherwise as in orc_knn_self. */
namespace {
bool ref_continue_query(const GridGeom& g, int s, float x, float y, float z, const int* cells, int c, int k,
                        float kth, float thres_n) {
    const int lvl = s == 0 ? 0 : s - 1;
    const int off = s == 0 ? 0 : 1;
    const int Bl = g.B >> lvl;
    const float h = g.ba / (float)Bl;
    const float coords[3] = {x, y, z};
    const float r = c < k ? std::sqrt(thres_n) : std::sqrt(kth);
    for (int ax = 0; ax < 3; ++ax) {
        const float v = coords[ax];
        const int ci = cells[ax];
        if (ci - off >= 0 && r >= std::fmaf(-(float)(ci - off), h, v - g.a)) return true;
        if (ci + off + 1 <= Bl && r >= std::fmaf((float)(ci + off + 1), h, g.a - v)) return true;
    }
    return false;
}
}  // namespace

int orc_knn_query(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n,
                  const float* qx_raw, const float* qy_raw, const float* qz_raw, int64_t nq, float min_extent,
                  float max_extent, float eps, int B, int k, float thres_n, int maxc, int32_t* knn, float* kth_out,
                  int32_t* invalid, int64_t* stats) {
    if (n < 1 || nq < 0 || k < 1 || B < 1) return -1;
    GridGeom g;
    g.a = -eps;
    g.b = 1.0f + eps;
    g.ba = g.b - g.a;
    g.B = B;
    int nlev = 0;
    for (int b = B; b >= 1; b >>= 1) ++nlev;
    std::vector<Level> lev(nlev);
    for (int l = 0; l < nlev; ++l) lev[l].build(xs, ys, zs, n, g.a, g.ba, B >> l);
    const int S = num_ring_stages(B);
    const float dext = max_extent - min_extent;
    int64_t n_invalid = 0, n_literal_diff = 0, n_evals = 0;
    const double margin = 2e-6;

#pragma omp parallel for schedule(dynamic, 256) reduction(+ : n_invalid, n_literal_diff, n_evals)
    for (int64_t q = 0; q < nq; ++q) {
        const float x = (qx_raw[q] - min_extent) / dext;
        const float y = (qy_raw[q] - min_extent) / dext;
        const float z = (qz_raw[q] - min_extent) / dext;
        TopK L(k);
        bool lit_done = false, lit_invalid = false, exact = false;
        std::vector<Entry> lit_list;
        for (int s = 0;; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            if (lvl >= nlev) break;
            const Level& Lv = lev[lvl];
            const int off = s == 0 ? 0 : 1;
            const int Bl = B >> lvl;
            const int cx = cell_of(x, g.a, g.ba, Bl), cy = cell_of(y, g.a, g.ba, Bl), cz = cell_of(z, g.a, g.ba, Bl);
            bool big = false;
            for (int dz = -off; dz <= off; ++dz)
                for (int dy = -off; dy <= off; ++dy)
                    for (int dx = -off; dx <= off; ++dx) {
                        if (s > 0 && dx == 0 && dy == 0 && dz == 0) continue;
                        int32_t st, en;
                        const int32_t cnt = Lv.lookup(cx + dx, cy + dy, cz + dz, st, en);
                        if (cnt <= 0) continue;
                        if (cnt >= maxc) big = true;
                        for (int32_t p = st; p < en; ++p) {
                            const int32_t j = Lv.order[p];
                            const float d = sqf(x, y, z, xs[j], ys[j], zs[j]);
                            ++n_evals;
                            if (d < thres_n) L.offer(d, j);
                        }
                    }
            const int cells[3] = {cx, cy, cz};
            if (!lit_done && s <= S) {
                if (big) {
                    lit_done = true;
                    lit_invalid = true;
                } else {
                    const int c = (int)L.e.size();
                    const float kth = c > 0 ? L.e.back().d : 0.0f;
                    if (!ref_continue_query(g, s, x, y, z, cells, c, k, kth, thres_n)) {
                        lit_done = true;
                        lit_list = L.e;
                    } else if (s == S) {
                        lit_done = true;
                        lit_invalid = true;
                    }
                }
            }
            // exactness of the list over the scanned boxes (double, with margin)
            const double hl = (double)g.ba / (double)Bl;
            const double bound = (int)L.e.size() == k ? (double)L.e.back().d : (double)thres_n;
            const double R = std::sqrt(bound) + margin;
            exact = true;
            const double qc[3] = {x, y, z};
            for (int ax = 0; ax < 3 && exact; ++ax) {
                if (cells[ax] - off >= 1) {
                    const double face = (double)g.a + (double)(cells[ax] - off) * hl;
                    if (!(R < qc[ax] - face)) exact = false;
                }
                if (cells[ax] + off <= Bl - 2) {
                    const double face = (double)g.a + (double)(cells[ax] + off + 1) * hl;
                    if (!(R < face - qc[ax])) exact = false;
                }
                // a box that misses the grid on this axis certifies nothing
                if (cells[ax] + off < 0 || cells[ax] - off > Bl - 1) exact = false;
            }
            if (exact && lit_done) break;
        }
        const int fq[3] = {cell_of(x, g.a, g.ba, B), cell_of(y, g.a, g.ba, B), cell_of(z, g.a, g.ba, B)};
        const bool outside = fq[0] < 0 || fq[1] < 0 || fq[2] < 0 || fq[0] >= B || fq[1] >= B || fq[2] >= B;
        if (outside || !exact) {
            // Outside the grid the ring stages do not cover the boxes: below the grid a level's own cell
            // (truncated toward zero) can hold points that no earlier box held, and the rings skip it. The
            // exact list then comes from brute force.
            TopK A(k);
            for (int64_t j = 0; j < n; ++j) {
                const float d = sqf(x, y, z, xs[j], ys[j], zs[j]);
                if (d < thres_n) A.offer(d, (int32_t)j);
            }
            L = A;
        }
        // reference validity from the exact list
        const int c = (int)L.e.size();
        const float kth = c > 0 ? L.e.back().d : 0.0f;
        int lox = 0, hix = 0, loy = 0, hiy = 0, loz = 0, hiz = 0;
        if (c > 0) {
            lox = loy = loz = B;
            hix = hiy = hiz = -1;
            for (const Entry& e : L.e) {
                const int32_t kk = key[e.j];
                const int mx_ = kk % B, my_ = (kk / B) % B, mz_ = kk / (B * B);
                lox = std::min(lox, mx_); hix = std::max(hix, mx_);
                loy = std::min(loy, my_); hiy = std::max(hiy, my_);
                loz = std::min(loz, mz_); hiz = std::max(hiz, mz_);
            }
        }
        bool valid = false, decided = false;
        for (int s = 0; s <= S && !decided; ++s) {
            const int lvl = s == 0 ? 0 : s - 1;
            const int off = s == 0 ? 0 : 1;
            const int Bl = B >> lvl;
            const Level& Lv = lev[lvl];
            const int cx = cell_of(x, g.a, g.ba, Bl), cy = cell_of(y, g.a, g.ba, Bl), cz = cell_of(z, g.a, g.ba, Bl);
            for (int dz = -off; dz <= off && !decided; ++dz)
                for (int dy = -off; dy <= off && !decided; ++dy)
                    for (int dx = -off; dx <= off; ++dx) {
                        if (s > 0 && dx == 0 && dy == 0 && dz == 0) continue;
                        int32_t st, en;
                        if (Lv.lookup(cx + dx, cy + dy, cz + dz, st, en) >= maxc) {
                            decided = true;
                            valid = false;
                            break;
                        }
                    }
            if (decided) break;
            bool inside = true;
            if (c > 0) {
                inside = (lox >> lvl) >= cx - off && (hix >> lvl) <= cx + off && (loy >> lvl) >= cy - off &&
                         (hiy >> lvl) <= cy + off && (loz >> lvl) >= cz - off && (hiz >> lvl) <= cz + off;
            }
            const int cells[3] = {cx, cy, cz};
            if (inside && !ref_continue_query(g, s, x, y, z, cells, c, k, kth, thres_n)) {
                decided = true;
                valid = true;
            }
        }
        int32_t* row = knn + (size_t)q * k;
        if (valid) {
            for (int i = 0; i < k; ++i) row[i] = i < c ? L.e[i].j : -1;
            kth_out[q] = c > 0 ? kth : -1.0f;
            invalid[q] = 0;
        } else {
            for (int i = 0; i < k; ++i) row[i] = -1;
            kth_out[q] = -1.0f;
            invalid[q] = 1;
            ++n_invalid;
        }
        bool same = (lit_invalid == !valid);
        if (same && valid) {
            same = lit_list.size() == L.e.size();
            for (size_t i = 0; same && i < L.e.size(); ++i) same = lit_list[i].j == L.e[i].j;
        }
        if (!same) ++n_literal_diff;
    }
    if (stats) {
        stats[0] = n_invalid;
        stats[1] = n_literal_diff;
        stats[2] = n_evals;
    }
    return 0;
}

/* Brute-force query kNN (tests only): every reference point with d < thres_n (normalized query coordinates
 * as orc_knn_query computes them), ordered by (d, sorted index). */
int orc_knn_query_brute(const float* xs, const float* ys, const float* zs, int64_t n, const float* qx_raw,
                        const float* qy_raw, const float* qz_raw, int64_t nq, float min_extent, float max_extent,
                        int k, float thres_n, int32_t* knn) {
    const float dext = max_extent - min_extent;
    parallel_for(int64_t q = 0; q < nq; ++q) {
        const float x = (qx_raw[q] - min_extent) / dext;
        const float y = (qy_raw[q] - min_extent) / dext;
        const float z = (qz_raw[q] - min_extent) / dext;
        TopK L(k);
        for (int64_t j = 0; j < n; ++j) {
            const float d = sqf(x, y, z, xs[j], ys[j], zs[j]);
            if (d < thres_n) L.offer(d, (int32_t)j);
        }
        for (int i = 0; i < k; ++i) knn[(size_t)q * k + i] = i < (int)L.e.size() ? L.e[i].j : -1;
    }
    return 0;
}

/* Brute-force exact kNN (tests only): every p != q with d <= thres_n, ordered by (d, sorted index). */
int orc_knn_brute(const float* xs, const float* ys, const float* zs, int64_t n, int k, float thres_n, int32_t* knn) {
    parallel_for(int64_t q = 0; q < n; ++q) {
        TopK L(k);
        for (int64_t j = 0; j < n; ++j) {
            if (j == q) continue;
            const float d = sqf(xs[q], ys[q], zs[q], xs[j], ys[j], zs[j]);
            if (d <= thres_n) L.offer(d, (int32_t)j);
        }
        for (int i = 0; i < k; ++i) knn[(size_t)q * k + i] = i < (int)L.e.size() ? L.e[i].j : -1;
    }
    return 0;
}

namespace {
inline void normalize3(float& x, float& y, float& z) {
    const float nrm = std::sqrt(std::fmaf(z, z, std::fmaf(y, y, x * x)));
    x /= nrm;
    y /= nrm;
    z /= nrm;
}
inline float rowt(float a1, float a2, float a3, float t, float x, float y, float z) {
    return std::fmaf(a3, z, std::fmaf(a2, y, a1 * x)) + t;
}
// isEllipicGabrielNeighbor_GKNN, EGT_CUDA.cu:136-170 (+ helpers :55-134)
bool egg_keep(const float* xs, const float* ys, const float* zs, const int32_t* row, int j, float x, float y, float z,
              float alpha) {
    const int32_t q = row[j];
    const float xi = xs[q], yi = ys[q], zi = zs[q];
    float xo = 0.5f * (xi + x), yo = 0.5f * (yi + y), zo = 0.5f * (zi + z);
    const float d = std::sqrt(sqf(xi, yi, zi, x, y, z)) / 2.0f;
    if ((double)d < 1e-6) return true;
    float ax = xi - x, ay = yi - y, az = zi - z;
    // :149's second test (|q-p| < 1e-6) is implied by d < 1e-6 above and never fires.
    normalize3(ax, ay, az);
    // GetNormalizedPerpendicularVectorToVector_GKNN :61-95
    float mxv = std::fabs(ax);
    int ci = 0;
    if (mxv < std::fabs(ay)) {
        ci = 1;
        mxv = std::fabs(ay);
    }
    if (mxv < std::fabs(az)) ci = 2;
    float bx = 1.0f, by = 1.0f, bz = 1.0f;
    if (ci == 0) bx = (-ay - az) / ax;
    else if (ci == 1) by = (-ax - az) / ay;
    else bz = (-ax - ay) / az;
    normalize3(bx, by, bz);
    // cross_GKNN :97-104, then normalize
    float cx = ay * bz - by * az;
    float cy = bx * az - ax * bz;
    float cz = ax * by - bx * ay;
    normalize3(cx, cy, cz);
    // CreateLocalCoordinateSystem_GKNN :117-134: translation = -(R * mid)
    const float tx = -rowt(ax, ay, az, 0.0f, xo, yo, zo);
    const float ty = -rowt(bx, by, bz, 0.0f, xo, yo, zo);
    const float tz = -rowt(cx, cy, cz, 0.0f, xo, yo, zo);
    const float aa = alpha * alpha;
    const float dd = d * d;
    for (int m = 0; m < j; ++m) {
        const int32_t nb = row[m];
        if (nb < 0) continue;
        const float px = xs[nb], py = ys[nb], pz = zs[nb];
        const float ex = rowt(ax, ay, az, tx, px, py, pz);
        const float ey = rowt(bx, by, bz, ty, px, py, pz);
        const float ez = rowt(cx, cy, cz, tz, px, py, pz);
        const float v = ex * ex + ey * ey / aa + ez * ez / aa;
        if (v < dd) return false;
    }
    return true;
}
}  // namespace

/* a6: EGG filter in place, calculateEGG_GKNN EGT_CUDA.cu:172-189. Slot j is removed (-1) when a valid
 * closer-ranked slot m < j lies strictly inside the ellipsoid on segment p-q_j. */
int orc_egg(const float* xs, const float* ys, const float* zs, int64_t n, int k, float alpha, int32_t* knn) {
    parallel_for(int64_t i = 0; i < n; ++i) {
        int32_t* row = knn + (size_t)i * k;
        const float x = xs[i], y = ys[i], z = zs[i];
        for (int j = k - 1; j >= 0; --j)
            if (row[j] != -1 && !egg_keep(xs, ys, zs, row, j, x, y, z, alpha)) row[j] = -1;
    }
    return 0;
}

/* The EGG filter of rows that belong to the points qsel[0..nsel) (row t is point qsel[t]'s). */
int orc_egg_subset(const float* xs, const float* ys, const float* zs, const int32_t* qsel, int64_t nsel, int k,
                   float alpha, int32_t* knn) {
    parallel_for(int64_t t = 0; t < nsel; ++t) {
        int32_t* row = knn + (size_t)t * k;
        const int64_t i = qsel[t];
        const float x = xs[i], y = ys[i], z = zs[i];
        for (int j = k - 1; j >= 0; --j)
            if (row[j] != -1 && !egg_keep(xs, ys, zs, row, j, x, y, z, alpha)) row[j] = -1;
    }
    return 0;
}

/* a7: Taubin lambda/mu Jacobi iterations, taubin_step_GKNN EGT_CUDA.cu:229-276 with the loop :545-634
 * (one iteration = lambda pass then mu pass, ping-pong). Weights w = exp(-|pn - pi|^2). */
int orc_taubin_f32(float* x, float* y, float* z, const int32_t* knn, int64_t n, int k, int iters, float lambda,
                   float mu, int regularized) {
    std::vector<float> ox(n), oy(n), oz(n);
    float *ix = x, *iy = y, *iz = z, *px = ox.data(), *py = oy.data(), *pz = oz.data();
    for (int it = 0; it < 2 * iters; ++it) {
        const float scale = (it & 1) ? mu : lambda;
        parallel_for(int64_t i = 0; i < n; ++i) {
            float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, sum = 0.0f;
            const float x1 = ix[i], y1 = iy[i], z1 = iz[i];
            const int32_t* row = knn + (size_t)i * k;
            for (int s = 0; s < k; ++s) {
                const int32_t nb = row[s];
                if (nb == -1) continue;
                const float x2 = ix[nb], y2 = iy[nb], z2 = iz[nb];
                const float dist = sqf(x1, y1, z1, x2, y2, z2);
                // isRegularized (EGT_CUDA.cu:263): w = 1 in the lambda pass, 1 / (d^2 + 1e-8) in the mu pass
                const float w = regularized ? (scale < 0.0f ? 1.0f / (dist + 1e-8f) : 1.0f) : egt_expf(-dist);
                cgx = std::fmaf(w, x2 - x1, cgx);
                cgy = std::fmaf(w, y2 - y1, cgy);
                cgz = std::fmaf(w, z2 - z1, cgz);
                sum += w;
            }
            if (sum == 0.0f) sum = 1e-8f;
            px[i] = x1 + scale * cgx / sum;
            py[i] = y1 + scale * cgy / sum;
            pz[i] = z1 + scale * cgz / sum;
        }
        std::swap(ix, px);
        std::swap(iy, py);
        std::swap(iz, pz);
    }
    // 2*iters swaps: the result is back in x, y, z.
    return 0;
}

/* FP64 twin of orc_taubin_f32 (error reference only). */
int orc_taubin_f64(double* x, double* y, double* z, const int32_t* knn, int64_t n, int k, int iters, double lambda,
                   double mu, int regularized) {
    std::vector<double> ox(n), oy(n), oz(n);
    double *ix = x, *iy = y, *iz = z, *px = ox.data(), *py = oy.data(), *pz = oz.data();
    for (int it = 0; it < 2 * iters; ++it) {
        const double scale = (it & 1) ? mu : lambda;
        parallel_for(int64_t i = 0; i < n; ++i) {
            double cgx = 0, cgy = 0, cgz = 0, sum = 0;
            const double x1 = ix[i], y1 = iy[i], z1 = iz[i];
            const int32_t* row = knn + (size_t)i * k;
            for (int s = 0; s < k; ++s) {
                const int32_t nb = row[s];
                if (nb == -1) continue;
                const double dx = x1 - ix[nb], dy = y1 - iy[nb], dz = z1 - iz[nb];
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double w = regularized ? (scale < 0.0 ? 1.0 / (d2 + 1e-8) : 1.0) : std::exp(-d2);
                cgx += w * (ix[nb] - x1);
                cgy += w * (iy[nb] - y1);
                cgz += w * (iz[nb] - z1);
                sum += w;
            }
            if (sum == 0.0) sum = 1e-8;
            px[i] = x1 + scale * cgx / sum;
            py[i] = y1 + scale * cgy / sum;
            pz[i] = z1 + scale * cgz / sum;
        }
        std::swap(ix, px);
        std::swap(iy, py);
        std::swap(iz, pz);
    }
    return 0;
}

/* a8: de-normalize and scatter to original order, getOutputPoints.cu:9-30. */
int orc_output(const float* xs, const float* ys, const float* zs, const int32_t* perm, int64_t n, float mn, float mx,
               float* ox, float* oy, float* oz) {
    const float d = mx - mn;
    for (int64_t s = 0; s < n; ++s) {
        const int32_t o = perm[s];
        ox[o] = xs[s] * d + mn;
        oy[o] = ys[s] * d + mn;
        oz[o] = zs[s] * d + mn;
    }
    return 0;
}

/* a12: kNN rows in original index space, GetKNNIndexes GridKNNCuda.cu:1941-1968 +
 * rearrangeReferenceNeighbors :862-872. */
int orc_knn_to_original(const int32_t* knn, const int32_t* perm, int64_t n, int k, int32_t* out) {
    for (int64_t s = 0; s < n; ++s) {
        const int32_t o = perm[s];
        for (int i = 0; i < k; ++i) {
            const int32_t v = knn[(size_t)s * k + i];
            out[(size_t)o * k + i] = v >= 0 ? perm[v] : -1;
        }
    }
    return 0;
}

/* a11: mesh one-ring adjacency in sorted space, getMeshNeihbors GridMeshCuda1.cu:99-138 with the batch loop
 * :416-511. Each vertex lists the distinct other corners of its valid triangles (a repeated corner lists the
 * vertex itself at distance 0, as the reference does), distance-ascending by (d, sorted index), cut to k,
 * padded with -1. Triangles with an index >= n are skipped (the reference remaps before its 0xFFFFFFFF
 * test, :224-232 vs :106, so such triangles are undefined there). inv_perm maps original -> sorted. */
int orc_mesh_adjacency(const float* xs, const float* ys, const float* zs, int64_t n, const uint32_t* tri,
                       int64_t ntri, const int32_t* inv_perm, int k, int32_t* knn) {
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t t = 0; t < ntri; ++t) {
        const uint32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
        if (v0 >= (uint64_t)n || v1 >= (uint64_t)n || v2 >= (uint64_t)n) continue;
        cnt[inv_perm[v0] + 1] += 2;
        cnt[inv_perm[v1] + 1] += 2;
        cnt[inv_perm[v2] + 1] += 2;
    }
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> adj(cnt[n]);
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    for (int64_t t = 0; t < ntri; ++t) {
        const uint32_t v0 = tri[3 * t], v1 = tri[3 * t + 1], v2 = tri[3 * t + 2];
        if (v0 >= (uint64_t)n || v1 >= (uint64_t)n || v2 >= (uint64_t)n) continue;
        const int32_t s[3] = {inv_perm[v0], inv_perm[v1], inv_perm[v2]};
        for (int c = 0; c < 3; ++c)
            for (int o = 0; o < 3; ++o)
                if (o != c) adj[cur[s[c]]++] = s[o];
    }
    parallel_for(int64_t i = 0; i < n; ++i) {
        TopK L(k);
        for (int64_t p = cnt[i]; p < cnt[i + 1]; ++p) {
            const int32_t j = adj[p];
            L.offer(sqf(xs[i], ys[i], zs[i], xs[j], ys[j], zs[j]), j);
        }
        for (int s = 0; s < k; ++s) knn[(size_t)i * k + s] = s < (int)L.e.size() ? L.e[s].j : -1;
    }
    return 0;
}

/* a9: PCA normal per point, calculateNormal Normal_CUDA.cu:87-144: centroid of the point and its valid
 * neighbours, covariance of the neighbours only (self excluded), eigenvector of the smallest eigenvalue,
 * normalized. Sorted space in and out. */
int orc_pca_normals(const float* xs, const float* ys, const float* zs, const int32_t* knn, int64_t n, int k,
                    float* nx, float* ny, float* nz) {
    parallel_for(int64_t i = 0; i < n; ++i) {
        const int32_t* row = knn + (size_t)i * k;
        float sx = xs[i], sy = ys[i], sz = zs[i];
        int counter = 1;
        for (int j = 0; j < k; ++j) {
            const int32_t nb = row[j];
            if (nb >= 0) {
                sx += xs[nb];
                sy += ys[nb];
                sz += zs[nb];
                ++counter;
            }
        }
        const float cx = sx / (float)counter, cy = sy / (float)counter, cz = sz / (float)counter;
        float a6[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < k; ++j) {
            const int32_t nb = row[j];
            if (nb < 0) continue;
            const float dx = xs[nb] - cx, dy = ys[nb] - cy, dz = zs[nb] - cz;
            a6[0] = std::fmaf(dx, dx, a6[0]);
            a6[1] = std::fmaf(dx, dy, a6[1]);
            a6[2] = std::fmaf(dx, dz, a6[2]);
            a6[3] = std::fmaf(dy, dy, a6[3]);
            a6[4] = std::fmaf(dy, dz, a6[4]);
            a6[5] = std::fmaf(dz, dz, a6[5]);
        }
        float ev[3], V[9];
        const int m = orc_eig3(a6, ev, V);
        const float vx = V[3 * m], vy = V[3 * m + 1], vz = V[3 * m + 2];
        const float nrm = std::sqrt(std::fmaf(vz, vz, std::fmaf(vy, vy, vx * vx)));
        nx[i] = vx / nrm;
        ny[i] = vy / nrm;
        nz[i] = vz / nrm;
    }
    return 0;
}

/* a10: normal orientation, Normal_CUDA.cu:304-711 on one card.
 * Undirected edges {i, n} of the (EGG-filtered) rows, weight max(0, 1 - |n_i . n_n|) (fillEdges :146-175),
 * deduplicated; minimum spanning forest under the total order (weight, u, v) with u < v (the reference's
 * stable weight sort after its u*N+v hash sort :363-385 gives the same order; its GPU Boruvka
 * CUDA_MST.cu:296-418 resolves equal weights by (w, v)); per tree the root is u of the tree's lightest edge
 * (the reference uses u of the globally lightest edge for the first tree :438-440 and a racy pick
 * findNextRoot :232-240 for the others); the root keeps its PCA sign and every tree edge (p, c) flips c
 * when n_p' . n_c < 0 (BFS_Bqueue_kernel bfs.cu:8-52), where n_p' is the parent's oriented normal.
 * A dot product of exactly 0 keeps the parent's sign (the reference resets the child to its own PCA sign
 * there). On one card the component vote of :514-711 never finds an edge between two trees, so it is a
 * no-op and is not restated. Vertices without edges keep their PCA sign.
 * stats: [0] unique edges, [1] forest edges, [2] trees, [3] flipped normals; wsum = forest weight (double). */
int orc_orient_normals(const int32_t* knn, int64_t n, int k, float* nx, float* ny, float* nz, int64_t* stats,
                       double* wsum) {
    struct E {
        float w;
        int32_t u, v;
        uint8_t rel;
    };
    std::vector<std::pair<int32_t, int32_t>> raw;
    raw.reserve((size_t)n * 4);
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < k; ++j) {
            const int32_t nb = knn[(size_t)i * k + j];
            if (nb < 0 || nb == i) continue;
            raw.emplace_back(std::min<int32_t>((int32_t)i, nb), std::max<int32_t>((int32_t)i, nb));
        }
    std::sort(raw.begin(), raw.end());
    raw.erase(std::unique(raw.begin(), raw.end()), raw.end());
    std::vector<E> edges(raw.size());
    for (size_t e = 0; e < raw.size(); ++e) {
        const int32_t u = raw[e].first, v = raw[e].second;
        const float dot = std::fmaf(nz[u], nz[v], std::fmaf(ny[u], ny[v], nx[u] * nx[v]));
        edges[e] = E{std::fmax(0.0f, 1.0f - std::fabs(dot)), u, v, (uint8_t)(dot < 0.0f)};
    }
    std::stable_sort(edges.begin(), edges.end(), [](const E& l, const E& r) { return l.w < r.w; });
    std::vector<int32_t> par(n);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int32_t a) {
        while (par[a] != a) {
            par[a] = par[par[a]];
            a = par[a];
        }
        return a;
    };
    std::vector<int64_t> off(n + 1, 0);
    std::vector<size_t> forest;
    double W = 0.0;
    for (size_t e = 0; e < edges.size(); ++e) {
        const int32_t ru = find(edges[e].u), rv = find(edges[e].v);
        if (ru == rv) continue;
        par[ru] = rv;
        forest.push_back(e);
        W += edges[e].w;
        off[edges[e].u + 1]++;
        off[edges[e].v + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    std::vector<int32_t> adj(off[n]);
    std::vector<uint8_t> arel(off[n]);
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (size_t f : forest) {
        const E& e = edges[f];
        adj[cur[e.u]] = e.v;
        arel[cur[e.u]++] = e.rel;
        adj[cur[e.v]] = e.u;
        arel[cur[e.v]++] = e.rel;
    }
    std::vector<int8_t> sign(n, -1);  // -1 unvisited, else parity
    int64_t trees = 0, flipped = 0;
    std::vector<int32_t> queue;
    for (size_t f : forest) {  // forest edges are in (w, u, v) order: first edge of a tree gives its root
        const int32_t r = edges[f].u;
        if (sign[r] >= 0) continue;
        ++trees;
        sign[r] = 0;
        queue.assign(1, r);
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t p = queue[h];
            for (int64_t a = off[p]; a < off[p + 1]; ++a) {
                const int32_t c = adj[a];
                if (sign[c] >= 0) continue;
                sign[c] = (int8_t)(sign[p] ^ arel[a]);
                queue.push_back(c);
            }
        }
    }
    for (int64_t i = 0; i < n; ++i)
        if (sign[i] == 1) {
            nx[i] = -nx[i];
            ny[i] = -ny[i];
            nz[i] = -nz[i];
            ++flipped;
        }
    if (stats) {
        stats[0] = (int64_t)edges.size();
        stats[1] = (int64_t)forest.size();
        stats[2] = trees;
        stats[3] = flipped;
    }
    if (wsum) *wsum = W;
    return 0;
}

}  // extern "C"
