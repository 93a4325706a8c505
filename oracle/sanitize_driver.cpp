// Test infrastructure only: drives every entry point of the oracle once on small seeded inputs, for the
// AddressSanitizer / UndefinedBehaviorSanitizer build (make -C oracle sanitize). A sanitizer report aborts
// the run with a non-zero status; a clean run prints "sanitize ok".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
float egt_expf(float x);
int orc_eig3(const float* a6, float* evals, float* evecs);
int orc_grid_build(const float* x, const float* y, const float* z, int64_t n, float eps, int B, float* xs, float* ys,
                   float* zs, int32_t* perm, int32_t* key, float* extent2);
int orc_knn_self(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, float eps, int B,
                 int k, float thres_n, int maxc, int32_t* knn, float* kth_out, int32_t* invalid, int64_t* stats);
int orc_knn_query(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, const float* qx,
                  const float* qy, const float* qz, int64_t nq, float min_extent, float max_extent, float eps, int B,
                  int k, float thres_n, int maxc, int32_t* knn, float* kth_out, int32_t* invalid, int64_t* stats);
int orc_knn_query_brute(const float* xs, const float* ys, const float* zs, int64_t n, const float* qx,
                        const float* qy, const float* qz, int64_t nq, float min_extent, float max_extent, int k,
                        float thres_n, int32_t* knn);
int orc_knn_brute(const float* xs, const float* ys, const float* zs, int64_t n, int k, float thres_n, int32_t* knn);
int orc_knn_self_subset(const float* xs, const float* ys, const float* zs, const int32_t* key, int64_t n, float eps,
                        int B, int k, float thres_n, int maxc, const int32_t* qsel, int64_t nsel, int32_t* knn,
                        float* kth_out, int32_t* invalid, int64_t* stats);
int orc_egg_subset(const float* xs, const float* ys, const float* zs, const int32_t* qsel, int64_t nsel, int k,
                   float alpha, int32_t* knn);
int orc_egg(const float* xs, const float* ys, const float* zs, int64_t n, int k, float alpha, int32_t* knn);
int orc_taubin_f32(float* x, float* y, float* z, const int32_t* knn, int64_t n, int k, int iters, float lambda,
                   float mu, int regularized);
int orc_taubin_f64(double* x, double* y, double* z, const int32_t* knn, int64_t n, int k, int iters, double lambda,
                   double mu, int regularized);
int orc_output(const float* xs, const float* ys, const float* zs, const int32_t* perm, int64_t n, float mn, float mx,
               float* ox, float* oy, float* oz);
int orc_knn_to_original(const int32_t* knn, const int32_t* perm, int64_t n, int k, int32_t* out);
int orc_mesh_adjacency(const float* xs, const float* ys, const float* zs, int64_t n, const uint32_t* tri,
                       int64_t ntri, const int32_t* inv_perm, int k, int32_t* knn);
int orc_pca_normals(const float* xs, const float* ys, const float* zs, const int32_t* knn, int64_t n, int k,
                    float* nx, float* ny, float* nz);
int orc_orient_normals(const int32_t* knn, int64_t n, int k, float* nx, float* ny, float* nz, int64_t* stats,
                       double* wsum);
}

int main() {
    const int64_t n = 3000, nq = 500;
    const int k = 16, B = 64;
    std::mt19937 rng(12345);
    std::normal_distribution<float> g(0.0f, 1.0f);
    std::vector<float> x(n), y(n), z(n);
    for (int64_t i = 0; i < n; ++i) {  // noisy unit sphere, with a few exact duplicates
        float a = g(rng), b = g(rng), c = g(rng);
        const float r = std::sqrt(a * a + b * b + c * c);
        x[i] = a / r + 0.01f * g(rng);
        y[i] = b / r + 0.01f * g(rng);
        z[i] = c / r + 0.01f * g(rng);
        if (i >= 10 && i % 97 == 0) {
            x[i] = x[i - 10];
            y[i] = y[i - 10];
            z[i] = z[i - 10];
        }
    }
    std::vector<float> xs(n), ys(n), zs(n), ext(2);
    std::vector<int32_t> perm(n), key(n), inv(n);
    int bad = 0;
    bad |= orc_grid_build(x.data(), y.data(), z.data(), n, 1e-6f, B, xs.data(), ys.data(), zs.data(), perm.data(),
                          key.data(), ext.data());
    for (int64_t s = 0; s < n; ++s) inv[perm[s]] = (int32_t)s;
    const float thres_n = 1e8f;
    std::vector<int32_t> knn(n * k), brute(n * k), invalid(n), orig(n * k);
    std::vector<float> kth(n);
    std::vector<int64_t> stats(16);
    bad |= orc_knn_self(xs.data(), ys.data(), zs.data(), key.data(), n, 1e-6f, B, k, thres_n, 100000, knn.data(),
                        kth.data(), invalid.data(), stats.data());
    bad |= orc_knn_self(xs.data(), ys.data(), zs.data(), key.data(), n, 1e-6f, B, k, thres_n, 20, brute.data(),
                        kth.data(), invalid.data(), stats.data());  // population rule hit
    bad |= orc_knn_brute(xs.data(), ys.data(), zs.data(), n, k, thres_n, brute.data());
    bad |= orc_knn_to_original(knn.data(), perm.data(), n, k, orig.data());
    std::vector<float> qx(nq), qy(nq), qz(nq);
    for (int64_t i = 0; i < nq; ++i) {  // inside and far outside the reference cube
        const float s = i % 5 == 0 ? 3.0f : 1.0f;
        qx[i] = s * g(rng);
        qy[i] = s * g(rng);
        qz[i] = s * g(rng);
    }
    std::vector<int32_t> qknn(nq * k), qbrute(nq * k), qinv(nq);
    std::vector<float> qkth(nq);
    bad |= orc_knn_query(xs.data(), ys.data(), zs.data(), key.data(), n, qx.data(), qy.data(), qz.data(), nq, ext[0],
                         ext[1], 1e-6f, B, k, thres_n, 100000, qknn.data(), qkth.data(), qinv.data(), stats.data());
    bad |= orc_knn_query_brute(xs.data(), ys.data(), zs.data(), n, qx.data(), qy.data(), qz.data(), nq, ext[0],
                               ext[1], k, thres_n, qbrute.data());
    std::vector<int32_t> egg = knn;
    bad |= orc_egg(xs.data(), ys.data(), zs.data(), n, k, 0.65f, egg.data());
    {  // every third query through the subset entry points
        std::vector<int32_t> sel;
        for (int64_t i = 0; i < n; i += 3) sel.push_back((int32_t)i);
        const int64_t m = (int64_t)sel.size();
        std::vector<int32_t> sk(m * k), sinv(m);
        std::vector<float> skth(m);
        bad |= orc_knn_self_subset(xs.data(), ys.data(), zs.data(), key.data(), n, 1e-6f, B, k, thres_n, 100000,
                                   sel.data(), m, sk.data(), skth.data(), sinv.data(), stats.data());
        bad |= orc_egg_subset(xs.data(), ys.data(), zs.data(), sel.data(), m, k, 0.65f, sk.data());
        for (int64_t t = 0; t < m; ++t)
            for (int j = 0; j < k; ++j) bad |= sk[t * k + j] != egg[(int64_t)sel[t] * k + j];
    }
    std::vector<float> tx = xs, ty = ys, tz = zs;
    bad |= orc_taubin_f32(tx.data(), ty.data(), tz.data(), egg.data(), n, k, 5, 0.5f, -0.53f, 0);
    bad |= orc_taubin_f32(tx.data(), ty.data(), tz.data(), egg.data(), n, k, 2, 0.5f, -0.53f, 1);
    std::vector<double> dx(xs.begin(), xs.end()), dy(ys.begin(), ys.end()), dz(zs.begin(), zs.end());
    bad |= orc_taubin_f64(dx.data(), dy.data(), dz.data(), egg.data(), n, k, 5, 0.5, -0.53, 0);
    std::vector<float> nx(n), ny(n), nz(n), ox(n), oy(n), oz(n);
    bad |= orc_pca_normals(tx.data(), ty.data(), tz.data(), egg.data(), n, k, nx.data(), ny.data(), nz.data());
    double w = 0.0;
    bad |= orc_orient_normals(egg.data(), n, k, nx.data(), ny.data(), nz.data(), stats.data(), &w);
    bad |= orc_output(tx.data(), ty.data(), tz.data(), perm.data(), n, ext[0], ext[1], ox.data(), oy.data(),
                      oz.data());
    // a small mesh: a triangle fan over consecutive original vertices, one degenerate triangle, one
    // triangle with an index past the vertex count (skipped)
    std::vector<uint32_t> tri;
    for (uint32_t i = 1; i + 1 < 200; ++i) tri.insert(tri.end(), {0u, i, i + 1});
    tri.insert(tri.end(), {5u, 5u, 6u});
    tri.insert(tri.end(), {7u, 8u, (uint32_t)n + 3u});
    std::vector<int32_t> mrows(n * k);
    bad |= orc_mesh_adjacency(xs.data(), ys.data(), zs.data(), n, tri.data(), (int64_t)tri.size() / 3, inv.data(), k,
                              mrows.data());
    float a6[6] = {2.0f, 0.5f, 0.1f, 1.0f, 0.2f, 0.3f}, ev[3], V[9];
    (void)orc_eig3(a6, ev, V);
    float e = 0.0f;
    for (float t = -20.0f; t <= 0.0f; t += 0.37f) e += egt_expf(t);
    if (bad || !std::isfinite(e)) {
        std::printf("oracle call failed\n");
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
