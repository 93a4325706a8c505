// KNNCUDA command line on the gfx950 library, argument-compatible with the reference's main.cpp:76-136:
//
//   KNNCUDA file.ply k iterations alpha isMesh numCards [lambda mu]
//
// Point cloud (isMesh 0): grid kNN (eps 1e-6, bucket count 1024, thres 1e8), EGG-filtered Taubin smoothing,
// PCA normals with MST orientation. Mesh (isMesh 1): triangle adjacency instead of the kNN, no normals.
// Writes output.ply in the working directory (layout of MeshPLY.cpp:67-127) and prints "Succeeded!".
// lambda / mu default to the reference's hard-coded 0.63 / -0.64 (main.cpp:116); numCards must be 1.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "egt.hpp"
#include "ply_io.hpp"

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s file.ply k iterations alpha isMesh numCards [lambda mu]\n", argv[0]);
        return 2;
    }
    const std::string path = argv[1];
    const int k = std::atoi(argv[2]);
    const int iterations = std::atoi(argv[3]);
    const float alpha = (float)std::atof(argv[4]);
    const bool isMesh = std::atoi(argv[5]) != 0;
    const int numcards = std::atoi(argv[6]);
    const float lambda = argc > 7 ? (float)std::atof(argv[7]) : 0.63f;
    const float mu = argc > 8 ? (float)std::atof(argv[8]) : -0.64f;
    try {
        Mesh mesh;
        ply::read(path, &mesh);
        const size_t n = mesh.nV();
        std::vector<float> nx(n), ny(n), nz(n), ox(n), oy(n), oz(n);
        const auto t0 = std::chrono::steady_clock::now();
        std::shared_ptr<KNNInterface> knn;
        if (!isMesh) {
            auto g = std::make_shared<GridStructure>(mesh, numcards, 1e-6f, 1024, 4);
            g->GRIDCUDAKNNSELF_COMPACT(k, 1e8f);
            knn = g;
        } else {
            auto g = std::make_shared<GridMeshCuda>(mesh, numcards, 1e-6f, 1024);
            g->GetNeighborsFromTriangles(k);
            knn = g;
        }
        EGT_CUDA smoother(*knn);
        smoother.PerformSmoothing(iterations, lambda, mu, alpha);
        if (!isMesh) {
            Normal_CUDA normals(*knn);
            normals.GetNormals(nx.data(), ny.data(), nz.data());
        }
        getOutputPoints(ox.data(), oy.data(), oz.data(), knn.get());
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!isMesh) mesh.mN.resize(n);
        for (size_t i = 0; i < n; ++i) {
            mesh.mV[i] = V3{ox[i], oy[i], oz[i]};
            if (!isMesh) mesh.mN[i] = V3{nx[i], ny[i], nz[i]};
        }
        ply::write("output.ply", &mesh);
        std::fprintf(stderr, "%zu points, %.3f s on the GPU path\n", n, secs);
        std::printf("Succeeded!\n");
        return 0;
    } catch (const egt::Error& e) {
        std::fprintf(stderr, "KNNCUDA: library error %d: %s\n", (int)e.status, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "KNNCUDA: %s\n", e.what());
        return 1;
    }
}
