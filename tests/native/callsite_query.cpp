// A reference-style call site (main.cpp / an ICP loop) written against the C++ mirror: it must compile
// unchanged. Compiled with -fsyntax-only; nothing runs.
#include <vector>
#include "egt.hpp"

void icp_step(const Mesh& reference, std::vector<float>& qx, std::vector<float>& qy, std::vector<float>& qz) {
    GridStructure grid(reference, 1, 1e-6f, 1024, 4);
    grid.GRIDCUDAKNNCompact(qx.data(), qy.data(), qz.data(), (int)qx.size(), 1, 1e8f);
    std::vector<int> nn((size_t)grid.pointsQuerCount() * grid.NeighborCount());
    grid.GetKNNIndexes(nn.data());
    int** rows = grid.GetKNNIndexesInPartitions();
    (void)rows;
    grid.GRIDCUDAKNNSELF_COMPACT(40, 1e8f);
    EGT_CUDA smoother(grid);
    smoother.PerformSmoothing(10, 0.63f, -0.64f, 0.65f);
}
