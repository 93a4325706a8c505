// egt.hpp -- header-only C++ mirror of the reference's provider / consumer classes over the C ABI in egt.h,
// so that code written against the reference (its main.cpp and the like) compiles unchanged against the
// gfx950 library with one GPU:
//
//   KNNInterface                 KNNInterface.cuh:8-31
//   GridStructure                GridKNNCuda.cuh:13-43 (constructor, GRIDCUDAKNNSELF_COMPACT,
//                                GRIDCUDAKNNCompact, accessors)
//   GridMeshCuda                 GridMeshCuda1.cuh:13-30 (constructor, GetNeighborsFromTriangles, accessors)
//   EGT_CUDA                     EGT_CUDA.cuh:11-15
//   Normal_CUDA                  Normal_CUDA.cuh:10-14
//   getOutputPoints              getOutputPoints.cuh:10
//   Mesh, V3, I3, C3, C4         Mesh.h:9-50
//
// Errors: the reference prints and carries on or aborts; here every failing library call throws
// egt::Error (status code + message). num_cards must be 1.
#ifndef EGT_HPP
#define EGT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "egt.h"

struct V3 {
    float x;
    float y;
    float z;
};
struct I3 {
    uint32_t v0;
    uint32_t v1;
    uint32_t v2;
};
struct C3 {
    uint8_t r;
    uint8_t g;
    uint8_t b;
};
struct C4 {
    uint8_t a;
    uint8_t r;
    uint8_t g;
    uint8_t b;
};

struct Mesh {
    V3* VData() { return mV.data(); }
    I3* FData() { return mF.data(); }
    V3* NData() { return mN.data(); }
    C4* CData() { return mC.data(); }
    size_t nV() const { return mV.size(); }
    size_t nF() const { return mF.size(); }
    size_t nN() const { return mN.size(); }
    size_t nC() const { return mC.size(); }
    void clear() {
        mV.clear();
        mF.clear();
        mN.clear();
        mC.clear();
    }
    std::vector<V3> mV;
    std::vector<I3> mF;
    std::vector<V3> mN;
    std::vector<C4> mC;
};

namespace egt {

class Error : public std::runtime_error {
public:
    Error(egt_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
    egt_status status;
};

inline void check(egt_status st) {
    if (st != EGT_OK) {
        const char* m = egt_last_error_string();
        throw Error(st, m ? m : "egt error");
    }
}

// SoA copies of a Mesh's vertices (the library takes x, y, z arrays).
struct SoA {
    std::vector<float> x, y, z;
    explicit SoA(const Mesh& m) : x(m.nV()), y(m.nV()), z(m.nV()) {
        for (size_t i = 0; i < m.nV(); ++i) {
            x[i] = m.mV[i].x;
            y[i] = m.mV[i].y;
            z[i] = m.mV[i].z;
        }
    }
};

}  // namespace egt

class KNNInterface {
public:
    virtual ~KNNInterface() {
        if (h_) egt_destroy(h_);
    }
    KNNInterface(const KNNInterface&) = delete;
    KNNInterface& operator=(const KNNInterface&) = delete;

    virtual bool isReady() const { return h_ != nullptr; }
    virtual int GetNumberOfCards() const { return 1; }
    virtual float** GetRefPointsX() {
        refresh_points();
        return &px_;
    }
    virtual float** GetRefPointsY() {
        refresh_points();
        return &py_;
    }
    virtual float** GetRefPointsZ() {
        refresh_points();
        return &pz_;
    }
    virtual float** GetQuerPointsX() { return GetRefPointsX(); }
    virtual float** GetQuerPointsY() { return GetRefPointsY(); }
    virtual float** GetQuerPointsZ() { return GetRefPointsZ(); }
    virtual int** GetKNNIndexesInPartitions() {
        int32_t* p = nullptr;
        egt::check(egt_get_device_knn(h_, &p));
        knn_ = p;
        return &knn_;
    }
    virtual int GetPartitionSize() const { return pointsRefCount(); }
    virtual int* GetPointCountInCards() {
        count_ = pointsRefCount();
        return &count_;
    }
    virtual int** GetOriginalRefIndexes() {
        int32_t* p = nullptr;
        egt::check(egt_get_device_original_indexes(h_, &p));
        orig_ = p;
        return &orig_;
    }
    virtual int** GetOriginalQuerIndexes() { return GetOriginalRefIndexes(); }
    virtual void GetKNNIndexes(int* knnIndexes) { egt::check(egt_get_knn_original(h_, knnIndexes)); }
    virtual float GetMaxExtent() const {
        float mn = 0, mx = 0;
        egt::check(egt_get_extent(h_, &mn, &mx));
        return mx;
    }
    virtual float GetMinExtent() const {
        float mn = 0, mx = 0;
        egt::check(egt_get_extent(h_, &mn, &mx));
        return mn;
    }
    virtual int pointsRefCount() const {
        int64_t n = 0;
        egt::check(egt_get_counts(h_, &n, nullptr, nullptr, nullptr));
        return (int)n;
    }
    virtual int pointsQuerCount() const { return pointsRefCount(); }
    virtual int NeighborCount() const {
        int32_t k = 0;
        egt::check(egt_get_counts(h_, nullptr, &k, nullptr, nullptr));
        return k;
    }

    // Not in the reference: the library handle behind this provider.
    egt_handle handle() const { return h_; }

protected:
    KNNInterface() = default;
    egt_handle h_ = nullptr;

private:
    void refresh_points() { egt::check(egt_get_device_points(h_, &px_, &py_, &pz_)); }
    float *px_ = nullptr, *py_ = nullptr, *pz_ = nullptr;
    int* knn_ = nullptr;
    int* orig_ = nullptr;
    int count_ = 0;
};

class GridStructure : public KNNInterface {
public:
    GridStructure(const Mesh& mesh, int num_cards, float eps = 1e-6, int bucketCount = 1024, int maxDivision = 32) {
        egt::SoA s(mesh);
        egt::check(egt_grid_create(s.x.data(), s.y.data(), s.z.data(), (int64_t)mesh.nV(), num_cards, eps,
                                   bucketCount, maxDivision, &h_));
    }
    void GRIDCUDAKNNSELF_COMPACT(int k, float thres) {
        egt::check(egt_grid_knn_self(h_, k, thres));
        query_ = false;
    }
    // GRIDCUDAKNNCompact GridKNNCuda.cuh:19: k nearest reference points of each query (host arrays, input
    // units). Afterwards the row accessors describe the query rows, as in the reference.
    void GRIDCUDAKNNCompact(float* pointsQ_x, float* pointsQ_y, float* pointsQ_z, int pointsCountQuery, int k,
                            float thres) {
        egt::check(egt_grid_knn_query(h_, pointsQ_x, pointsQ_y, pointsQ_z, pointsCountQuery, k, thres));
        query_ = true;
    }
    void GetKNNIndexes(int* knnIndexes) override {
        if (query_) egt::check(egt_get_knn_query_original(h_, knnIndexes));
        else KNNInterface::GetKNNIndexes(knnIndexes);
    }
    int** GetKNNIndexesInPartitions() override {
        if (!query_) return KNNInterface::GetKNNIndexesInPartitions();
        int32_t* p = nullptr;
        egt::check(egt_get_device_query_knn(h_, &p, nullptr, nullptr));
        qknn_ = p;
        return &qknn_;
    }
    int pointsQuerCount() const override {
        if (!query_) return pointsRefCount();
        int64_t nq = 0;
        egt::check(egt_get_device_query_knn(h_, nullptr, &nq, nullptr));
        return (int)nq;
    }
    int NeighborCount() const override {
        if (!query_) return KNNInterface::NeighborCount();
        int32_t k = 0;
        egt::check(egt_get_device_query_knn(h_, nullptr, nullptr, &k));
        return k;
    }

private:
    bool query_ = false;
    int* qknn_ = nullptr;
};

class GridMeshCuda : public KNNInterface {
public:
    GridMeshCuda(const Mesh& mesh, int num_cards, float eps = 1e-6, int bucketCount = 1024) {
        egt::SoA s(mesh);
        const uint32_t* tri = mesh.nF() ? &mesh.mF[0].v0 : nullptr;
        egt::check(egt_mesh_create(s.x.data(), s.y.data(), s.z.data(), (int64_t)mesh.nV(), tri, (int64_t)mesh.nF(),
                                   num_cards, eps, bucketCount, &h_));
    }
    void GetNeighborsFromTriangles(int k) { egt::check(egt_mesh_neighbors(h_, k)); }
};

class EGT_CUDA {
public:
    explicit EGT_CUDA(KNNInterface& knnInterface) : p_(knnInterface) {}
    void PerformSmoothing(int iterationCount, float lambda = 0.63f, float mu = -0.64f, float alpha = 1e-8f) {
        egt::check(egt_smooth(p_.handle(), iterationCount, lambda, mu, alpha));
    }
    virtual ~EGT_CUDA() = default;

private:
    KNNInterface& p_;
};

class Normal_CUDA {
public:
    explicit Normal_CUDA(KNNInterface& knnInterface) : p_(knnInterface) {}
    void GetNormals(float* normal_x, float* normal_y, float* normal_z) {
        egt::check(egt_normals(p_.handle(), normal_x, normal_y, normal_z));
    }

private:
    KNNInterface& p_;
};

inline void getOutputPoints(float* pnts_x, float* pnts_y, float* pnts_z, KNNInterface* knnInterface) {
    egt::check(egt_get_output_points(knnInterface->handle(), pnts_x, pnts_y, pnts_z));
}

#endif  // EGT_HPP
