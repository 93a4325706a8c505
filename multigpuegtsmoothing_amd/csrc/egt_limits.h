// Size limits of the library's 32-bit index spaces (host-only, no HIP: the CPU tests compile it with g++).
//
// The EGG-kept adjacency is a CSR with int32 row offsets (Taubin, PCA and the normals' edge list index it
// with int32), and a mesh's corner lists (two entries per triangle corner) use uint32 offsets. A count past
// these limits would wrap and make the kernels index out of bounds, so the entry points refuse it with
// EGT_ERR_UNSUPPORTED instead.
#pragma once
#include <cstdint>

namespace egt {

// Largest number of kept neighbour slots (CSR entries) of one handle, all clouds together.
constexpr int64_t CSR_MAX_ENTRIES = INT32_MAX;
// Largest triangle count of a mesh handle: 6 corner-list entries per triangle within int32.
constexpr int64_t MESH_MAX_TRIANGLES = INT32_MAX / 6;

// Null when a CSR of `total` entries fits, else the error message.
inline const char* csr_total_error(uint64_t total) {
    return total > (uint64_t)CSR_MAX_ENTRIES
               ? "the kept neighbour slots exceed 2^31 - 1 (int32 CSR offsets): split the point set or lower k"
               : nullptr;
}

// Null when a mesh of ntri triangles fits, else the error message.
inline const char* mesh_triangles_error(int64_t ntri) {
    return ntri > MESH_MAX_TRIANGLES ? "more than (2^31 - 1) / 6 triangles (int32 corner-list offsets)" : nullptr;
}

}  // namespace egt
