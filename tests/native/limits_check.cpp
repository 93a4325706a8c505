// Host check of the 32-bit index-space guards (egt_limits.h): counts at the limit pass, one past it fails.
#include <cstdio>
#include "egt_limits.h"

int main() {
    int bad = 0;
    bad += egt::csr_total_error(0) != nullptr;
    bad += egt::csr_total_error((uint64_t)INT32_MAX) != nullptr;
    bad += egt::csr_total_error((uint64_t)INT32_MAX + 1) == nullptr;
    bad += egt::csr_total_error(2190000000ull) == nullptr;  // 54.7M points x k 40, unfiltered
    bad += egt::csr_total_error(1ull << 36) == nullptr;     // beyond 32 bits: must not wrap into "fits"
    bad += egt::mesh_triangles_error(1310720) != nullptr;   // 8-fold icosphere
    bad += egt::mesh_triangles_error(egt::MESH_MAX_TRIANGLES) != nullptr;
    bad += egt::mesh_triangles_error(egt::MESH_MAX_TRIANGLES + 1) == nullptr;
    bad += 6 * egt::MESH_MAX_TRIANGLES > (int64_t)INT32_MAX;
    std::printf("%d\n", bad);
    return bad;
}
