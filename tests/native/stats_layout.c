/* Prints the size of egt_stats and the offset of each field, one "name offset" line each, so a test can
 * compare the C layout with the ctypes mirror in multigpuegtsmoothing_amd/egt.py. */
#include <stddef.h>
#include <stdio.h>
#include "egt.h"

#define F(name) printf(#name " %zu\n", offsetof(egt_stats, name))

int main(void) {
    printf("sizeof %zu\n", sizeof(egt_stats));
    F(n_points); F(k); F(bucket_count); F(invalid_rows); F(knn_candidates); F(egg_kept); F(learned_queries);
    F(learned_fallbacks); F(msf_edges); F(msf_trees); F(msf_weight); F(knn_ms); F(egg_ms); F(taubin_ms);
    F(normals_ms); F(grid_ms); F(propose_ms); F(knn_cells); F(taubin_tile_ms); F(taubin_tile);
    F(taubin_fallback_tiles); F(query_count); F(query_candidates); F(query_invalid_rows); F(query_ms);
    F(query_k); F(taubin_pass_ms); F(taubin_setup_ms); F(taubin_overflow_tiles); F(learned_bound_retries);
    F(knn_insertions); F(knn_max_candidates); F(knn_long_queries);
    return 0;
}
