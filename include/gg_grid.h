/*
 * gg_grid.h -- the reference's grid-graph evaluator ABI, served by
 * libgg_grid.so (a thin shim over libggeval.so).
 */
#ifndef GG_GRID_H
#define GG_GRID_H
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces grid/evaluation.c:192-209; declared at grid/common.h:88-90.
 * rotate_hash is the permutation built by create_rotate_hash()
 * (grid/common.c:312-341).  Connected: returns true with *diameter and
 * *ASPL set.  Disconnected: returns false, *diameter = nodes+1 for the
 * matrix path (grid/evaluation.c:173,182) or the largest eccentricity for
 * enable_bfs (grid/evaluation.c:95,112); *ASPL is left untouched. */
bool evaluation(const int nodes, const int degree, const int groups, const int *adjacency,
                const int based_nodes, const int height, const int based_height, int *diameter,
                double *ASPL, const bool enable_bfs, const int *rotate_hash);

#ifdef __cplusplus
}
#endif
#endif
