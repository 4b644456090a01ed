/*
 * gg_nonregular.h -- the evaluator ABI of the reference's non-regular
 * variants, served by two shim libraries over libggeval.so:
 *
 *   libgg_nonregular_general.so  variants/not_regular_general_graph
 *   libgg_nonregular_grid.so     variants/not_regular_grid_graph
 *
 * Both export the symbol `evaluation` (each replaces its variant's
 * evaluation.o), so a program links one of them.  Row v of the adjacency
 * table holds the variant's per-vertex degree (num_degrees[v] / degree[v])
 * neighbours followed by unused slots; the shims fill the unused slots with
 * v itself (a self-loop adds nothing to a reach set) and evaluate on the
 * device.  Results and return conventions are those of the general and grid
 * evaluators respectively.
 */
#ifndef GG_NONREGULAR_H
#define GG_NONREGULAR_H
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef GG_NONREGULAR_GRID
/* Replaces variants/not_regular_grid_graph/evaluation.c:194-208 (declared
 * in that variant's common.h:86).  adjacency is [nodes][max_degree]; row v
 * uses its first degree[v] entries. */
bool evaluation(const int nodes, const int max_degree, const int *degree, const int groups,
                const int *adjacency, const int based_nodes, const int height,
                const int based_height, int *diameter, double *ASPL, const bool enable_bfs,
                const int *rotate_hash);
#else
/* Replaces variants/not_regular_general_graph/evaluation.c:378-396
 * (declared in that variant's common.h:103).  adj is [nodes][degree]; row v
 * uses its first num_degrees[v] entries.  algo as in gg_general.h: 0 = the
 * variant's bfs() (:57-150), 1/2 = its matrix_op. */
bool evaluation(const int nodes, int based_nodes, const int groups, const int lines,
                const int degree, int *adj, int *diameter, double *ASPL, const int added_centers,
                const int *num_degrees, const int algo);
#endif

#ifdef __cplusplus
}
#endif
#endif
