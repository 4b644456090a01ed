/*
 * gg_host.h -- host-side C ABI of libggeval.so: seeded graph construction
 * and the simulated-annealing drivers that call the device evaluator.
 *
 * Graphs are edge lists int32 [lines][2] (lines = nodes*degree/2) and
 * adjacency tables int32 [nodes][degree] in edge order, the reference's
 * formats (general/main.c:112-126, general/sa.c:61-74).
 */
#ifndef GG_HOST_H
#define GG_HOST_H
#include <stdint.h>

#include "gg_eval.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Random d-regular simple graph on n vertices by configuration-model
 * pairing with self-loops and multi-edges rejected (pairs drawn uniformly
 * among the remaining stubs; restart when no admissible pair is left).
 * Connected graphs only (retries with the next seed otherwise).
 * edges: [n*d/2][2].  Returns GG_OK or an error code. */
int gg_random_regular(int nodes, int degree, uint64_t seed, int32_t *edges);

/* Random d-regular simple graph with rotational symmetry v -> v+B (mod n),
 * B = nodes/groups: a g-fold cyclic lift of a random base multigraph with
 * edge voltages.  Edge list ordered orbit-major: line o*groups + i is the
 * i-th rotation of orbit o.  This is the symmetry the reference's `-g`
 * graphs have (general/main.c:175-210). */
int gg_random_symmetric(int nodes, int degree, int groups, uint64_t seed, int32_t *edges);

/* Random grid graph: vertices v = x*height + y on a width x height grid,
 * every vertex of degree `degree`, every edge of Manhattan length <= length
 * (grid/common.c:44-51), connected.  Built by seeded length-limited pairing. */
int gg_random_grid(int width, int height, int degree, int length, uint64_t seed, int32_t *edges);

/* Edge list -> adjacency (general/sa.c:61-74).  Rows of vertices with
 * fewer than `degree` edges are padded with the vertex's own id; returns
 * GG_ERR_INVALID for an out-of-range id or a vertex above `degree`. */
int gg_edges_to_adj(int nodes, int lines, int degree, const int32_t *edges, int32_t *adj);

/* ---- Simulated annealing over 2-opt moves --------------------------------
 *
 * gg_sa_general: the general driver (general/sa.c:413-520 with
 * exchange_edge_2opt general/sa.c:217-369 and accept general/sa.c:371-411),
 * drawing from libc random() exactly as the reference does, so a run seeded
 * with srandom(seed) follows the reference's move sequence.  groups == 1.
 *
 * The evaluator is pluggable: `eval` is called for every proposal; pass
 * NULL to use the device evaluator (gg_apsp).  A per-step log of
 * (diameter, sum) of every evaluated proposal can be recorded. */
typedef int (*gg_eval_fn)(void *ctx, int nodes, int degree, const int32_t *adj, int32_t *diameter,
                          uint64_t *sum, int32_t *connected);

typedef struct gg_sa_params {
  int64_t ncalcs;        /* SA steps */
  double max_temp;       /* general/main.c:327 default 100 */
  double min_temp;       /* default 0.217147 */
  int32_t cooling_cycle; /* default 1 */
  int32_t hill_climbing; /* -H */
  uint32_t seed;         /* srandom(seed) */
  int32_t log_capacity;  /* entries available in log_* arrays (0 = no log) */
} gg_sa_params;

typedef struct gg_sa_stats {
  int64_t steps;
  int64_t accepts;
  int64_t evaluations;
  int32_t best_diameter;
  int32_t pad;
  double best_aspl;
  double seconds_total;
  double seconds_eval;
} gg_sa_stats;

int gg_sa_general(int nodes, int degree, int32_t *edges, const gg_sa_params *params,
                  gg_eval_fn eval, void *eval_ctx, gg_sa_stats *stats, int32_t *log_diam,
                  uint64_t *log_sum);

#ifdef __cplusplus
}
#endif
#endif
