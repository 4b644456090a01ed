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
  int32_t incremental;   /* general SA on the device evaluator: score each proposal
                            with the incremental scorer (include/gg_incr.h; simple graphs,
                            degree <= 32).  Same exact results, so the same trajectory. */
  int32_t detect_temp;   /* -D / -d: record the largest energy increase met
                            (stats->max_diff_energy); the grid SA then accepts every
                            proposal by the Metropolis test alone (grid/sa.c:231-233) */
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
  double max_diff_energy; /* with detect_temp: max over tested proposals of -diff */
} gg_sa_stats;

int gg_sa_general(int nodes, int degree, int32_t *edges, const gg_sa_params *params,
                  gg_eval_fn eval, void *eval_ctx, gg_sa_stats *stats, int32_t *log_diam,
                  uint64_t *log_sum);

/* Returned by the *_program entry points when the start graph already meets
 * the lower bound: nothing is run, stats hold the start graph's values
 * ("The input file has already optimum solution.", general/sa.c:568-569). */
#define GG_ALREADY_OPTIMAL 1

/* The general program's whole run after its argument checks, groups = 1
 * (general/main.c:360-462), on one libc random() stream seeded with
 * params->seed: unless `halfway` (-M), 100 random 2-opt moves and then moves
 * until the graph is connected (create_symmetric_edge, general/main.c:175-210);
 * the start-graph check (check_current_edge, general/sa.c:557-572); the five
 * timing moves on a copy (estimate_elapse_time, general/sa.c:523-554; their
 * evaluations are skipped, they do not touch the stream); then sa() as
 * gg_sa_general without re-seeding.  `edges` must be a simple d-regular graph
 * (GG_ERR_INVALID otherwise); on return it holds the best graph, as the
 * reference's -o file does. */
int gg_general_program(int nodes, int degree, int32_t *edges, int halfway,
                       const gg_sa_params *params, gg_eval_fn eval, void *eval_ctx,
                       gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum);

/* ---- General SA with groups and added centres (general/sa.c:217-520,
 * general/common.c:103-276): rotationally symmetric graphs (vertex v and
 * v + based_nodes*k are the same vertex in group k; centres are the last
 * added_centers ids), 2g moves and 1g moves as the reference draws them. */
typedef struct gg_general_shape {
  int32_t nodes;          /* all vertices, centres included */
  int32_t degree;
  int32_t groups;         /* -g */
  int32_t added_centers;  /* -v */
  int32_t not_simple;     /* -N: repeated edges allowed; the adjacency is rebuilt
                             after every move and moves are undone from a copy */
} gg_general_shape;

/* Evaluator callback with the symmetry inputs (based_nodes =
 * (nodes - added_centers) / groups). */
typedef int (*gg_eval_sym_fn)(void *ctx, int nodes, int degree, int based_nodes, int groups,
                              int added_centers, const int32_t *adj, int32_t *diameter,
                              uint64_t *sum, int32_t *connected);

/* sa() after srandom(params->seed); eval NULL = device evaluator. */
int gg_sa_general_sym(const gg_general_shape *shape, int32_t *edges, const gg_sa_params *params,
                      gg_eval_sym_fn eval, void *eval_ctx, gg_sa_stats *stats, int32_t *log_diam,
                      uint64_t *log_sum);

/* The general program's run (general/main.c:360-462) for any groups and
 * centres.  edges: room for all nodes*degree/2 lines; source:
 *   GG_GENERAL_FULL_FILE     -M: edges hold the whole graph;
 *   GG_GENERAL_BASE_FILE     the first lines/groups lines are the base graph:
 *                            rotated copies, 100 random moves, moves until
 *                            connected (create_symmetric_edge);
 *   GG_GENERAL_BASE_CENTRES  -v: the first file_lines lines are the file's;
 *                            centre edges and the rest of the base lines are
 *                            laid down, then the copies
 *                            (create_symmetric_edge_with_vertexes);
 * then the degree / self-loop checks, the start-graph check (groups = 1), the
 * five timing moves' draws, and sa(). */
#define GG_GENERAL_FULL_FILE 0
#define GG_GENERAL_BASE_FILE 1
#define GG_GENERAL_BASE_CENTRES 2
int gg_general_program_sym(const gg_general_shape *shape, int32_t *edges, int source,
                           int file_lines, const gg_sa_params *params, gg_eval_sym_fn eval,
                           void *eval_ctx, gg_sa_stats *stats, int32_t *log_diam,
                           uint64_t *log_sum);

/* ---- Grid SA (grid/sa.c:283-361), same RNG draws (libc random() after
 * srandom(seed)) and moves as the reference: length-limited 2-opt with
 * groups-fold rotational symmetry (exchange_edge, grid/sa.c:137-221;
 * edge_1g_opt, :34-135), disconnected proposals redrawn, accept
 * grid/sa.c:223-268, stop at the lower bound (grid/main.c:480-529). */
typedef struct gg_grid_shape {
  int32_t width, height; /* nodes = width * height, vertex v at (v / height, v % height) */
  int32_t groups;        /* 1, 2 or 4 */
  int32_t length;        /* max Manhattan length of an edge (-r) */
} gg_grid_shape;

/* Grid evaluator callback: like gg_eval_fn plus the symmetry inputs. */
typedef int (*gg_eval_grid_fn)(void *ctx, int nodes, int degree, int groups, int based_nodes,
                               const int32_t *adj, const int32_t *rotate_hash, int32_t *diameter,
                               uint64_t *sum, int32_t *connected);

/* A groups-fold symmetric grid graph (create_symmetric_edge,
 * grid/main.c:393-445): a seeded base lattice (as gg_random_grid) on the
 * base sub-grid, its rotated copies, lines*10 symmetric exchanges, then
 * exchanges until connected (libc random() after srandom(seed)). */
int gg_random_grid_symmetric(int width, int height, int degree, int length, int groups,
                             uint64_t seed, int32_t *edges);

/* The same from a given base-grid edge list: base[] holds the lines/groups
 * edges of the base sub-grid (width x height/2 for g = 2, width/2 x height/2
 * for g = 4, vertex v at (v / base_height, v % base_height)), as read from a
 * -f file (grid/main.c:393-445 with srandom(seed)). */
int gg_grid_symmetric_from_base(int width, int height, int degree, int length, int groups,
                                uint32_t seed, const int32_t *base, int32_t *edges);

/* create_lattice (grid/main.c:232-330) after srandom(seed): the reference's
 * random length-limited lattice on a width x height grid, vertex v at
 * (v / height, v % height). */
int gg_grid_lattice(int width, int height, int degree, int length, uint32_t seed, int32_t *edges);

/* lower_bound_of_diam_aspl (grid/main.c:480-529). */
int gg_grid_lower_bound(int width, int height, int degree, int length, int *low_diam,
                        double *low_aspl);

/* params: as for gg_sa_general; fixed_temp > 0 uses that temperature in the
 * Metropolis test (-F).  rotate_hash: create_rotate_hash's permutation
 * (grid/common.c:312-341), NULL for groups = 1.  eval NULL = device
 * evaluator.  On return `edges` holds the best graph. */
int gg_sa_grid(const gg_grid_shape *shape, int degree, int32_t *edges, const int32_t *rotate_hash,
               const gg_sa_params *params, double fixed_temp, gg_eval_grid_fn eval, void *eval_ctx,
               gg_sa_stats *stats, int32_t *log_diam, uint64_t *log_sum);

/* The grid program's whole run after its argument checks (grid/main.c:630-727)
 * on one libc random() stream seeded with params->seed.  source:
 *   GG_GRID_LATTICE    no -f: create_lattice on the base grid, written to edges;
 *   GG_GRID_BASE_FILE  -f without -M: edges[0 .. lines/groups) hold the base
 *                      grid's edges (base coordinates, as read from the file);
 *   GG_GRID_FULL_FILE  -f with -M: edges hold the whole graph.
 * then, unless GG_GRID_FULL_FILE or groups = 1, create_symmetric_edge
 * (grid/main.c:393-445); verfy_graph's degree and length checks; the
 * start-graph check; the five timing moves, each on a fresh copy
 * (estimated_elapse_time, grid/sa.c:350-376, evaluations skipped); then sa()
 * as gg_sa_grid without re-seeding.  edges: [lines][2]; on return the best
 * graph. */
#define GG_GRID_LATTICE 0
#define GG_GRID_BASE_FILE 1
#define GG_GRID_FULL_FILE 2
int gg_grid_program(const gg_grid_shape *shape, int degree, int32_t *edges, int source,
                    const int32_t *rotate_hash, const gg_sa_params *params, double fixed_temp,
                    gg_eval_grid_fn eval, void *eval_ctx, gg_sa_stats *stats, int32_t *log_diam,
                    uint64_t *log_sum);

/* ---- Batched, surrogate-ranked simulated annealing (no reference
 * counterpart; the accept rule is the reference's, general/sa.c:371-411).
 *
 * Every step scores `exact` candidate graphs in one patched plan launch.
 * A candidate is the current graph plus one or more vertex-disjoint 2-opt
 * moves.  Explore steps (the first `warmup` steps, every `explore_every`-th
 * step after them, and every step without the surrogate) score random single
 * moves.  Ranked steps draw `proposals` random moves, compute their features
 * on the device (per-vertex distance statistics of the four endpoints and
 * 1-/2-hop ball overlaps of the removed and added pairs), rank them with the
 * MLP (include/gg_mlp.h), and merge the best-ranked vertex-disjoint moves,
 * up to `combine` per candidate, into each candidate.  The best candidate
 * then faces the reference's accept rule: smaller diameter wins; equal
 * diameter accepts a non-increasing S, else Metropolis exp(-dS/T).  Scored
 * singles train the MLP on their exact dS; a scored combination trains its
 * members on an equal share of its prediction error.  exact = 1 without the
 * surrogate is plain single-graph GPU SA on a device-resident adjacency.
 * Simple d-regular graphs, groups = 1, no added centres. */
typedef struct gg_bsa_params {
  int64_t max_steps;
  double time_budget_s;   /* stop after this wall time (0: none) */
  double target_aspl;     /* stop once best ASPL <= target (0: none) */
  int32_t proposals;      /* moves ranked per ranked step */
  int32_t exact;          /* candidate graphs scored per step */
  int32_t combine;        /* max moves merged into one candidate (1..128) */
  int32_t warmup;         /* initial explore steps */
  int32_t explore_every;  /* every k-th later step explores (0: never) */
  int32_t use_surrogate;
  int32_t random_ranking; /* ablation: rank ranked-step proposals at random
                             (the MLP still trains but does not choose) */
  double max_temp, min_temp; /* geometric schedule over max_steps */
  uint64_t seed;
  int32_t train_steps;    /* Adam steps per SA step */
  int32_t batch_train;    /* rows per Adam step */
  float lr;
  int32_t hist_maxd;      /* distance histogram depth for the statistics (<= 15) */
  int32_t feature_refresh;/* recompute per-vertex statistics every k accepts */
  int32_t log_capacity;   /* entries in the log arrays */
  int32_t grid_height;    /* > 0: a grid graph, vertex v at (v / height, v % height) */
  int32_t grid_length;    /* grid: max Manhattan length of every edge (grid/sa.c:196-201) */
  int32_t incremental;    /* score single-move candidates with the incremental scorer
                             (include/gg_incr.h; simple graphs, degree <= 32) -- same
                             exact results, faster at large n.  Without the surrogate:
                             every step; with it: the explore steps up to the first
                             ranked step (merged candidates flag nearly every source) */
  int32_t groups;         /* > 1: a rotationally symmetric graph (general: v -> v + n/groups;
                             grid: 2 or 4 groups, ROTATE of grid/common.c:63-86); every move
                             is the 2g-edge 2-opt of exchange_edge_2opt (general/sa.c:217-369,
                             grid/sa.c:137-221) on two lines and their rotated copies, drawn on
                             the device; scored from the base sources.  0 or 1: none */
} gg_bsa_params;

typedef struct gg_bsa_stats {
  int64_t steps, accepts, evaluations;
  int32_t best_diameter, pad;
  double best_aspl;
  double seconds_total;
  double seconds_exact;     /* candidate scoring: patch upload + plan replay (the
                               MLP trains on its own stream meanwhile) */
  double seconds_surrogate; /* feature kernel, MLP ranking, training-row storage */
  double seconds_features;  /* host: proposal draws, candidate assembly, statistics refresh */
  double seconds_to_target; /* -1 when the target was not reached */
  double surrogate_rank_corr; /* Spearman, prediction vs exact dS of explore singles (held out) */
  int64_t moves_accepted;     /* 2-opt moves in accepted candidates */
  double seconds_train;       /* MLP training, overlapped with candidate scoring */
} gg_bsa_stats;

/* log_time[k], log_best_aspl[k], log_best_diam[k]: after step k.
 * accepted[k*3 + {0,1,2}] = (line0, line1, variant) of the k-th accepted move
 * in application order; accepted_diam[k] / accepted_sum[k] are the exact
 * diameter and S of the graph after it when it ends an accepted candidate,
 * and -1 / 0 for the earlier members of a merged candidate, so the run can be
 * replayed from the initial edges and every accepted graph re-checked.
 * On return `edges` holds the best graph found. */
int gg_sa_batched(int nodes, int degree, int32_t *edges, const gg_bsa_params *params,
                  gg_bsa_stats *stats, double *log_time, double *log_best_aspl,
                  int32_t *log_best_diam, int32_t *accepted, int32_t *accepted_diam,
                  uint64_t *accepted_sum, int accepted_capacity);

/* The surrogate's features of `count` moves props[i] = (a, b, c, d, r) of the
 * graph adj (row-major nodes x degree), as gg_sa_batched computes them on the
 * device: feats[i][0..63] (64 floats per move; layout in gg_bsa.hip).  vstat:
 * per-vertex statistics [nodes][10], or NULL for zeros.  Exposed for tests. */
int gg_sa_features(int nodes, int degree, const int32_t *adj, int count, const int32_t *props,
                   const float *vstat, float *feats);

/* Ranked steps draw their proposals on the device: lines t0 != t1 and a
 * variant r from a counter-based stream of (seed, step, proposal), redrawn
 * (up to 64 times) while the lines share a vertex, valid when neither added
 * edge exists.  Grid graphs (grid_height > 0) draw t1 among the lines of a
 * vertex within grid_length of t0's first endpoint and need both added edges
 * within grid_length.  props[i] = (a, b, c, d, r) with (a, b) = line t0 and
 * (c, d) = line t1; lines_out[i] = (t0, t1, valid).  Exposed for tests. */
int gg_sa_proposals(int nodes, int degree, const int32_t *edges, const int32_t *adj, int count,
                    uint64_t seed, uint64_t step, int grid_height, int grid_length,
                    int32_t *props, int32_t *lines_out);

/* Applies one logged move (line0, line1, variant) to an edge list the way
 * gg_sa_batched does (for replays). */
int gg_apply_move(int lines, int32_t *edges, int line0, int line1, int variant);
/* The same for a symmetric graph's move (gg_bsa_params.groups > 1): lines
 * line0, line1 and their rotated copies re-paired with `variant`; the edge
 * list must be invariant under the rotation (grid_height > 0: a grid). */
int gg_apply_move_sym(int nodes, int degree, int groups, int grid_height, int32_t *edges, int line0,
                      int line1, int variant);

#ifdef __cplusplus
}
#endif
#endif
