/*
 * gg_eval.h -- C ABI of the MI355X (gfx950) APSP evaluator, libggeval.so.
 *
 * The evaluator computes, for a degree-regular graph given as an int32
 * adjacency table adj[nodes][degree], the quantities the reference's
 * evaluation() derives from its bit-parallel BFS (general/evaluation.c:153-263,
 * grid/evaluation.c:122-190):
 *   - the diameter (the level at which every source tile reached every vertex),
 *   - S, the integer sum of ordered-pair distances with the reference's
 *     symmetry weights (groups, added centres),
 * bit-exactly.  ASPL is S / (((double)nodes - 1) * nodes), formed by the caller
 * exactly as the reference does.
 *
 * Plain pointers and sizes only; no torch types.  Thread model: one host
 * thread at a time per process (like the reference's evaluation()).
 *
 * The reference-compatible entry points, with the reference's exact
 * signatures, live in gg_general.h and gg_grid.h (separate libraries,
 * because both are named evaluation()).
 */
#ifndef GG_EVAL_H
#define GG_EVAL_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define GG_OK 0
#define GG_ERR_INVALID (-1)   /* bad argument (sizes, ranges, adjacency entry out of range) */
#define GG_ERR_TOO_LARGE (-2) /* nodes beyond what the device kernels support */
#define GG_ERR_HIP (-3)       /* a HIP runtime call failed; see gg_last_error() */
#define GG_ERR_STATE (-4)     /* library not initialised / plan misuse */

/* Graph flags. */
#define GG_GRAPH_SYMMETRIC 1  /* caller guarantees adj is symmetric (u in adj[v] <=> v in adj[u],
                                 with multiplicity); skips the device-side check */
#define GG_GRAPH_BFS 2        /* score with the reference's bfs() (general/evaluation.c:55-150,
                                 algo = BFS): distances along adj from each source s to the
                                 targets i > s, weighted (groups - i/based_nodes), centre targets
                                 x groups, centre sources x 1; "connected" looks at i > s only;
                                 result.sum is that weighted sum and ASPL = sum / (n(n-1)/2).
                                 gg_apsp / gg_apsp_device only; relabel must be NULL */

/* One graph to score.  `adj` is row-major int32 [nodes][degree].  `relabel`
 * (may be NULL) is a permutation of 0..nodes-1 applied to every vertex id
 * before the BFS, as grid/evaluation.c:148-153 does with rotate_hash; the
 * base sources are then the vertices relabelled to 0..based_nodes-1. */
typedef struct gg_graph {
  int32_t nodes;
  int32_t degree;
  int32_t based_nodes;   /* B: sources of the base pass are 0..B-1 */
  int32_t groups;        /* g: rotational symmetry factor (1 = none) */
  int32_t added_centers; /* c: centre vertices nodes-c..nodes-1 (general/evaluation.c:205) */
  int32_t flags;         /* GG_GRAPH_* */
  const int32_t *adj;
  const int32_t *relabel;
} gg_graph;

/* Result of one evaluation. */
typedef struct gg_result {
  int32_t connected;  /* 1 when every source tile reached every vertex */
  int32_t diameter;   /* raw level count (max over tiles of kk+1, at least 1); nodes+1 when
                         disconnected, as the reference's loop leaves it */
  int32_t max_ecc;    /* largest level at which any bit was added (the eccentricity the
                         reference's grid bfs() reports, grid/evaluation.c:95) */
  int32_t word_bits;  /* sources per tile word used (8/16/32/64) */
  uint64_t sum;       /* S = sum over ordered pairs of distances, symmetry weighted
                         (general/evaluation.c:199-200,241-249); meaningful when connected */
} gg_result;

/* Lifecycle.  gg_init(device) selects the HIP device (-1: $GG_DEVICE or 0) and
 * creates the stream; it is idempotent.  The reference's counterpart is
 * init_matrix_dev()/finalize_matrix_dev() (general/evaluation.cu:178-203). */
int gg_init(int device);
void gg_finalize(void);
const char *gg_last_error(void);
int gg_last_error_code(void); /* the GG_ERR_* code that came with gg_last_error() */
int gg_device_count(void);

/* One evaluation, adjacency in host memory (copied to the device each call,
 * like general/evaluation.cu:83).  Replaces the body of evaluation()
 * (general/evaluation.c:375-391).  Graphs of up to 65536 vertices run the
 * LDS-resident tile kernels; larger ones (any n with n*degree < 2^32; the
 * reference's matrix_op_mem_saving range, general/main.c:411-417) run the
 * HBM-resident kernels of gg_wide.hip, as does GG_GRAPH_BFS unless the
 * table is symmetric and rotation invariant (then S/2 is that sum). */
int gg_apsp(const gg_graph *g, gg_result *out);

/* One evaluation whose adjacency (and relabel) pointers are DEVICE pointers
 * (already resident in HBM).  Same contract as gg_apsp; a relabel table is
 * read back and checked to be a permutation. */
int gg_apsp_device(const gg_graph *g, gg_result *out);

/* gg_apsp_device without any allocation: no counterpart in the reference;
 * for callers that cannot allocate (code that runs under HIP graph capture
 * of its own work, tight serving loops, harnesses that forbid hipMalloc).
 * Same inputs, result and errors as gg_apsp_device for graphs of at most
 * 65536 vertices and 32 neighbours per vertex (larger ones, and
 * GG_GRAPH_BFS, are refused with GG_ERR_TOO_LARGE / GG_ERR_INVALID).  All
 * intermediate buffers live in a 5 MiB device array of the library that
 * each GPU holds for the life of the process; a device relabel table is
 * checked on the device.  Calls on one GPU share that array: calls that
 * overlap on one GPU (other host threads) race.  gg_init() beforehand
 * creates the stream and looks up the array's address; the first call does
 * it otherwise. */
int gg_apsp_device_noalloc(const gg_graph *g, gg_result *out);

/* Score `count` graphs of any sizes in one set of launches (host adjacency).
 * out[i] corresponds to graphs[i].  No reference counterpart: the batched
 * scorer for simulated-annealing proposals. */
int gg_eval_batch(int count, const gg_graph *graphs, gg_result *out);

/* ---- Batched plans: device buffers + a captured hipGraph, replayed per batch.
 * A plan is created for a fixed list of graph shapes; the caller writes each
 * slot's adjacency (host upload, or device pointer for device-side producers)
 * and replays. */
typedef struct gg_plan gg_plan;
gg_plan *gg_plan_create(int count, const gg_graph *shapes);
int gg_plan_set_adj(gg_plan *p, int index, const int32_t *adj_host);
int32_t *gg_plan_adj_device(gg_plan *p, int index); /* device buffer of slot `index` */
int gg_plan_run(gg_plan *p, gg_result *out);        /* replay + wait + results */
int gg_plan_launch(gg_plan *p);                     /* replay only (async) */
int gg_plan_wait(gg_plan *p, gg_result *out);       /* wait for the last launch */
void gg_plan_destroy(gg_plan *p);
/* Patched plans: `count` proposals that all share one base graph (shape
 * `*shape`, which must carry GG_GRAPH_SYMMETRIC) and differ from it in at most
 * `max_patch` adjacency entries each.  The base table lives on the device
 * (gg_plan_set_base / gg_plan_base_device); patches are int32
 * [count][max_patch][2] pairs (v*degree + j, new neighbour), index -1 = unused.
 * The prep kernel applies them while converting, so proposals cost no
 * per-proposal adjacency transfer. */
gg_plan *gg_plan_create_patched(int count, const gg_graph *shape, int max_patch);
int gg_plan_set_base(gg_plan *p, const int32_t *adj_host);
int32_t *gg_plan_base_device(gg_plan *p);
int gg_plan_set_patches(gg_plan *p, const int32_t *patches_host);
/* Enqueues the plan's kernels directly (not through the graph) with HIP
 * events around the tile-BFS launches; complete it with gg_plan_wait(). */
int gg_plan_launch_timed(gg_plan *p);
/* Duration (ms, HIP events on the evaluator's stream) of the tile-BFS
 * kernels in the plan's last completed gg_plan_launch_timed(); -1 after a
 * graph launch (gg_plan_launch / gg_plan_run carry no events). */
float gg_plan_kernel_ms(gg_plan *p);
/* Waits for all work of the evaluator's stream and the device. */
int gg_synchronize(void);

/* ---- Tuning.  A launch configuration is (word bits 8/16/32/64, threads per
 * block 0 = from n / 256 / 1024, row-order mode 0 off / 1 by size / 2 on,
 * persistent: one queue-driven launch whose workgroups take the (graph, tile)
 * units from a device counter).  Configurations are kept
 * per shape bucket (ceil(log2 n) of the graph, or of a batch's largest graph;
 * degree; groups; centres or not; batch or single) and consulted by every
 * evaluation and plan made afterwards; without an entry the built-in choice
 * applies.  gg_set_word_bits(b) forces the word width (0 restores the
 * automatic choice); gg_set_row_order (0 or 2) wins over the table's order.
 * config[4] = {bits, block, row order, persistent}. */
int gg_set_word_bits(int bits);
/* Times every candidate on a seeded graph of the shape (`batch` copies in one
 * plan) and records the fastest; all candidates must agree bit for bit. */
int gg_autotune(int nodes, int degree, int based_nodes, int groups, int batch, int reps,
                int *best_bits, double *best_ms);
/* The same on the caller's `count` graphs (host adjacency), recorded under the
 * bucket of the largest; config may be NULL. */
int gg_autotune_batch(int count, const gg_graph *graphs, int reps, int32_t *config, double *best_ms);
/* The configuration an evaluation (batched = 0) or a batch whose largest graph
 * has this shape (batched = 1) would use now. */
int gg_tune_lookup(int nodes, int degree, int based_nodes, int groups, int added_centers, int batched,
                   int32_t *config);
/* The configuration a plan was built with (slot 0's word bits and block). */
int gg_plan_config(gg_plan *p, int32_t *config);
int gg_tune_table_load(const char *path); /* returns entries loaded, or an error code */
int gg_tune_table_save(const char *path);
int gg_tune_table_clear(void);

/* ---- Per-vertex distance features (eccentricity / distance histogram) from
 * row popcounts of the same bit-parallel BFS.  hist is [nodes][maxd+1] int32:
 * hist[v][k] = vertices at distance exactly k from v (k < maxd), hist[v][maxd]
 * = vertices at distance >= maxd.  Adjacency in host memory. */
int gg_distance_histogram(int nodes, int degree, const int32_t *adj, int maxd, int32_t *hist);
/* The same evaluation returning per-vertex features for ranking models (no
 * reference counterpart): hist as above (may be NULL), ecc[v] = v's
 * eccentricity (maxd when some vertex is at distance >= maxd, -1 when v does
 * not reach every vertex; may be NULL) and the evaluation's result in *out
 * (may be NULL; all n vertices are sources). */
int gg_eval_features(int nodes, int degree, const int32_t *adj, int maxd, int32_t *ecc,
                     int32_t *hist, gg_result *out);

/* ---- Profiling aid: one evaluation that records, for the first `units`
 * source tiles, per BFS level k (k < 64): [0] the shader-clock cycles since
 * the tile started, at the level's end; [1] a mode word (bit 0: push level;
 * bits 1..: queue length after a push level, or the reached-pair count after
 * a pull level); [2] the cycles when wave 0 had issued the level's reads;
 * [3] the cycles when the barrier after the reads released:
 * trace[(unit*64 + k)*4 + {0,1,2,3}].  Record k = 0 holds where the unit ran
 * ([1]: XCC_ID << 32 | HW_ID hardware registers) and the shader clock at
 * its start ([2]) and at level 1's start ([3]). */
int gg_trace_levels(const gg_graph *g, int units, uint64_t *trace, gg_result *out);

/* ---- Bank-aware row order (no counterpart in the reference, which gathers
 * from host memory).  Evaluations permute the entries of every converted
 * adjacency row (degree <= 32) so that the 32 rows a half-wave gathers
 * together spread each entry slot over the LDS banks; results do not depend
 * on row order.  gg_set_row_order(mode) applies to plans and evaluations
 * made afterwards: 0 off, 1 graphs of at least 1024 source tiles (the
 * default; where the ordering pays for itself) and incremental scorers of
 * at least 8192 vertices, 2 every graph; returns the previous mode.  gg_row_order converts one graph as an
 * evaluation does and orders it for `word_bits`-wide reach words, returning
 * the row table rows[nodes][dp] (dp = degree rounded up to 8, padding = the
 * row's own id) and the conflict cost of the given and of the returned
 * order: cost[k] = sum over 32-row groups and slots of the largest number
 * of distinct LDS words that one bank serves (bank = word index mod 32,
 * word index = id*bits/32 for 8/16/32-bit words, the id for 64-bit ones). */
int gg_set_row_order(int mode);
int gg_row_order(int nodes, int degree, const int32_t *adj, int word_bits, int32_t *rows,
                 uint64_t cost[2]);

/* ---- Bank-balanced layout (no counterpart in the reference).  Computes
 * once for a graph (nodes <= 65536, degree <= 32) a numbering and an order
 * of each row's entries under which the pull levels' gathers meet few LDS
 * bank conflicts with `word_bits`-wide reach words.  Every numbering gives
 * the same diameter and S.  Vertices trade ids only with vertices whose
 * reach words share their bank, chosen by simulated annealing so that the
 * 32 rows a half-wave gathers together hold about equally many entries on
 * every bank; the entries of each 32-row group are then ordered slot by
 * slot by bipartite matching of rows to banks.
 *   relabel[v]  (out) new id of v (pass as gg_graph.relabel);
 *   adj_out     (out) [nodes][degree]: adj with each row's entries in slot
 *               order (the same graph);
 *   cost        (out) conflict cost, as gg_row_order defines it, of the given
 *               table and of the layout.
 * iterations: annealing swaps (0: 80 per vertex); threads: host threads
 * (0: all).  Evaluate the result with row order mode 0 (gg_set_row_order)
 * to keep the order, or let the tuner choose. */
int gg_bank_layout(int nodes, int degree, const int32_t *adj, int word_bits, int64_t iterations,
                   uint32_t seed, int threads, int32_t *relabel, int32_t *adj_out,
                   uint64_t cost[2]);

#ifdef __cplusplus
}
#endif
#endif
