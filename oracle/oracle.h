/*
 * oracle.h -- CPU restatement of GraphGolf's APSP evaluator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker / CPU timing reference.
 *
 * Every function states the reference file:line it restates.  Results are
 * integers (diameter, S = sum of ordered-pair distances); ASPL is formed by
 * the caller with the reference's own division.
 */
#ifndef GG_ORACLE_H
#define GG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Outcome of one evaluation.  `reached` mirrors the bool the reference
 * returns; `diameter` is the raw level count before the reference maps a
 * disconnected graph to its sentinels (so tests can see it); `S` is the
 * integer the reference accumulates in a double (exact below 2^53). */
typedef struct {
  int reached;
  int diameter;
  uint64_t S;
} orc_result;

/* Bit-parallel level BFS over 64-source words (general/evaluation.c:153-263
 * with tile_words = 0 meaning "one tile of all words", and :265-373 with
 * tile_words = 64, the memory-saving variant).  Handles groups and added
 * centres exactly as the reference does. */
orc_result orc_matrix_op(int nodes, int based_nodes, int groups, int degree,
                         const int *adj, int added_centers, int tile_words);

/* Scalar per-source BFS with the symmetry weights (general/evaluation.c:55-150).
 * S here is the reference's unordered-pair sum; ASPL = S / (((n-1)*n)/2). */
orc_result orc_bfs_general(int nodes, int based_nodes, int groups, int degree,
                           const int *adj, int added_centers);

/* Grid evaluator, matrix variant (grid/evaluation.c:122-190): gathers through
 * rotate_hash.  reached=0 -> reference returns false with diameter = nodes+1. */
orc_result orc_grid_matrix(int nodes, int degree, int groups, const int *adj,
                           int based_nodes, const int *rotate_hash);

/* Grid evaluator, BFS variant (grid/evaluation.c:64-120). */
orc_result orc_grid_bfs(int nodes, int degree, int groups, const int *adj,
                        int based_nodes, int height, int based_height);

/* Independent check: plain BFS from every vertex, no symmetry, no tiling. */
orc_result orc_apsp_plain(int nodes, int degree, const int *adj);

/* Per-vertex distance histogram from plain BFS: hist[v*(maxd+1)+k] = number
 * of vertices at distance exactly k from v (k<=maxd; farther ones are
 * counted in k=maxd).  Returns the largest eccentricity seen. */
int orc_distance_histogram(int nodes, int degree, const int *adj, int maxd,
                           int32_t *hist);

/* Edge list -> adjacency table in edge order (general/sa.c:61-74,
 * grid/common.c:14-27); short rows padded with the vertex itself.
 * Returns -1 on an out-of-range id or a row above `degree`. */
int orc_create_adj(int nodes, int lines, int degree, const int *edge, int *adj);

/* grid/common.c:63-86 and :312-341. */
int orc_rotate(int v, int height, int width, int groups, int degree);
void orc_rotate_hash(int nodes, int height, int width, int groups, int *hash);

/* Number of OpenMP threads the oracle runs with (1 when built without). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
