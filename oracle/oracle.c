/*
 * oracle.c -- CPU restatement of GraphGolf's APSP evaluator.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The algorithms are restated from
 * the reference's C sources; the product never links this file.
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_matrix_op     -- reference binary (oracle/_ref) on fixtures and
 *                        seeded graphs, closed forms (K_n, C_n, Petersen,
 *                        Hoffman-Singleton, Q_k, meshes, tori), plain-BFS
 *                        brute force, tiling invariance, symmetry identity.
 *   orc_bfs_general   -- reference binary; equality with orc_matrix_op on
 *                        g-symmetric graphs.
 *   orc_grid_matrix   -- reference binary (grid build); relabelling identity.
 *   orc_grid_bfs      -- reference binary (grid build).
 *   orc_apsp_plain    -- closed forms above.
 *   orc_distance_histogram -- closed forms (row sums = n, moments = S).
 *   orc_rotate_hash   -- reference binary (grid build) permutation check.
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Sets the OpenMP thread count of later calls (no-op without OpenMP). */
void orc_set_threads(int k) {
#ifdef _OPENMP
  if (k > 0) omp_set_num_threads(k);
#else
  (void)k;
#endif
}

static inline uint64_t pop64(uint64_t x) { return (uint64_t)__builtin_popcountll(x); }

/*
 * One source tile of the level-synchronous bit-parallel BFS
 * (general/evaluation.c:166-201 for the base pass, :212-244 for the centre
 * pass; grid/evaluation.c:135-174 with `perm` = rotate_hash).
 *
 * Row r of A/B holds `words` 64-bit words; bit s%64 of word s/64 says
 * "source s of this tile reaches the vertex stored at row r".  Sources sit at
 * rows first_row .. first_row+l-1.  Rows below `split` are "low" rows, the
 * rest "high" rows (the added centres).  Per level the low/high popcounts are
 * compared with their targets; the base pass stops when every row is full,
 * the centre pass when the high rows are full.  Until then the unreached
 * counts are accumulated into *acc_lo / *acc_hi.
 *
 * Returns kk+1 for the level index kk at which the tile stopped, or
 * nodes+1 when it never stopped (a disconnected graph).
 */
static int tile_levels(int nodes, int degree, const int *adj, const int *perm,
                       int first_row, int l, int words, int split, int centre_pass,
                       uint64_t *A, uint64_t *B, uint64_t *acc_lo, uint64_t *acc_hi) {
  const size_t cells = (size_t)nodes * (size_t)words;
  memset(A, 0, cells * sizeof(uint64_t));
  memset(B, 0, cells * sizeof(uint64_t));
  for (int s = 0; s < l; s++) {
    size_t at = (size_t)(first_row + s) * words + (size_t)(s / 64);
    A[at] = B[at] = 1ULL << (s % 64);
  }
  const uint64_t lo_full = (uint64_t)split * (uint64_t)l;
  const uint64_t hi_full = (uint64_t)(nodes - split) * (uint64_t)l;

  int kk;
  for (kk = 0; kk < nodes; kk++) {
    /* B[row(i)] |= A[row(nbr)] for every neighbour (the gather-OR). */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nodes; i++) {
      uint64_t *dst = B + (size_t)(perm ? perm[i] : i) * words;
      const int *nb = adj + (size_t)i * degree;
      for (int j = 0; j < degree; j++) {
        const uint64_t *src = A + (size_t)(perm ? perm[nb[j]] : nb[j]) * words;
        for (int k = 0; k < words; k++) dst[k] |= src[k];
      }
    }
    uint64_t lo = 0, hi = 0;
    const size_t split_cells = (size_t)split * words;
#pragma omp parallel for reduction(+ : lo) schedule(static)
    for (size_t x = 0; x < split_cells; x++) lo += pop64(B[x]);
#pragma omp parallel for reduction(+ : hi) schedule(static)
    for (size_t x = split_cells; x < cells; x++) hi += pop64(B[x]);

    int done = centre_pass ? (hi == hi_full) : (lo + hi == lo_full + hi_full);
    if (done) break;
    uint64_t *t = A; A = B; B = t;
    if (!centre_pass) *acc_lo += lo_full - lo;
    *acc_hi += hi_full - hi;
  }
  return kk + 1;
}

/* Runs every tile of one pass.  Sources are first_src .. first_src+count-1
 * (as rows), tiled by `tile_words` words (0 = a single tile of all words). */
static int run_pass(int nodes, int degree, const int *adj, const int *perm, int first_src,
                    int count, int tile_words, int split, int centre_pass,
                    uint64_t *acc_lo, uint64_t *acc_hi) {
  int words_all = (count + 63) / 64;
  int words = tile_words > 0 ? tile_words : words_all;
  if (words < 1) words = 1;
  int tiles = (words_all + words - 1) / words;
  size_t bytes = (size_t)nodes * words * sizeof(uint64_t);
  uint64_t *A = (uint64_t *)malloc(bytes), *B = (uint64_t *)malloc(bytes);
  int diam = 1;
  for (int t = 0; t < tiles; t++) {
    int lo = t * 64 * words;
    int l = count - lo < 64 * words ? count - lo : 64 * words;
    int k = tile_levels(nodes, degree, adj, perm, first_src + lo, l, words, split,
                        centre_pass, A, B, acc_lo, acc_hi);
    if (k > diam) diam = k;
  }
  free(A);
  free(B);
  return diam;
}

orc_result orc_matrix_op(int nodes, int based_nodes, int groups, int degree, const int *adj,
                         int added_centers, int tile_words) {
  /* general/evaluation.c:153-263 (tile_words=0) and :265-373 (tile_words=64). */
  orc_result r;
  const int split = based_nodes * groups;
  uint64_t lo = 0, hi = 0, c_lo = 0, c_hi = 0;
  int diam = run_pass(nodes, degree, adj, NULL, 0, based_nodes, tile_words, split, 0, &lo, &hi);
  if (added_centers) {
    int dc = run_pass(nodes, degree, adj, NULL, nodes - added_centers, added_centers,
                      tile_words, split, 1, &c_lo, &c_hi);
    if (dc > diam) diam = dc;
  }
  r.S = lo * (uint64_t)groups + hi * 2ULL * (uint64_t)groups + c_hi +
        (uint64_t)nodes * (uint64_t)(nodes - 1);
  r.diameter = diam;
  r.reached = diam < nodes; /* general/evaluation.c:254 */
  return r;
}

orc_result orc_grid_matrix(int nodes, int degree, int groups, const int *adj, int based_nodes,
                           const int *rotate_hash) {
  /* grid/evaluation.c:122-190: all rows are low rows (no centres). */
  orc_result r;
  uint64_t lo = 0, hi = 0;
  int diam = run_pass(nodes, degree, adj, rotate_hash, 0, based_nodes, 0, nodes, 0, &lo, &hi);
  r.S = lo * (uint64_t)groups + (uint64_t)nodes * (uint64_t)(nodes - 1);
  r.diameter = diam;
  r.reached = !(diam > nodes); /* grid/evaluation.c:182 */
  return r;
}

/* Level-by-level BFS from one source; fills dist (-1 = unreached) and
 * returns the eccentricity over the reached vertices. */
static int bfs_from(int nodes, int degree, const int *adj, int src, int *dist, int *queue) {
  for (int i = 0; i < nodes; i++) dist[i] = -1;
  int head = 0, tail = 0, ecc = 0;
  dist[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    int v = queue[head++];
    const int *nb = adj + (size_t)v * degree;
    for (int j = 0; j < degree; j++) {
      int u = nb[j];
      if (dist[u] < 0) {
        dist[u] = dist[v] + 1;
        if (dist[u] > ecc) ecc = dist[u];
        queue[tail++] = u;
      }
    }
  }
  return ecc;
}

orc_result orc_bfs_general(int nodes, int based_nodes, int groups, int degree, const int *adj,
                           int added_centers) {
  /* general/evaluation.c:55-150: per-source BFS; each unordered pair (s,i),
   * i>s, is weighted by the number of rotated copies it stands for. */
  orc_result r;
  int reached = 1, diam = 0;
  uint64_t sum = 0;
  const int split = based_nodes * groups;
  const int nsrc = based_nodes + added_centers;
#pragma omp parallel reduction(max : diam) reduction(+ : sum) reduction(&& : reached)
  {
    int *dist = (int *)malloc(sizeof(int) * nodes);
    int *queue = (int *)malloc(sizeof(int) * nodes);
#pragma omp for schedule(dynamic, 4)
    for (int q = 0; q < nsrc; q++) {
      int s = q < based_nodes ? q : split + (q - based_nodes);
      int ecc = bfs_from(nodes, degree, adj, s, dist, queue);
      if (ecc > diam) diam = ecc;
      for (int i = s + 1; i < nodes; i++) {
        if (dist[i] < 0) { reached = 0; continue; }
        uint64_t w;
        if (q >= based_nodes) w = 1;                   /* centre source */
        else if (i < split) w = (uint64_t)(groups - i / based_nodes);
        else w = (uint64_t)groups;                      /* centre target */
        sum += (uint64_t)dist[i] * w;
      }
    }
    free(dist);
    free(queue);
  }
  r.reached = reached;
  r.diameter = diam;
  r.S = sum;
  return r;
}

orc_result orc_grid_bfs(int nodes, int degree, int groups, const int *adj, int based_nodes,
                        int height, int based_height) {
  /* grid/evaluation.c:64-120: base-quadrant sources, every other vertex
   * weighted by `groups`. */
  orc_result r;
  int reached = 1, diam = 0;
  uint64_t sum = 0;
#pragma omp parallel reduction(max : diam) reduction(+ : sum) reduction(&& : reached)
  {
    int *dist = (int *)malloc(sizeof(int) * nodes);
    int *queue = (int *)malloc(sizeof(int) * nodes);
#pragma omp for schedule(dynamic, 4)
    for (int s = 0; s < based_nodes; s++) {
      int s1 = (s / based_height) * height + (s % based_height);
      int ecc = bfs_from(nodes, degree, adj, s1, dist, queue);
      if (ecc > diam) diam = ecc;
      for (int i = 0; i < nodes; i++) {
        if (i == s1) continue;
        if (dist[i] < 0) { reached = 0; continue; }
        sum += (uint64_t)dist[i] * (uint64_t)groups;
      }
    }
    free(dist);
    free(queue);
  }
  r.reached = reached;
  r.diameter = diam;
  r.S = sum;
  return r;
}

orc_result orc_apsp_plain(int nodes, int degree, const int *adj) {
  orc_result r;
  int reached = 1, diam = 0;
  uint64_t sum = 0;
#pragma omp parallel reduction(max : diam) reduction(+ : sum) reduction(&& : reached)
  {
    int *dist = (int *)malloc(sizeof(int) * nodes);
    int *queue = (int *)malloc(sizeof(int) * nodes);
#pragma omp for schedule(dynamic, 8)
    for (int s = 0; s < nodes; s++) {
      int ecc = bfs_from(nodes, degree, adj, s, dist, queue);
      if (ecc > diam) diam = ecc;
      for (int i = 0; i < nodes; i++) {
        if (dist[i] < 0) reached = 0;
        else sum += (uint64_t)dist[i];
      }
    }
    free(dist);
    free(queue);
  }
  r.reached = reached;
  r.diameter = diam;
  r.S = sum;
  return r;
}

int orc_distance_histogram(int nodes, int degree, const int *adj, int maxd, int32_t *hist) {
  int diam = 0;
  memset(hist, 0, sizeof(int32_t) * (size_t)nodes * (size_t)(maxd + 1));
#pragma omp parallel reduction(max : diam)
  {
    int *dist = (int *)malloc(sizeof(int) * nodes);
    int *queue = (int *)malloc(sizeof(int) * nodes);
#pragma omp for schedule(dynamic, 8)
    for (int v = 0; v < nodes; v++) {
      int ecc = bfs_from(nodes, degree, adj, v, dist, queue);
      if (ecc > diam) diam = ecc;
      int32_t *h = hist + (size_t)v * (maxd + 1);
      for (int i = 0; i < nodes; i++)
        if (dist[i] >= 0) h[dist[i] < maxd ? dist[i] : maxd]++;
    }
    free(dist);
    free(queue);
  }
  return diam;
}

int orc_create_adj(int nodes, int lines, int degree, const int *edge, int *adj) {
  /* general/sa.c:61-74: fill each vertex's row in edge order; short rows
   * are padded with the vertex itself (no distance changes). */
  int *fill = (int *)calloc((size_t)nodes, sizeof(int));
  int rc = 0;
  for (int e = 0; e < lines && !rc; e++) {
    int a = edge[2 * e], b = edge[2 * e + 1];
    if (a < 0 || a >= nodes || b < 0 || b >= nodes || fill[a] >= degree || fill[b] >= degree) {
      rc = -1;
      break;
    }
    adj[(size_t)a * degree + fill[a]++] = b;
    adj[(size_t)b * degree + fill[b]++] = a;
  }
  for (int v = 0; v < nodes && !rc; v++)
    for (int k = fill[v]; k < degree; k++) adj[(size_t)v * degree + k] = v;
  free(fill);
  return rc;
}

int orc_rotate(int v, int height, int width, int groups, int degree) {
  /* grid/common.c:63-86 (v = column*height + row). */
  int w = v / height, h = v % height;
  if (groups == 2) return (width - w - 1) * height + (height - h - 1);
  if (degree == 90) return h * height + (height - w - 1);
  if (degree == 180) return (height - w - 1) * height + (height - h - 1);
  return (height - h - 1) * height + w;
}

void orc_rotate_hash(int nodes, int height, int width, int groups, int *hash) {
  /* grid/common.c:312-341: base quadrant -> 0..B-1, its images -> next blocks. */
  int based = nodes / groups;
  if (groups == 1) {
    for (int i = 0; i < based; i++) hash[i] = i;
    return;
  }
  int bh = height / 2;
  for (int i = 0; i < based; i++) {
    int j = (i / bh) * height + (i % bh);
    hash[j] = i;
    if (groups == 2) {
      hash[orc_rotate(j, height, width, 2, 180)] = i + based;
    } else {
      hash[orc_rotate(j, height, width, 4, 90)] = i + based;
      hash[orc_rotate(j, height, width, 4, 180)] = i + 2 * based;
      hash[orc_rotate(j, height, width, 4, 270)] = i + 3 * based;
    }
  }
}
