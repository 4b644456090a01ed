/*
 * gg_incr.h -- incremental re-scoring of 2-opt moves (part of libggeval.so).
 *
 * No counterpart in the reference: its SA re-runs the whole evaluation()
 * for every proposal (general/sa.c:465-475).  This scorer keeps one simple
 * graph resident on the device together with every vertex's distance sum
 * and eccentricity, and scores a 2-opt move by re-running the BFS only from
 * the sources whose distances the move can change.  Results are the exact
 * (connected, diameter, S) of the moved graph -- the same integers as
 * gg_apsp() on it, and the reference's ASPL = S / (((double)n-1)*n).
 *
 * Which sources: with d(s,.) the distances of the current graph, a source s
 * keeps all its distances under the move unless
 *   - a removed edge (u,v) with d(s,u) < d(s,v) was v's only link to the
 *     level before it (no other neighbour w of v has d(s,w) = d(s,v) - 1), or
 *     the same with u and v exchanged; or
 *   - an added edge (x,y) has |d(s,x) - d(s,y)| >= 2.
 * (Removal: a vertex that keeps a parent at the level before it keeps its
 * distance, and level by level so does every vertex; the two removed edges
 * share no vertex, so their conditions combine.  Addition: an edge between
 * vertices at most one level apart shortens no path from s.)  The levels of
 * the four endpoints and of their neighbours come from 16-source probe BFS
 * runs: per endpoint ceil((degree-1)/15) of them, each with the endpoint and
 * up to 15 of its other neighbours.
 *
 * Graphs: simple (no self-loop, no repeated edge), degree-regular adjacency
 * int32 [nodes][degree], 2 <= degree <= 32, nodes <= 65536, connected.
 * Not thread-safe; one pending move at a time.
 */
#ifndef GG_INCR_H
#define GG_INCR_H
#include <stdint.h>

#include "gg_eval.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_incr gg_incr;

/* Uploads the graph and computes every vertex's distance sum and
 * eccentricity (one full per-source run).  NULL on error (gg_last_error()). */
gg_incr *gg_incr_create(int nodes, int degree, const int32_t *adj);

/* The current graph's (connected = 1, diameter, S). */
int gg_incr_state(const gg_incr *s, gg_result *out);

/* Scores the current graph with edges rem = {u1,v1, u2,v2} removed and
 * add = {x1,y1, x2,y2} added, where the four endpoints are distinct, both
 * removed edges exist, each endpoint appears once in the added edges and no
 * added edge exists yet.  `out` receives the moved graph's exact result
 * (connected = 0 when it is disconnected); `affected` (may be NULL) the
 * number of sources re-run.  The move stays pending until gg_incr_accept
 * or gg_incr_reject. */
int gg_incr_eval(gg_incr *s, const int32_t rem[4], const int32_t add[4], gg_result *out,
                 int32_t *affected);

/* Makes the pending move part of the current graph (its distance sums and
 * eccentricities are recomputed for the affected sources). */
int gg_incr_accept(gg_incr *s);

/* Drops the pending move. */
int gg_incr_reject(gg_incr *s);

/* Copies the current graph's adjacency (nodes x degree, in row order as
 * moves left it) to adj. */
int gg_incr_adjacency(const gg_incr *s, int32_t *adj);

void gg_incr_destroy(gg_incr *s);

#ifdef __cplusplus
}
#endif
#endif
