/*
 * gg_general.h -- the reference's general-graph evaluator ABI, served by
 * libgg_general.so (a thin shim over libggeval.so).
 *
 * Link libgg_general.so in place of the reference's evaluation.o /
 * evaluation_cuda.o and these symbols resolve exactly as before.
 */
#ifndef GG_GENERAL_H
#define GG_GENERAL_H
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces general/evaluation.c:375-391 (CPU) and general/evaluation.cu:205-218
 * (CUDA); declared at general/common.h:112-113.  algo (general/common.h:58-60):
 * 1 = MATRIX_OP and 2 = MATRIX_OP_MEM_SAVING give matrix_op's result
 * (:153-373; any n, tiles are always used); 0 = BFS gives bfs()'s
 * (:55-150): distances along adj from each source to the targets i > s,
 * symmetry weighted, over n(n-1)/2 -- for directed or non-rotation-invariant
 * tables a different number than matrix_op's, as in the reference.  Returns
 * true when connected, with *diameter and *ASPL set; false with
 * *diameter = INT_MAX, *ASPL = DBL_MAX otherwise (general/evaluation.c:137-147,
 * :254-262).  Brackets TIMER_APSP with the caller's timer_start/timer_stop
 * (general/timer.c:23-38) when they are linked in (for algo = BFS on a
 * disconnected graph the timer is stopped twice, as bfs() does at :146). */
bool evaluation(const int nodes, const int based_nodes, const int groups, const int lines,
                const int degree, int *adj, int *diameter, double *ASPL,
                const int added_centers, const int algo);

/* Replaces general/evaluation.cu:178-194 / :196-203 (called from
 * general/main.c:420 and :493).  Optional: evaluation() initialises lazily. */
void init_matrix_dev(const int nodes, const int degree, const int based_nodes, const int algo);
void finalize_matrix_dev(void);

#ifdef __cplusplus
}
#endif
#endif
