// gg_nonregular_shim.cpp -- the non-regular variants' evaluation() ABI
// (include/gg_nonregular.h) over libggeval.so.  Built twice: with
// GG_NONREGULAR_GRID for libgg_nonregular_grid.so, without it for
// libgg_nonregular_general.so.
#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gg_eval.h"
#include "gg_nonregular.h"

extern "C" __attribute__((weak)) void timer_start(const int) {}
extern "C" __attribute__((weak)) void timer_stop(const int) {}
static const int kTimerApsp = 2;  // TIMER_APSP in both variants' common.h

static void fatal(const char *what) {
  fprintf(stderr, "gg non-regular evaluation: %s: %s\n", what, gg_last_error());
  fflush(stderr);
  exit(1);
}

// The device is brought up with libc random() switched to a private state:
// the reference's callers seed random() before their first evaluation()
// (general/main.c:360, grid/main.c:630) and replay its stream, and the HIP
// runtime's start-up must not draw from it.
static void init_keeping_random() {
  static bool done = false;
  if (done) return;
  static char own_state[256];
  char *caller = initstate(1u, own_state, sizeof own_state);
  const int rc = gg_init(-1);
  setstate(caller);
  if (rc != GG_OK) fatal("device initialisation");
  done = true;
}

// Row v: its first len[v] entries, then v itself in the unused slots.
static const int32_t *padded(int nodes, int width, const int *adj, const int *len) {
  static thread_local std::vector<int32_t> buf;
  buf.resize((size_t)nodes * width);
  for (int v = 0; v < nodes; ++v) {
    const int k = len[v];
    if (k < 0 || k > width) {
      fprintf(stderr, "gg non-regular evaluation: vertex %d has degree %d outside [0,%d]\n", v, k,
              width);
      exit(1);
    }
    const int *row = adj + (size_t)v * width;
    int32_t *out = buf.data() + (size_t)v * width;
    for (int j = 0; j < k; ++j) out[j] = row[j];
    for (int j = k; j < width; ++j) out[j] = v;
  }
  return buf.data();
}

#ifdef GG_NONREGULAR_GRID
extern "C" bool evaluation(const int nodes, const int max_degree, const int *degree, const int groups,
                           const int *adjacency, const int based_nodes, const int height,
                           const int based_height, int *diameter, double *ASPL,
                           const bool enable_bfs, const int *rotate_hash) {
  (void)height;
  (void)based_height;
  timer_start(kTimerApsp);
  init_keeping_random();
  gg_graph g = {nodes, max_degree, based_nodes, groups, 0, 0,
                padded(nodes, max_degree, adjacency, degree), rotate_hash};
  gg_result r;
  if (gg_apsp(&g, &r) != GG_OK) fatal("evaluation");
  const bool ok = r.connected && !(r.diameter > nodes);  // grid/evaluation.c:182 in the variant
  if (ok) {
    *diameter = r.diameter;
    *ASPL = (double)r.sum / (((double)nodes - 1) * nodes);
  } else {
    *diameter = enable_bfs ? r.max_ecc : nodes + 1;
  }
  timer_stop(kTimerApsp);
  return ok;
}
#else
extern "C" bool evaluation(const int nodes, int based_nodes, const int groups, const int lines,
                           const int degree, int *adj, int *diameter, double *ASPL,
                           const int added_centers, const int *num_degrees, const int algo) {
  (void)lines;
  timer_start(kTimerApsp);
  init_keeping_random();
  const bool bfs = algo == 0;  // BFS: the variant's bfs() (evaluation.c:57-150)
  gg_graph g = {nodes, degree, based_nodes, groups, added_centers, bfs ? GG_GRAPH_BFS : 0,
                padded(nodes, degree, adj, num_degrees), nullptr};
  gg_result r;
  if (gg_apsp(&g, &r) != GG_OK) fatal("evaluation");
  const bool ok = bfs ? r.connected != 0 : (r.connected && r.diameter < nodes);
  if (ok) {
    *diameter = r.diameter;
    *ASPL = bfs ? (double)r.sum / ((((double)nodes - 1) * nodes) / 2)
                : (double)r.sum / (((double)nodes - 1) * nodes);
  } else {
    *diameter = INT_MAX;
    *ASPL = DBL_MAX;
    if (bfs) timer_stop(kTimerApsp);  // the variant's bfs() stops the timer here too
  }
  timer_stop(kTimerApsp);
  return ok;
}
#endif
