// gg_general_shim.cpp -- libgg_general.so: the reference's general
// evaluation() ABI (general/common.h:112-113) over libggeval.so.
#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstdlib>

#include "gg_eval.h"
#include "gg_general.h"

// The reference links its own timer.c (general/timer.c:23-38); when it is
// present those strong definitions win over these no-op defaults.
extern "C" __attribute__((weak)) void timer_start(const int) {}
extern "C" __attribute__((weak)) void timer_stop(const int) {}
static const int kTimerApsp = 2;  // TIMER_APSP, general/common.h:41
static const int kAlgoBfs = 0;    // BFS, general/common.h:58

// Fatal errors end the process, as the reference's ERROR() does
// (general/common.h:67-70).
static void fatal(const char *what) {
  fprintf(stderr, "gg evaluation: %s: %s\n", what, gg_last_error());
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

extern "C" void init_matrix_dev(const int, const int, const int, const int) {
  init_keeping_random();
}

extern "C" void finalize_matrix_dev(void) { gg_finalize(); }

extern "C" bool evaluation(const int nodes, const int based_nodes, const int groups, const int lines,
                           const int degree, int *adj, int *diameter, double *ASPL,
                           const int added_centers, const int algo) {
  (void)lines;
  timer_start(kTimerApsp);
  init_keeping_random();
  const bool bfs = algo == kAlgoBfs;
  gg_graph g = {nodes, degree, based_nodes, groups, added_centers, bfs ? GG_GRAPH_BFS : 0, adj,
                nullptr};
  gg_result r;
  if (gg_apsp(&g, &r) != GG_OK) fatal("evaluation");
  bool ok;
  if (bfs) {
    // general/evaluation.c:137-147: "reached" covers the targets i > s.
    ok = r.connected != 0;
    if (ok) {
      *diameter = r.diameter;
      *ASPL = (double)r.sum / ((((double)nodes - 1) * nodes) / 2);
    } else {
      *diameter = INT_MAX;
      *ASPL = DBL_MAX;
      timer_stop(kTimerApsp);  // bfs() stops the timer here too (:146), then evaluation() again
    }
  } else {
    ok = r.connected && r.diameter < nodes;  // general/evaluation.c:254
    if (ok) {
      *diameter = r.diameter;
      *ASPL = (double)r.sum / (((double)nodes - 1) * nodes);
    } else {
      *diameter = INT_MAX;
      *ASPL = DBL_MAX;
    }
  }
  timer_stop(kTimerApsp);
  return ok;
}
