// gg_grid_shim.cpp -- libgg_grid.so: the reference's grid evaluation() ABI
// (grid/common.h:88-90) over libggeval.so.
#include <cstdio>
#include <cstdlib>

#include "gg_eval.h"
#include "gg_grid.h"

extern "C" __attribute__((weak)) void timer_start(const int) {}
extern "C" __attribute__((weak)) void timer_stop(const int) {}
static const int kTimerApsp = 2;  // TIMER_APSP, grid/common.h:30

static void fatal(const char *what) {
  fprintf(stderr, "gg grid evaluation: %s: %s\n", what, gg_last_error());
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

extern "C" bool evaluation(const int nodes, const int degree, const int groups, const int *adjacency,
                           const int based_nodes, const int height, const int based_height,
                           int *diameter, double *ASPL, const bool enable_bfs,
                           const int *rotate_hash) {
  (void)height;
  (void)based_height;  // the base quadrant is rotate_hash^-1(0..based_nodes-1)
  timer_start(kTimerApsp);
  init_keeping_random();
  gg_graph g = {nodes, degree, based_nodes, groups, 0, 0, adjacency, rotate_hash};
  gg_result r;
  if (gg_apsp(&g, &r) != GG_OK) fatal("evaluation");
  bool ok = r.connected && !(r.diameter > nodes);  // grid/evaluation.c:182
  if (ok) {
    *diameter = r.diameter;
    *ASPL = (double)r.sum / (((double)nodes - 1) * nodes);
  } else {
    *diameter = enable_bfs ? r.max_ecc : nodes + 1;
  }
  timer_stop(kTimerApsp);
  return ok;
}
