/*
 * ref_glue.c -- a few lines that let tests call the reference's own,
 * unmodified evaluator sources (compiled by oracle/Makefile straight from
 * /root/reference into oracle/_ref/).  TEST INFRASTRUCTURE ONLY.
 *
 * The reference's translation units expect an initialised MPI world and the
 * globals `rank` / `procs` (general/common.h:27, grid/common.h:22).  This
 * file provides the setup call; the reference's `evaluation()` and `sa()`
 * are then called directly through ctypes.
 */
#include <mpi.h>
#include <stdlib.h>

extern int rank, procs;

int ref_setup(void) {
  int inited = 0;
  MPI_Initialized(&inited);
  if (!inited) {
    int argc = 0;
    char **argv = NULL;
    if (MPI_Init(&argc, &argv) != MPI_SUCCESS) return -1;
  }
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  MPI_Comm_size(MPI_COMM_WORLD, &procs);
  return procs;
}

/* libc random() seeding, so the reference SA and the product SA can be
 * driven from the same seed in one process. */
void ref_srandom(unsigned seed) { srandom(seed); }
