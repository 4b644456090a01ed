"""Times the reference's own general ``evaluation()`` on the host cores.

TEST / BASELINE INFRASTRUCTURE ONLY: the CPU-baseline leg of the benchmark
runs this module in a child process, so that the reference build's MPI
singleton start-up (``ref_glue.c``) stays out of the process that drives the
GPU and a stall there is ended by the parent's time limit.

    python -m oracle.ref_timing ADJ.npy BUDGET_SECONDS

``ADJ.npy`` is an int32 ``[n][d]`` adjacency table.  The library is
``oracle/_ref/libref_general.so``, compiled by ``oracle/Makefile`` from the
reference's unmodified ``general/evaluation.c`` (``evaluation()`` at
general/evaluation.c:375-391, dispatching to ``matrix_op()`` at :153-263,
OpenMP over the source tiles).  The algorithm argument follows the
reference program's own choice (general/main.c:411-417): MATRIX_OP when
the n x ceil(n/64) uint64 matrix fits 2 GiB, MATRIX_OP_MEM_SAVING otherwise.

Prints one JSON line: the per-evaluation times at the OpenMP runtime's
default thread count, the thread count, one evaluation on one thread, and
the (diameter, S) of the last run.
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

from . import ref_available, ref_evaluation, set_threads, threads

MATRIX_OP, MATRIX_OP_MEM_SAVING = 1, 2
MATRIX_OP_THRESHOLD = 2147483648  # bytes, general/common.h:61


def algo_for(n: int, based_nodes: int, groups: int) -> int:
    """The reference program's algorithm choice (general/main.c:411-417)."""
    elements = (based_nodes + 63) // 64
    s = float(based_nodes) * groups * elements * 8
    return MATRIX_OP if s <= MATRIX_OP_THRESHOLD else MATRIX_OP_MEM_SAVING


def main(argv) -> int:
    path, budget = argv[0], float(argv[1])
    if not ref_available("general"):
        print(json.dumps({"error": "oracle/_ref/libref_general.so is not built"}))
        return 1
    adj = np.load(path, allow_pickle=False).astype(np.int32)
    n = adj.shape[0]
    algo = algo_for(n, n, 1)
    times = []
    t0 = time.perf_counter()
    r = None
    while True:
        t = time.perf_counter()
        r = ref_evaluation(adj, algo=algo)
        times.append(time.perf_counter() - t)
        if time.perf_counter() - t0 > budget or len(times) >= 50:
            break
    cores = int(threads())
    set_threads(1)
    t = time.perf_counter()
    ref_evaluation(adj, algo=algo)
    one = time.perf_counter() - t
    set_threads(cores)
    print(json.dumps({"times": times, "threads": cores, "one_thread_s": one, "algo": algo,
                      "ok": bool(r.ok), "diameter": int(r.raw_diameter), "S": int(r.S)}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
