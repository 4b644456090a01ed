"""Simulated-annealing drivers over 2-opt moves (host C++ in libggeval.so).

``sa_general`` follows the reference's sa() (general/sa.c:413-520) step for
step, drawing from libc random() in the same order, so that with the same
seed and start graph it visits the same graphs; the evaluator is the device
kernels, or any callable plugged in (used by the tests to run the oracle).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import _lib as L


@dataclass
class SaResult:
    edges: np.ndarray
    steps: int
    accepts: int
    evaluations: int
    best_diameter: int
    best_aspl: float
    seconds_total: float
    seconds_eval: float
    log_diameter: Optional[np.ndarray] = None
    log_sum: Optional[np.ndarray] = None


def sa_general(edges, nodes: int, degree: int, seed: int, ncalcs: int, max_temp: float = 100.0,
               min_temp: float = 0.217147, cooling_cycle: int = 1, hill_climbing: bool = False,
               evaluator: Optional[Callable] = None, log: int = 0) -> SaResult:
    """Plain SA on a general graph (groups = 1).  `evaluator(adj) ->
    (connected, diameter, S)` replaces the device evaluator when given."""
    e = L.c32(edges).reshape(-1, 2).copy()
    p = L.GGSaParams(ncalcs, max_temp, min_temp, cooling_cycle, int(hill_climbing), seed, log)
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = ctypes.cast(None, L.EVAL_FN)
    keep = None
    if evaluator is not None:
        def _cb(ctx, n, d, adj_p, diam_p, sum_p, conn_p):
            try:
                adj = np.ctypeslib.as_array(adj_p, shape=(n, d)).copy()
                ok, dm, S = evaluator(adj)
                diam_p[0] = int(dm)
                sum_p[0] = int(S)
                conn_p[0] = int(bool(ok))
                return 0
            except Exception:  # surfaced as an error code
                return -1
        keep = L.EVAL_FN(_cb)
        cb = keep
    rc = L.lib().gg_sa_general(nodes, degree, L.ptr(e), ctypes.byref(p), cb, None,
                               ctypes.byref(st), L.ptr(ld),
                               ls.ctypes.data_as(L.u64p))
    L.check(rc)
    n_log = min(log, st.evaluations)
    return SaResult(e, st.steps, st.accepts, st.evaluations, st.best_diameter, st.best_aspl,
                    st.seconds_total, st.seconds_eval,
                    ld[:n_log] if log else None, ls[:n_log] if log else None)
