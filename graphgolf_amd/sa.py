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
    already_optimal: bool = False
    max_diff_energy: float = 0.0  # with detect_temp: the largest energy increase tested


def _general_callback(evaluator):
    """ctypes evaluator callback (None: the device evaluator)."""
    if evaluator is None:
        return ctypes.cast(None, L.EVAL_FN)

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
    return L.EVAL_FN(_cb)


def _grid_callback(evaluator):
    if evaluator is None:
        return ctypes.cast(None, L.EVAL_GRID_FN)

    def _cb(ctx, n, d, g, b, adj_p, rh_p, diam_p, sum_p, conn_p):
        try:
            adj = np.ctypeslib.as_array(adj_p, shape=(n, d)).copy()
            rhash = np.ctypeslib.as_array(rh_p, shape=(n,)).copy()
            ok, dm, S = evaluator(adj, g, b, rhash)
            diam_p[0] = int(dm)
            sum_p[0] = int(S)
            conn_p[0] = int(bool(ok))
            return 0
        except Exception:  # surfaced as an error code
            return -1
    return L.EVAL_GRID_FN(_cb)


def _sym_callback(evaluator):
    if evaluator is None:
        return ctypes.cast(None, L.EVAL_SYM_FN)

    def _cb(ctx, n, d, b, g, c, adj_p, diam_p, sum_p, conn_p):
        try:
            adj = np.ctypeslib.as_array(adj_p, shape=(n, d)).copy()
            ok, dm, S = evaluator(adj, b, g, c)
            diam_p[0] = int(dm)
            sum_p[0] = int(S)
            conn_p[0] = int(bool(ok))
            return 0
        except Exception:  # surfaced as an error code
            return -1
    return L.EVAL_SYM_FN(_cb)


def _result(rc, e, st, log, ld, ls) -> SaResult:
    if rc != L.GG_ALREADY_OPTIMAL:
        L.check(rc)
    n_log = min(log, st.evaluations)
    return SaResult(e, st.steps, st.accepts, st.evaluations, st.best_diameter, st.best_aspl,
                    st.seconds_total, st.seconds_eval,
                    ld[:n_log] if log else None, ls[:n_log] if log else None,
                    already_optimal=rc == L.GG_ALREADY_OPTIMAL,
                    max_diff_energy=float(st.max_diff_energy))


def sa_general(edges, nodes: int, degree: int, seed: int, ncalcs: int, max_temp: float = 100.0,
               min_temp: float = 0.217147, cooling_cycle: int = 1, hill_climbing: bool = False,
               evaluator: Optional[Callable] = None, log: int = 0,
               incremental: bool = False) -> SaResult:
    """Plain SA on a general graph (groups = 1).  `evaluator(adj) ->
    (connected, diameter, S)` replaces the device evaluator when given;
    `incremental` scores proposals with the incremental scorer (same exact
    results, so the same trajectory)."""
    e = L.c32(edges).reshape(-1, 2).copy()
    p = L.GGSaParams(ncalcs, max_temp, min_temp, cooling_cycle, int(hill_climbing), seed, log,
                     int(incremental))
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _general_callback(evaluator)
    rc = L.lib().gg_sa_general(nodes, degree, L.ptr(e), ctypes.byref(p), cb, None,
                               ctypes.byref(st), L.ptr(ld), ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


def general_program(edges, nodes: int, degree: int, seed: int, ncalcs: int, halfway: bool = False,
                    max_temp: float = 100.0, min_temp: float = 0.217147, hill_climbing: bool = False,
                    evaluator: Optional[Callable] = None, log: int = 0,
                    incremental: bool = False) -> SaResult:
    """The general program's run from its start graph (general/main.c:360-462):
    the pre-run moves unless `halfway` (-M), the start-graph check, then SA, all
    on one random() stream seeded with `seed`.  `already_optimal` is set when
    the start graph meets the Moore bound (nothing run)."""
    e = L.c32(edges).reshape(-1, 2).copy()
    p = L.GGSaParams(ncalcs, max_temp, min_temp, 1, int(hill_climbing), seed, log,
                     int(incremental))
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _general_callback(evaluator)
    rc = L.lib().gg_general_program(nodes, degree, L.ptr(e), int(halfway), ctypes.byref(p), cb,
                                    None, ctypes.byref(st), L.ptr(ld), ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


def sa_grid(edges, width: int, height: int, degree: int, length: int, seed: int, ncalcs: int,
            groups: int = 1, max_temp: float = 100.0, min_temp: float = 0.217147,
            cooling_cycle: int = 1, hill_climbing: bool = False, fixed_temp: float = 0.0,
            evaluator: Optional[Callable] = None, log: int = 0) -> SaResult:
    """Plain SA on a grid graph (grid/sa.c:283-361).  `evaluator(adj, groups,
    based_nodes, rotate_hash) -> (connected, diameter, S)` replaces the device
    evaluator when given."""
    from .graphs import rotate_hash as make_rotate_hash

    nodes = width * height
    e = L.c32(edges).reshape(-1, 2).copy()
    rh = make_rotate_hash(nodes, height, width, groups)
    shape = L.GGGridShape(width, height, groups, length)
    p = L.GGSaParams(ncalcs, max_temp, min_temp, cooling_cycle, int(hill_climbing), seed, log)
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _grid_callback(evaluator)
    rc = L.lib().gg_sa_grid(ctypes.byref(shape), degree, L.ptr(e), L.ptr(rh), ctypes.byref(p),
                            fixed_temp, cb, None, ctypes.byref(st), L.ptr(ld),
                            ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


GENERAL_FULL_FILE, GENERAL_BASE_FILE, GENERAL_BASE_CENTRES = 0, 1, 2


def sa_general_sym(edges, nodes: int, degree: int, groups: int, added_centers: int, seed: int,
                   ncalcs: int, max_temp: float = 100.0, min_temp: float = 0.217147,
                   cooling_cycle: int = 1, hill_climbing: bool = False,
                   evaluator: Optional[Callable] = None, log: int = 0) -> SaResult:
    """The reference's sa() for rotationally symmetric graphs with centres
    (general/sa.c:413-520).  `evaluator(adj, based_nodes, groups,
    added_centers) -> (connected, diameter, S)` replaces the device evaluator."""
    e = L.c32(edges).reshape(-1, 2).copy()
    shape = L.GGGeneralShape(nodes, degree, groups, added_centers)
    p = L.GGSaParams(ncalcs, max_temp, min_temp, cooling_cycle, int(hill_climbing), seed, log, 0)
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _sym_callback(evaluator)
    rc = L.lib().gg_sa_general_sym(ctypes.byref(shape), L.ptr(e), ctypes.byref(p), cb, None,
                                   ctypes.byref(st), L.ptr(ld), ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


def general_program_sym(file_edges, nodes: int, degree: int, groups: int, added_centers: int,
                        source: int, seed: int, ncalcs: int, max_temp: float = 100.0,
                        min_temp: float = 0.217147, hill_climbing: bool = False,
                        evaluator: Optional[Callable] = None, log: int = 0,
                        not_simple: bool = False, detect_temp: bool = False) -> SaResult:
    """The general program's run for any groups and centres (general/main.c:
    360-462); file_edges are the lines read from the file."""
    f = L.c32(file_edges).reshape(-1, 2)
    lines = nodes * degree // 2
    e = np.zeros((lines, 2), dtype=np.int32)
    e[:len(f)] = f
    shape = L.GGGeneralShape(nodes, degree, groups, added_centers, int(not_simple))
    p = L.GGSaParams(ncalcs, max_temp, min_temp, 1, int(hill_climbing), seed, log, 0,
                     int(detect_temp))
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _sym_callback(evaluator)
    rc = L.lib().gg_general_program_sym(ctypes.byref(shape), L.ptr(e), source, len(f),
                                        ctypes.byref(p), cb, None, ctypes.byref(st), L.ptr(ld),
                                        ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


GRID_LATTICE, GRID_BASE_FILE, GRID_FULL_FILE = 0, 1, 2


def grid_program(edges, width: int, height: int, degree: int, length: int, seed: int, ncalcs: int,
                 source: int, groups: int = 1, max_temp: float = 100.0, min_temp: float = 0.217147,
                 cooling_cycle: int = 1, hill_climbing: bool = False, fixed_temp: float = 0.0,
                 evaluator: Optional[Callable] = None, log: int = 0,
                 detect_temp: bool = False) -> SaResult:
    """The grid program's run (grid/main.c:630-727): the lattice (source
    GRID_LATTICE, `edges` ignored), the symmetric copies (GRID_BASE_FILE with
    groups > 1: `edges` holds the base grid's lines in base coordinates), the
    checks, then SA, all on one random() stream seeded with `seed`."""
    from .graphs import rotate_hash as make_rotate_hash

    nodes = width * height
    lines = nodes * degree // 2
    e = np.zeros((lines, 2), dtype=np.int32)
    if source != GRID_LATTICE:
        src = L.c32(edges).reshape(-1, 2)
        e[:len(src)] = src
    rh = make_rotate_hash(nodes, height, width, groups)
    shape = L.GGGridShape(width, height, groups, length)
    p = L.GGSaParams(ncalcs, max_temp, min_temp, cooling_cycle, int(hill_climbing), seed, log, 0,
                     int(detect_temp))
    st = L.GGSaStats()
    ld = np.zeros(max(log, 1), dtype=np.int32)
    ls = np.zeros(max(log, 1), dtype=np.uint64)
    cb = _grid_callback(evaluator)
    rc = L.lib().gg_grid_program(ctypes.byref(shape), degree, L.ptr(e), source, L.ptr(rh),
                                 ctypes.byref(p), fixed_temp, cb, None, ctypes.byref(st),
                                 L.ptr(ld), ls.ctypes.data_as(L.u64p))
    return _result(rc, e, st, log, ld, ls)


class BsaParams(ctypes.Structure):
    _fields_ = [("max_steps", ctypes.c_int64), ("time_budget_s", ctypes.c_double),
                ("target_aspl", ctypes.c_double), ("proposals", ctypes.c_int32),
                ("exact", ctypes.c_int32), ("combine", ctypes.c_int32),
                ("warmup", ctypes.c_int32), ("explore_every", ctypes.c_int32),
                ("use_surrogate", ctypes.c_int32), ("random_ranking", ctypes.c_int32),
                ("max_temp", ctypes.c_double),
                ("min_temp", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("train_steps", ctypes.c_int32), ("batch_train", ctypes.c_int32),
                ("lr", ctypes.c_float), ("hist_maxd", ctypes.c_int32),
                ("feature_refresh", ctypes.c_int32), ("log_capacity", ctypes.c_int32),
                ("grid_height", ctypes.c_int32), ("grid_length", ctypes.c_int32),
                ("incremental", ctypes.c_int32), ("groups", ctypes.c_int32)]


class BsaStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("accepts", ctypes.c_int64),
                ("evaluations", ctypes.c_int64), ("best_diameter", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("best_aspl", ctypes.c_double),
                ("seconds_total", ctypes.c_double), ("seconds_exact", ctypes.c_double),
                ("seconds_surrogate", ctypes.c_double), ("seconds_features", ctypes.c_double),
                ("seconds_to_target", ctypes.c_double), ("surrogate_rank_corr", ctypes.c_double),
                ("moves_accepted", ctypes.c_int64), ("seconds_train", ctypes.c_double)]


@dataclass
class BatchedResult:
    edges: np.ndarray
    stats: dict
    log_time: np.ndarray
    log_best_aspl: np.ndarray
    log_best_diam: np.ndarray
    accepted: np.ndarray        # [k, 3] (line0, line1, variant)
    accepted_diam: np.ndarray
    accepted_sum: np.ndarray


def sa_batched(edges, nodes: int, degree: int, *, max_steps: int = 100, time_budget_s: float = 0.0,
               target_aspl: float = 0.0, proposals: int = 4096, exact: int = 1, combine: int = 4,
               warmup: int = 256, explore_every: int = 8, use_surrogate: bool = True,
               random_ranking: bool = False,
               max_temp: float = 100.0, min_temp: float = 0.217147, seed: int = 1,
               train_steps: int = 1, batch_train: int = 256, lr: float = 1e-3, hist_maxd: int = 8,
               feature_refresh: int = 100, log_capacity: int = 100000,
               accepted_capacity: int = 100000, grid_height: int = 0,
               grid_length: int = 0, incremental: bool = False, groups: int = 1) -> BatchedResult:
    """Batched (and, with use_surrogate, MLP-ranked) SA on the GPU; see
    include/gg_host.h.  exact = 1 without the surrogate is plain single-graph
    GPU SA with a device-resident adjacency.  grid_height > 0 keeps every
    edge of a grid graph within grid_length Manhattan steps.  groups > 1: a
    rotationally symmetric graph (general: v -> v + n/groups; grid: 2 or 4
    groups), moved by 2g-edge 2-opts that keep the symmetry."""
    lib = L.lib()
    if not getattr(lib, "_bsa_sigs", False):
        lib.gg_sa_batched.argtypes = [ctypes.c_int, ctypes.c_int, L.i32p,
                                      ctypes.POINTER(BsaParams), ctypes.POINTER(BsaStats),
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), L.i32p, L.i32p, L.i32p,
                                      L.u64p, ctypes.c_int]
        lib.gg_sa_batched.restype = ctypes.c_int
        lib.gg_apply_move.argtypes = [ctypes.c_int, L.i32p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
        lib.gg_apply_move.restype = ctypes.c_int
        lib._bsa_sigs = True
    e = L.c32(edges).reshape(-1, 2).copy()
    p = BsaParams(max_steps, time_budget_s, target_aspl, proposals, exact, combine, warmup,
                  explore_every, int(use_surrogate), int(random_ranking), max_temp, min_temp,
                  seed, train_steps, batch_train, lr, hist_maxd, feature_refresh, log_capacity,
                  grid_height, grid_length, int(incremental), int(groups))
    st = BsaStats()
    lt = np.zeros(max(log_capacity, 1))
    la = np.zeros(max(log_capacity, 1))
    ld = np.zeros(max(log_capacity, 1), dtype=np.int32)
    acc = np.zeros((max(accepted_capacity, 1), 3), dtype=np.int32)
    ad = np.zeros(max(accepted_capacity, 1), dtype=np.int32)
    asum = np.zeros(max(accepted_capacity, 1), dtype=np.uint64)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.gg_sa_batched(nodes, degree, L.ptr(e), ctypes.byref(p), ctypes.byref(st),
                           lt.ctypes.data_as(dp), la.ctypes.data_as(dp), L.ptr(ld), L.ptr(acc),
                           L.ptr(ad), asum.ctypes.data_as(L.u64p), accepted_capacity)
    L.check(rc)
    stats = {k: getattr(st, k) for k, _ in BsaStats._fields_ if k != "pad"}
    ns = min(int(st.steps), log_capacity)
    na = min(int(st.moves_accepted), accepted_capacity)
    return BatchedResult(e, stats, lt[:ns], la[:ns], ld[:ns], acc[:na], ad[:na], asum[:na])


def apply_move(edges: np.ndarray, line0: int, line1: int, variant: int, *, nodes: int = 0,
               degree: int = 0, groups: int = 1, grid_height: int = 0) -> None:
    """Applies a logged batched-SA move to `edges` in place (groups > 1: the
    move and its rotated copies; nodes and degree then required)."""
    if groups > 1:
        lib = L.lib()
        lib.gg_apply_move_sym.argtypes = [ctypes.c_int] * 4 + [L.i32p] + [ctypes.c_int] * 3
        lib.gg_apply_move_sym.restype = ctypes.c_int
        L.check(lib.gg_apply_move_sym(nodes, degree, groups, grid_height, L.ptr(edges), int(line0),
                                      int(line1), int(variant)))
        return
    L.check(L.lib().gg_apply_move(edges.shape[0], L.ptr(edges), int(line0), int(line1),
                                  int(variant)))


def sa_features(adj: np.ndarray, props: np.ndarray, vstat: np.ndarray | None = None) -> np.ndarray:
    """The batched SA's device-computed move features ([count, 64] float32) for
    moves props[i] = (a, b, c, d, r) of the graph `adj` ([n, d] int32)."""
    lib = L.lib()
    lib.gg_sa_features.argtypes = [ctypes.c_int, ctypes.c_int, L.i32p, ctypes.c_int, L.i32p,
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.gg_sa_features.restype = ctypes.c_int
    adj = L.c32(adj)
    props = L.c32(props).reshape(-1, 5)
    n, d = adj.shape
    out = np.zeros((props.shape[0], 64), dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    vs = None
    if vstat is not None:
        vstat = np.ascontiguousarray(vstat, dtype=np.float32).reshape(n, 10)
        vs = vstat.ctypes.data_as(fp)
    L.check(lib.gg_sa_features(n, d, L.ptr(adj), props.shape[0], L.ptr(props), vs,
                               out.ctypes.data_as(fp)))
    return out


def sa_proposals(edges, nodes: int, degree: int, count: int, seed: int, step: int,
                 grid_height: int = 0, grid_length: int = 0):
    """The device move draw of ranked batched-SA steps: (props [count, 5] =
    (a, b, c, d, r), lines [count, 3] = (t0, t1, valid))."""
    lib = L.lib()
    lib.gg_sa_proposals.argtypes = [ctypes.c_int, ctypes.c_int, L.i32p, L.i32p, ctypes.c_int,
                                    ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                    L.i32p, L.i32p]
    lib.gg_sa_proposals.restype = ctypes.c_int
    from .graphs import edges_to_adj
    e = L.c32(edges).reshape(-1, 2)
    adj = edges_to_adj(nodes, degree, e)
    props = np.zeros((count, 5), dtype=np.int32)
    lines = np.zeros((count, 3), dtype=np.int32)
    L.check(lib.gg_sa_proposals(nodes, degree, L.ptr(e), L.ptr(adj), count, seed, step,
                                grid_height, grid_length, L.ptr(props), L.ptr(lines)))
    return props, lines
