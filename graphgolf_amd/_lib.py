"""ctypes bindings of libggeval.so (include/gg_eval.h, include/gg_host.h).

The libraries are built in-tree by ``graphgolf_amd.csrc.build``.  Loading
fails loudly when they are missing: there is no CPU fallback in the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libggeval.so")
GENERAL_PATH = os.path.join(PKG, "libgg_general.so")
GRID_PATH = os.path.join(PKG, "libgg_grid.so")
NONREGULAR_GENERAL_PATH = os.path.join(PKG, "libgg_nonregular_general.so")
NONREGULAR_GRID_PATH = os.path.join(PKG, "libgg_nonregular_grid.so")

GG_OK = 0
GG_ALREADY_OPTIMAL = 1  # *_program: the start graph already meets the lower bound
GG_ERR_INVALID = -1
GG_ERR_TOO_LARGE = -2
GG_ERR_HIP = -3
GG_ERR_STATE = -4
GG_GRAPH_SYMMETRIC = 1
GG_GRAPH_BFS = 2  # the reference's bfs() semantics (include/gg_eval.h)

i32p = ctypes.POINTER(ctypes.c_int32)
u64p = ctypes.POINTER(ctypes.c_uint64)


class GGGraph(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int32), ("degree", ctypes.c_int32),
                ("based_nodes", ctypes.c_int32), ("groups", ctypes.c_int32),
                ("added_centers", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("adj", i32p), ("relabel", i32p)]


class GGResult(ctypes.Structure):
    _fields_ = [("connected", ctypes.c_int32), ("diameter", ctypes.c_int32),
                ("max_ecc", ctypes.c_int32), ("word_bits", ctypes.c_int32),
                ("sum", ctypes.c_uint64)]


class GGGeneralShape(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int32), ("degree", ctypes.c_int32), ("groups", ctypes.c_int32),
                ("added_centers", ctypes.c_int32), ("not_simple", ctypes.c_int32)]


class GGSaParams(ctypes.Structure):
    _fields_ = [("ncalcs", ctypes.c_int64), ("max_temp", ctypes.c_double),
                ("min_temp", ctypes.c_double), ("cooling_cycle", ctypes.c_int32),
                ("hill_climbing", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("log_capacity", ctypes.c_int32), ("incremental", ctypes.c_int32),
                ("detect_temp", ctypes.c_int32)]


class GGSaStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("accepts", ctypes.c_int64),
                ("evaluations", ctypes.c_int64), ("best_diameter", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("best_aspl", ctypes.c_double),
                ("seconds_total", ctypes.c_double), ("seconds_eval", ctypes.c_double),
                ("max_diff_energy", ctypes.c_double)]


EVAL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, i32p,
                           i32p, u64p, i32p)
EVAL_GRID_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, u64p, i32p)
# (ctx, nodes, degree, based_nodes, groups, added_centers, adj, diameter, sum, connected)
EVAL_SYM_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p, u64p, i32p)


class GGGridShape(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("groups", ctypes.c_int32), ("length", ctypes.c_int32)]


class GGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gg error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded libggeval.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m graphgolf_amd.csrc.build` "
                           "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "gg_init": ([ctypes.c_int], ctypes.c_int),
        "gg_finalize": ([], None),
        "gg_last_error": ([], ctypes.c_char_p),
        "gg_device_count": ([], ctypes.c_int),
        "gg_apsp": ([ctypes.POINTER(GGGraph), ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_apsp_device": ([ctypes.POINTER(GGGraph), ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_eval_batch": ([ctypes.c_int, ctypes.POINTER(GGGraph), ctypes.POINTER(GGResult)],
                          ctypes.c_int),
        "gg_plan_create": ([ctypes.c_int, ctypes.POINTER(GGGraph)], ctypes.c_void_p),
        "gg_plan_set_adj": ([ctypes.c_void_p, ctypes.c_int, i32p], ctypes.c_int),
        "gg_plan_adj_device": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_void_p),
        "gg_plan_run": ([ctypes.c_void_p, ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_plan_launch": ([ctypes.c_void_p], ctypes.c_int),
        "gg_plan_wait": ([ctypes.c_void_p, ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_plan_destroy": ([ctypes.c_void_p], None),
        "gg_set_word_bits": ([ctypes.c_int], ctypes.c_int),
        "gg_plan_kernel_ms": ([ctypes.c_void_p], ctypes.c_float),
        "gg_plan_launch_timed": ([ctypes.c_void_p], ctypes.c_int),
        "gg_trace_levels": ([ctypes.POINTER(GGGraph), ctypes.c_int, u64p,
                             ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_synchronize": ([], ctypes.c_int),
        "gg_set_row_order": ([ctypes.c_int], ctypes.c_int),
        "gg_eval_features": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p, i32p,
                              ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_row_order": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p, u64p], ctypes.c_int),
        "gg_bank_layout": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int64,
                            ctypes.c_uint32, ctypes.c_int, i32p, i32p, u64p], ctypes.c_int),
        "gg_autotune": ([ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "gg_autotune_batch": ([ctypes.c_int, ctypes.POINTER(GGGraph), ctypes.c_int, i32p,
                               ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "gg_tune_lookup": ([ctypes.c_int] * 6 + [i32p], ctypes.c_int),
        "gg_plan_config": ([ctypes.c_void_p, i32p], ctypes.c_int),
        "gg_last_error_code": ([], ctypes.c_int),
        "gg_apsp_device_noalloc": ([ctypes.POINTER(GGGraph), ctypes.POINTER(GGResult)],
                                   ctypes.c_int),
        "gg_tune_table_load": ([ctypes.c_char_p], ctypes.c_int),
        "gg_tune_table_save": ([ctypes.c_char_p], ctypes.c_int),
        "gg_tune_table_clear": ([], ctypes.c_int),
        "gg_sa_general": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.POINTER(GGSaParams),
                           EVAL_FN, ctypes.c_void_p, ctypes.POINTER(GGSaStats), i32p, u64p],
                          ctypes.c_int),
        "gg_distance_histogram": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, i32p],
                                  ctypes.c_int),
        "gg_random_regular": ([ctypes.c_int, ctypes.c_int, ctypes.c_uint64, i32p], ctypes.c_int),
        "gg_random_symmetric": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, i32p],
                                ctypes.c_int),
        "gg_random_grid": ([ctypes.c_int] * 4 + [ctypes.c_uint64, i32p], ctypes.c_int),
        "gg_random_grid_symmetric": ([ctypes.c_int] * 5 + [ctypes.c_uint64, i32p], ctypes.c_int),
        "gg_grid_symmetric_from_base": ([ctypes.c_int] * 5 + [ctypes.c_uint32, i32p, i32p],
                                        ctypes.c_int),
        "gg_grid_lower_bound": ([ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int),
                                                      ctypes.POINTER(ctypes.c_double)],
                                ctypes.c_int),
        "gg_sa_grid": ([ctypes.POINTER(GGGridShape), ctypes.c_int, i32p, i32p,
                        ctypes.POINTER(GGSaParams), ctypes.c_double, EVAL_GRID_FN,
                        ctypes.c_void_p, ctypes.POINTER(GGSaStats), i32p, u64p], ctypes.c_int),
        "gg_edges_to_adj": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p, i32p], ctypes.c_int),
        "gg_general_program": ([ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int,
                                ctypes.POINTER(GGSaParams), EVAL_FN, ctypes.c_void_p,
                                ctypes.POINTER(GGSaStats), i32p, u64p], ctypes.c_int),
        "gg_grid_lattice": ([ctypes.c_int] * 4 + [ctypes.c_uint32, i32p], ctypes.c_int),
        "gg_sa_general_sym": ([ctypes.POINTER(GGGeneralShape), i32p, ctypes.POINTER(GGSaParams),
                               EVAL_SYM_FN, ctypes.c_void_p, ctypes.POINTER(GGSaStats), i32p,
                               u64p], ctypes.c_int),
        "gg_general_program_sym": ([ctypes.POINTER(GGGeneralShape), i32p, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(GGSaParams), EVAL_SYM_FN, ctypes.c_void_p,
                                    ctypes.POINTER(GGSaStats), i32p, u64p], ctypes.c_int),
        "gg_incr_create": ([ctypes.c_int, ctypes.c_int, i32p], ctypes.c_void_p),
        "gg_incr_state": ([ctypes.c_void_p, ctypes.POINTER(GGResult)], ctypes.c_int),
        "gg_incr_eval": ([ctypes.c_void_p, i32p, i32p, ctypes.POINTER(GGResult), i32p],
                         ctypes.c_int),
        "gg_incr_accept": ([ctypes.c_void_p], ctypes.c_int),
        "gg_incr_reject": ([ctypes.c_void_p], ctypes.c_int),
        "gg_incr_adjacency": ([ctypes.c_void_p, i32p], ctypes.c_int),
        "gg_incr_destroy": ([ctypes.c_void_p], None),
        "gg_grid_program": ([ctypes.POINTER(GGGridShape), ctypes.c_int, i32p, ctypes.c_int, i32p,
                             ctypes.POINTER(GGSaParams), ctypes.c_double, EVAL_GRID_FN,
                             ctypes.c_void_p, ctypes.POINTER(GGSaStats), i32p, u64p],
                            ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != GG_OK:
        raise GGError(rc, (lib().gg_last_error() or b"").decode())


def ptr(a: np.ndarray):
    return a.ctypes.data_as(i32p)


def c32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def exported_symbols(path: str) -> set:
    """Dynamic symbols a shared library defines (read with nm -D)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
