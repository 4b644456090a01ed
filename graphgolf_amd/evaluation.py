"""Host-side mirror of the reference's evaluator interface over the gfx950
kernels in libggeval.so.

``evaluation`` / ``grid_evaluation`` keep the reference's argument meaning
and return convention (general/evaluation.c:375-391, grid/evaluation.c:192-209);
``apsp`` / ``eval_batch`` / ``Plan`` expose the raw integers (diameter, S).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

INT_MAX = 2**31 - 1
DBL_MAX = 1.7976931348623157e308
BFS, MATRIX_OP, MATRIX_OP_MEM_SAVING = 0, 1, 2  # general/common.h:58-60


@dataclass
class Result:
    connected: bool
    diameter: int      # raw level count; nodes+1 when disconnected
    max_ecc: int
    sum: int           # S, the ordered-pair distance sum (symmetry weighted)
    word_bits: int
    nodes: int

    @property
    def aspl(self) -> float:
        n = self.nodes
        return self.sum / ((float(n) - 1) * n)


def init(device: int = -1) -> None:
    L.check(L.lib().gg_init(device))


def _graph(adj: np.ndarray, based_nodes, groups, added_centers, relabel, symmetric=False,
           bfs=False):
    n, d = adj.shape
    g = L.GGGraph()
    g.nodes, g.degree = n, d
    g.groups = groups
    g.added_centers = added_centers
    g.based_nodes = (n - added_centers) // groups if based_nodes is None else based_nodes
    g.flags = (L.GG_GRAPH_SYMMETRIC if symmetric else 0) | (L.GG_GRAPH_BFS if bfs else 0)
    g.adj = L.ptr(adj)
    g.relabel = L.ptr(relabel) if relabel is not None else ctypes.cast(None, L.i32p)
    return g


def _result(r: L.GGResult, n: int) -> Result:
    return Result(bool(r.connected), int(r.diameter), int(r.max_ecc), int(r.sum),
                  int(r.word_bits), n)


def apsp(adj, based_nodes: Optional[int] = None, groups: int = 1, added_centers: int = 0,
         relabel=None, symmetric: bool = False, bfs: bool = False) -> Result:
    """One evaluation on the GPU (adjacency from host memory).  bfs=True
    scores with the reference's bfs() (general/evaluation.c:55-150): `sum` is
    then its weighted sum over pairs s < i (ASPL = sum / (n(n-1)/2))."""
    adj = L.c32(adj)
    rel = L.c32(relabel) if relabel is not None else None
    g = _graph(adj, based_nodes, groups, added_centers, rel, symmetric, bfs)
    r = L.GGResult()
    L.check(L.lib().gg_apsp(ctypes.byref(g), ctypes.byref(r)))
    return _result(r, adj.shape[0])


def apsp_device(adj_dev_ptr: int, nodes: int, degree: int, based_nodes=None, groups=1,
                added_centers=0, symmetric=False) -> Result:
    """One evaluation whose adjacency already sits in device memory (an
    int32 [nodes][degree] buffer at ``adj_dev_ptr``, e.g. a torch tensor's
    data_ptr())."""
    g = L.GGGraph()
    g.nodes, g.degree, g.groups, g.added_centers = nodes, degree, groups, added_centers
    g.based_nodes = (nodes - added_centers) // groups if based_nodes is None else based_nodes
    g.flags = L.GG_GRAPH_SYMMETRIC if symmetric else 0
    g.adj = ctypes.cast(ctypes.c_void_p(adj_dev_ptr), L.i32p)
    g.relabel = ctypes.cast(None, L.i32p)
    r = L.GGResult()
    L.check(L.lib().gg_apsp_device(ctypes.byref(g), ctypes.byref(r)))
    return _result(r, nodes)


def eval_batch(graphs: Sequence[dict]) -> list:
    """Score many graphs in one launch set.  Each item: dict(adj=..., and
    optionally based_nodes, groups, added_centers, relabel, symmetric)."""
    keep = []
    arr = (L.GGGraph * len(graphs))()
    for i, item in enumerate(graphs):
        adj = L.c32(item["adj"])
        rel = item.get("relabel")
        rel = L.c32(rel) if rel is not None else None
        keep.append((adj, rel))
        arr[i] = _graph(adj, item.get("based_nodes"), item.get("groups", 1),
                        item.get("added_centers", 0), rel, item.get("symmetric", False))
    out = (L.GGResult * len(graphs))()
    L.check(L.lib().gg_eval_batch(len(graphs), arr, out))
    return [_result(out[i], keep[i][0].shape[0]) for i in range(len(graphs))]


class Plan:
    """A batch of fixed graph shapes with device buffers and a captured
    hipGraph; ``run`` uploads nothing, ``set_adj`` uploads one slot."""

    def __init__(self, shapes: Sequence[dict]):
        self.n = len(shapes)
        arr = (L.GGGraph * self.n)()
        self._keep = []
        self.nodes = []
        for i, s in enumerate(shapes):
            nodes, degree = s["nodes"], s["degree"]
            dummy = np.zeros((nodes, degree), dtype=np.int32)
            rel = s.get("relabel")
            rel = L.c32(rel) if rel is not None else None
            self._keep.append((dummy, rel))
            arr[i] = _graph(dummy, s.get("based_nodes"), s.get("groups", 1),
                            s.get("added_centers", 0), rel, s.get("symmetric", False))
            self.nodes.append(nodes)
        p = L.lib().gg_plan_create(self.n, arr)
        if not p:
            raise L.GGError(L.lib().gg_last_error_code() or -4,
                            (L.lib().gg_last_error() or b"").decode())
        self._p = ctypes.c_void_p(p)
        self._out = (L.GGResult * self.n)()

    def set_adj(self, index: int, adj) -> None:
        a = L.c32(adj)
        L.check(L.lib().gg_plan_set_adj(self._p, index, L.ptr(a)))
        self._keep.append((a, None))  # keep alive until the async copy is done

    def device_adj(self, index: int) -> int:
        return int(L.lib().gg_plan_adj_device(self._p, index))

    def run(self) -> list:
        L.check(L.lib().gg_plan_run(self._p, self._out))
        self._keep = self._keep[: self.n]
        return [_result(self._out[i], self.nodes[i]) for i in range(self.n)]

    def launch(self) -> None:
        L.check(L.lib().gg_plan_launch(self._p))

    def launch_timed(self) -> None:
        """Enqueue the same kernels eagerly with HIP events around the tile BFS."""
        L.check(L.lib().gg_plan_launch_timed(self._p))

    def wait(self) -> list:
        L.check(L.lib().gg_plan_wait(self._p, self._out))
        self._keep = self._keep[: self.n]
        return [_result(self._out[i], self.nodes[i]) for i in range(self.n)]

    def kernel_ms(self) -> float:
        """Tile-BFS kernel time (ms, HIP events) of the last completed
        launch_timed(); -1 after run()/launch(), whose graph carries no events."""
        return float(L.lib().gg_plan_kernel_ms(self._p))

    def config(self) -> dict:
        """The launch configuration the plan was built with."""
        cfg = np.zeros(4, dtype=np.int32)
        L.check(L.lib().gg_plan_config(self._p, cfg.ctypes.data_as(L.i32p)))
        return dict(zip(CONFIG_KEYS, (int(x) for x in cfg)))

    def close(self) -> None:
        if getattr(self, "_p", None):
            L.lib().gg_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace_levels(adj, units: int = 8, based_nodes=None, groups=1, added_centers=0,
                 symmetric=False, relabel=None):
    """Per-level shader-cycle timings of the first `units` source tiles of one
    evaluation: array [units, 64, 4] of (cycles since start at the level's
    end, mode word, cycles when wave 0 finished the level's reads, cycles
    when the barrier after them released)."""
    adj = L.c32(adj)
    rel = L.c32(relabel) if relabel is not None else None
    g = _graph(adj, based_nodes, groups, added_centers, rel, symmetric)
    out = np.zeros((units, 64, 4), dtype=np.uint64)
    r = L.GGResult()
    L.check(L.lib().gg_trace_levels(ctypes.byref(g), units,
                                    out.ctypes.data_as(L.u64p), ctypes.byref(r)))
    return out, _result(r, adj.shape[0])


def synchronize() -> None:
    L.check(L.lib().gg_synchronize())


CONFIG_KEYS = ("word_bits", "block", "row_order", "persistent")


def autotune(nodes: int, degree: int, based_nodes=None, groups: int = 1, batch: int = 1,
             reps: int = 5):
    """Time every candidate configuration (word width, block, row order,
    persistent batch launch) on a seeded graph of this shape (`batch` copies);
    the fastest is used for the shape's bucket from then on.  Returns
    (bits, ms per replay)."""
    bits = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    b = nodes // groups if based_nodes is None else based_nodes
    L.check(L.lib().gg_autotune(nodes, degree, b, groups, batch, reps, ctypes.byref(bits),
                                ctypes.byref(ms)))
    return bits.value, ms.value


def autotune_batch(graphs: Sequence[dict], reps: int = 3):
    """Tune on the caller's graphs (items as eval_batch); the winner is kept
    under the bucket of the largest.  Returns (config dict, ms per replay)."""
    keep = []
    arr = (L.GGGraph * len(graphs))()
    for i, item in enumerate(graphs):
        adj = L.c32(item["adj"])
        rel = item.get("relabel")
        rel = L.c32(rel) if rel is not None else None
        keep.append((adj, rel))
        arr[i] = _graph(adj, item.get("based_nodes"), item.get("groups", 1),
                        item.get("added_centers", 0), rel, item.get("symmetric", False))
    cfg = np.zeros(4, dtype=np.int32)
    ms = ctypes.c_double(0)
    L.check(L.lib().gg_autotune_batch(len(graphs), arr, reps, cfg.ctypes.data_as(L.i32p),
                                      ctypes.byref(ms)))
    return dict(zip(CONFIG_KEYS, (int(x) for x in cfg))), ms.value


def tune_lookup(nodes: int, degree: int, based_nodes=None, groups: int = 1, added_centers: int = 0,
                batched: bool = False) -> dict:
    """The configuration an evaluation (or a batch led by this shape) uses now."""
    cfg = np.zeros(4, dtype=np.int32)
    b = (nodes - added_centers) // groups if based_nodes is None else based_nodes
    L.check(L.lib().gg_tune_lookup(nodes, degree, b, groups, added_centers, int(batched),
                                   cfg.ctypes.data_as(L.i32p)))
    return dict(zip(CONFIG_KEYS, (int(x) for x in cfg)))


def tune_table_save(path: str) -> None:
    L.check(L.lib().gg_tune_table_save(path.encode()))


def tune_table_clear() -> None:
    L.check(L.lib().gg_tune_table_clear())


def tune_table_load(path: str) -> int:
    n = L.lib().gg_tune_table_load(path.encode())
    if n < 0:
        L.check(n)
    return n


def set_row_order(mode) -> int:
    """Bank-aware row order of plans and evaluations made afterwards:
    0/False off, 1/True graphs of at least 1024 source tiles and
    incremental scorers of at least 8192 vertices (the default), 2 every graph;
    returns the previous mode.  Results never depend on it."""
    rc = L.lib().gg_set_row_order(int(mode))
    if rc < 0:
        L.check(rc)
    return rc


def row_order(adj, word_bits: int = 16):
    """The converted row table of one graph as an evaluation orders it for
    `word_bits`-bit reach words: (rows [n][dp] int32, (cost of the given
    order, cost of the returned order)), cost = sum over 32-row groups and
    entry slots of the most distinct 4-byte LDS words on one of 32 banks."""
    adj = L.c32(adj)
    n, d = adj.shape
    dp = (d + 7) // 8 * 8
    rows = np.zeros((n, dp), dtype=np.int32)
    cost = np.zeros(2, dtype=np.uint64)
    L.check(L.lib().gg_row_order(n, d, adj.ctypes.data_as(L.i32p), word_bits,
                                 rows.ctypes.data_as(L.i32p), cost.ctypes.data_as(L.u64p)))
    return rows, (int(cost[0]), int(cost[1]))


def bank_layout(adj, word_bits: int = 16, iterations: int = 0, seed: int = 1, threads: int = 0):
    """Bank-balanced layout of one graph (include/gg_eval.h, gg_bank_layout):
    (relabel, adj_ordered, (cost of the given table, cost of the layout)).
    Evaluate ``adj_ordered`` with ``relabel``: the same graph, renumbered and
    with each row's entries in slot order; results equal those of ``adj``."""
    adj = L.c32(adj)
    n, d = adj.shape
    relabel = np.zeros(n, dtype=np.int32)
    out = np.zeros((n, d), dtype=np.int32)
    cost = np.zeros(2, dtype=np.uint64)
    L.check(L.lib().gg_bank_layout(n, d, adj.ctypes.data_as(L.i32p), word_bits, int(iterations),
                                   int(seed) & 0xffffffff, int(threads),
                                   relabel.ctypes.data_as(L.i32p), out.ctypes.data_as(L.i32p),
                                   cost.ctypes.data_as(L.u64p)))
    return relabel, out, (int(cost[0]), int(cost[1]))


def set_word_bits(bits: int) -> None:
    """Force the LDS word width (8/16/32/64 sources per word); 0 = automatic."""
    L.check(L.lib().gg_set_word_bits(bits))


def features(adj, maxd: int):
    """Per-vertex features of one evaluation: (ecc, hist, result) with
    ecc[v] = eccentricity of v (maxd when some vertex is at distance >= maxd,
    -1 when v does not reach every vertex) and hist as distance_histogram."""
    adj = L.c32(adj)
    n, d = adj.shape
    ecc = np.zeros(n, dtype=np.int32)
    hist = np.zeros((n, maxd + 1), dtype=np.int32)
    r = L.GGResult()
    L.check(L.lib().gg_eval_features(n, d, adj.ctypes.data_as(L.i32p), maxd, ecc.ctypes.data_as(L.i32p),
                                     hist.ctypes.data_as(L.i32p), ctypes.byref(r)))
    return ecc, hist, _result(r, n)


def distance_histogram(adj, maxd: int) -> np.ndarray:
    """hist[v, k] = vertices at distance exactly k from v (k < maxd);
    hist[v, maxd] = vertices at distance >= maxd."""
    adj = L.c32(adj)
    n, d = adj.shape
    h = np.zeros((n, maxd + 1), dtype=np.int32)
    L.check(L.lib().gg_distance_histogram(n, d, L.ptr(adj), maxd, L.ptr(h)))
    return h


# ---- The reference's entry points, same argument meaning -------------------

def evaluation(nodes: int, based_nodes: int, groups: int, lines: int, degree: int, adj,
               added_centers: int = 0, algo: int = MATRIX_OP):
    """general/evaluation.c:375-391 -> (connected, diameter, ASPL); a
    disconnected graph gives (False, INT_MAX, DBL_MAX).  algo = BFS scores
    with bfs() (general/evaluation.c:55-150), MATRIX_OP and
    MATRIX_OP_MEM_SAVING with matrix_op (:153-373)."""
    a = L.c32(adj).reshape(nodes, degree)
    if algo == BFS:
        r = apsp(a, based_nodes, groups, added_centers, bfs=True)
        if r.connected:
            return True, r.diameter, r.sum / (((float(nodes) - 1) * nodes) / 2)
        return False, INT_MAX, DBL_MAX
    r = apsp(a, based_nodes, groups, added_centers)
    if r.connected and r.diameter < nodes:
        return True, r.diameter, r.sum / ((float(nodes) - 1) * nodes)
    return False, INT_MAX, DBL_MAX


def grid_evaluation(nodes: int, degree: int, groups: int, adjacency, based_nodes: int,
                    height: int, based_height: int, enable_bfs: bool, rotate_hash):
    """grid/evaluation.c:192-209 -> (connected, diameter, ASPL or None).
    Disconnected: diameter = nodes+1 (matrix) or the largest eccentricity
    (enable_bfs); ASPL is not produced."""
    a = L.c32(adjacency).reshape(nodes, degree)
    r = apsp(a, based_nodes, groups, 0, relabel=rotate_hash)
    if r.connected and not (r.diameter > nodes):
        return True, r.diameter, r.sum / ((float(nodes) - 1) * nodes)
    return False, (r.max_ecc if enable_bfs else nodes + 1), None


def _rows_padded(adj, lens) -> np.ndarray:
    """Row v keeps its first lens[v] entries; the unused slots get v itself
    (what the non-regular shims hand to the device)."""
    a = np.array(L.c32(adj), copy=True)
    n, w = a.shape
    ln = np.asarray(lens, dtype=np.int64)
    if ln.shape != (n,) or (ln < 0).any() or (ln > w).any():
        raise ValueError("each row length must be in [0, width]")
    mask = np.arange(w)[None, :] >= ln[:, None]
    a[mask] = np.broadcast_to(np.arange(n, dtype=np.int32)[:, None], (n, w))[mask]
    return a


def evaluation_nonregular(nodes: int, based_nodes: int, groups: int, lines: int, degree: int,
                          adj, num_degrees, added_centers: int = 0, algo: int = MATRIX_OP):
    """variants/not_regular_general_graph/evaluation.c:378-396: row v of adj
    ([nodes][degree]) uses its first num_degrees[v] entries; results as
    evaluation()."""
    a = _rows_padded(L.c32(adj).reshape(nodes, degree), num_degrees)
    return evaluation(nodes, based_nodes, groups, lines, degree, a, added_centers, algo)


def grid_evaluation_nonregular(nodes: int, max_degree: int, degree, groups: int, adjacency,
                               based_nodes: int, height: int, based_height: int,
                               enable_bfs: bool, rotate_hash):
    """variants/not_regular_grid_graph/evaluation.c:194-208: row v of
    adjacency ([nodes][max_degree]) uses its first degree[v] entries; results
    as grid_evaluation()."""
    a = _rows_padded(L.c32(adjacency).reshape(nodes, max_degree), degree)
    return grid_evaluation(nodes, max_degree, groups, a, based_nodes, height, based_height,
                           enable_bfs, rotate_hash)
