"""CPU oracle for the APSP evaluator -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline
leg import this package, and only as the checker / CPU timing reference.
The product (``graphgolf_amd``) never imports it.

``liboracle.so`` is the C restatement in ``oracle.c``; ``_ref/`` (built only
where the reference sources exist) holds the reference's own unmodified
evaluator compiled by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INT_MAX = 2**31 - 1
DBL_MAX = 1.7976931348623157e308


class _Res(ctypes.Structure):
    _fields_ = [("reached", ctypes.c_int), ("diameter", ctypes.c_int), ("S", ctypes.c_uint64)]


@dataclass
class Eval:
    """What the reference's evaluation() hands back, plus the raw integers."""

    ok: bool
    diameter: int
    aspl: float | None
    S: int
    raw_diameter: int


_lib = None
_ref = {}


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference is present)."""
    if force or not os.path.exists(os.path.join(HERE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        for name, args in {
            "orc_matrix_op": [ctypes.c_int] * 4 + [ip, ctypes.c_int, ctypes.c_int],
            "orc_bfs_general": [ctypes.c_int] * 4 + [ip, ctypes.c_int],
            "orc_grid_matrix": [ctypes.c_int] * 3 + [ip, ctypes.c_int, ip],
            "orc_grid_bfs": [ctypes.c_int] * 3 + [ip] + [ctypes.c_int] * 3,
            "orc_apsp_plain": [ctypes.c_int, ctypes.c_int, ip],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _Res
        L.orc_distance_histogram.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int32)]
        L.orc_distance_histogram.restype = ctypes.c_int
        L.orc_create_adj.argtypes = [ctypes.c_int] * 3 + [ip, ip]
        L.orc_create_adj.restype = ctypes.c_int
        L.orc_rotate_hash.argtypes = [ctypes.c_int] * 4 + [ip]
        L.orc_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        _lib = L
    return _lib


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _c32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _general(r: _Res, n: int) -> Eval:
    # general/evaluation.c:254-262: false -> (INT_MAX, DBL_MAX)
    if r.reached:
        return Eval(True, r.diameter, r.S / ((float(n) - 1) * n), int(r.S), r.diameter)
    return Eval(False, INT_MAX, DBL_MAX, int(r.S), r.diameter)


def matrix_op(adj, based_nodes=None, groups=1, added_centers=0, tile_words=0) -> Eval:
    adj = _c32(adj)
    n, d = adj.shape
    b = n if based_nodes is None else based_nodes
    return _general(lib().orc_matrix_op(n, b, groups, d, _ip(adj), added_centers, tile_words), n)


def bfs_general(adj, based_nodes=None, groups=1, added_centers=0) -> Eval:
    adj = _c32(adj)
    n, d = adj.shape
    b = n if based_nodes is None else based_nodes
    r = lib().orc_bfs_general(n, b, groups, d, _ip(adj), added_centers)
    if r.reached:  # general/evaluation.c:139-141: unordered sum over half the pairs
        return Eval(True, r.diameter, r.S / (((float(n) - 1) * n) / 2), int(r.S) * 2, r.diameter)
    return Eval(False, INT_MAX, DBL_MAX, int(r.S) * 2, r.diameter)


def evaluation(adj, based_nodes=None, groups=1, added_centers=0, algo=1) -> Eval:
    """Mirror of general evaluation() (general/evaluation.c:375-391)."""
    if algo == 0:
        return bfs_general(adj, based_nodes, groups, added_centers)
    return matrix_op(adj, based_nodes, groups, added_centers, 0 if algo == 1 else 64)


def grid_evaluation(adj, groups, based_nodes, height, based_height, rotate_hash,
                    enable_bfs=False) -> Eval:
    """Mirror of grid evaluation() (grid/evaluation.c:192-209).  On a
    disconnected graph aspl is None: the reference leaves ASPL unwritten."""
    adj = _c32(adj)
    n, d = adj.shape
    if enable_bfs:
        r = lib().orc_grid_bfs(n, d, groups, _ip(adj), based_nodes, height, based_height)
    else:
        rh = _c32(rotate_hash)
        r = lib().orc_grid_matrix(n, d, groups, _ip(adj), based_nodes, _ip(rh))
    aspl = r.S / ((float(n) - 1) * n) if r.reached else None
    return Eval(bool(r.reached), r.diameter, aspl, int(r.S), r.diameter)


def apsp_plain(adj) -> Eval:
    adj = _c32(adj)
    n, d = adj.shape
    r = lib().orc_apsp_plain(n, d, _ip(adj))
    aspl = r.S / ((float(n) - 1) * n) if r.reached else None
    return Eval(bool(r.reached), r.diameter, aspl, int(r.S), r.diameter)


def distance_histogram(adj, maxd: int):
    adj = _c32(adj)
    n, d = adj.shape
    h = np.zeros((n, maxd + 1), dtype=np.int32)
    diam = lib().orc_distance_histogram(n, d, _ip(adj), maxd,
                                        h.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return diam, h


def create_adj(nodes: int, degree: int, edges) -> np.ndarray:
    e = _c32(edges).reshape(-1, 2)
    adj = np.zeros((nodes, degree), dtype=np.int32)
    if lib().orc_create_adj(nodes, e.shape[0], degree, _ip(e), _ip(adj)) != 0:
        raise ValueError("edge list does not fit an adjacency of this degree")
    return adj


def rotate_hash(nodes: int, height: int, width: int, groups: int) -> np.ndarray:
    h = np.zeros(nodes, dtype=np.int32)
    lib().orc_rotate_hash(nodes, height, width, groups, _ip(h))
    return h


def threads() -> int:
    return lib().orc_threads()


def set_threads(k: int) -> None:
    lib().orc_set_threads(int(k))


# ---------------------------------------------------------------------------
# The reference's own evaluator (oracle/_ref), where it was built.

def ref_available(kind: str = "general") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libref_{kind}.so"))


def ref_lib(kind: str = "general"):
    if kind not in _ref:
        lib()  # load the system OpenMP runtime first; the reference build binds to it
        L = ctypes.CDLL(os.path.join(HERE, "_ref", f"libref_{kind}.so"), mode=ctypes.RTLD_LOCAL)
        L.ref_setup.restype = ctypes.c_int
        if L.ref_setup() < 1:
            raise RuntimeError("MPI singleton init failed for the reference build")
        _ref[kind] = L
    return _ref[kind]


def ref_evaluation(adj, based_nodes=None, groups=1, added_centers=0, algo=1) -> Eval:
    """Call the reference's general evaluation() (general/evaluation.c:375)."""
    L = ref_lib("general")
    adj = _c32(adj)
    n, d = adj.shape
    b = n if based_nodes is None else based_nodes
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(0)
    L.evaluation.restype = ctypes.c_bool
    L.evaluation.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_double),
                                                  ctypes.c_int, ctypes.c_int]
    lines = n * d // 2
    ok = L.evaluation(n, b, groups, lines, d, _ip(adj), ctypes.byref(diam), ctypes.byref(aspl),
                      added_centers, algo)
    S = int(round(aspl.value * n * (n - 1))) if ok else -1
    return Eval(bool(ok), diam.value, aspl.value, S, diam.value)


def ref_grid_evaluation(adj, groups, based_nodes, height, based_height, rotate_hash,
                        enable_bfs=False) -> Eval:
    """Call the reference's grid evaluation() (grid/evaluation.c:192)."""
    L = ref_lib("grid")
    adj = _c32(adj)
    n, d = adj.shape
    rh = _c32(rotate_hash)
    diam = ctypes.c_int(-7)
    aspl = ctypes.c_double(-7.0)
    L.evaluation.restype = ctypes.c_bool
    L.evaluation.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)] + \
        [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                              ctypes.c_bool, ctypes.POINTER(ctypes.c_int)]
    ok = L.evaluation(n, d, groups, _ip(adj), based_nodes, height, based_height,
                      ctypes.byref(diam), ctypes.byref(aspl), enable_bfs, _ip(rh))
    S = int(round(aspl.value * n * (n - 1))) if ok else -1
    return Eval(bool(ok), diam.value, aspl.value if ok else None, S, diam.value)


def ref_rotate_hash(nodes, height, width, groups) -> np.ndarray:
    L = ref_lib("grid")
    h = np.zeros(nodes, dtype=np.int32)
    L.create_rotate_hash.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)]
    L.create_rotate_hash(nodes, height, width, groups, _ip(h))
    return h


def ref_sa_general(edges, nodes, degree, seed, ncalcs, max_temp=100.0, min_temp=0.217147,
                   cooling_cycle=1, hill_climbing=False):
    """Run the reference's sa() (general/sa.c:413-520) from srandom(seed) on a
    copy of `edges`.  Returns (best_edges, best_diam, best_aspl, steps)."""
    import math

    L = ref_lib("general")
    e = np.ascontiguousarray(edges, dtype=np.int32).copy()
    lines = e.shape[0]
    # lower bound (general/main.c:244-265) for the "found optimum" stop
    diam_lb, n_lb, r = -1, 1, 1
    aspl_lb = 0.0
    while True:
        tmp = n_lb + degree * int(math.pow(degree - 1, r - 1))
        if tmp >= nodes:
            break
        n_lb = tmp
        aspl_lb += r * degree * math.pow(degree - 1, r - 1)
        diam_lb = r
        r += 1
    diam_lb += 1
    aspl_lb += diam_lb * (nodes - n_lb)
    aspl_lb /= (nodes - 1)
    if hill_climbing:
        max_temp = min_temp = 0.0
        rate = 1.0
    else:
        rate = math.pow(min_temp / max_temp, cooling_cycle / ncalcs) if max_temp != min_temp else 1.0
    L.sa.restype = ctypes.c_longlong
    L.sa.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_longlong, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_bool,
                                          ctypes.c_bool, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_longlong), ctypes.c_bool,
                                          ctypes.c_int]
    L.ref_srandom.argtypes = [ctypes.c_uint]
    mde = ctypes.c_double(0)
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(0)
    tot = ctypes.c_longlong(0)
    L.ref_srandom(seed)
    steps = L.sa(nodes, lines, degree, 1, max_temp, ncalcs, rate, diam_lb, aspl_lb,
                 bool(hill_climbing), False, ctypes.byref(mde), _ip(e), ctypes.byref(diam),
                 ctypes.byref(aspl), cooling_cycle, 0, nodes, ctypes.byref(tot), True, 1)
    return e, diam.value, aspl.value, int(steps)


def grid_lower_bound(width, height, degree, length):
    """Restatement of lower_bound_of_diam_aspl (grid/main.c:480-529): per
    vertex, min(Moore count, grid vertices within k*length Manhattan steps)
    vertices lie within k hops; the diameter bound is read from the last
    vertex's profile as the reference does."""
    m, n = width, height
    mn = m * n
    moore = [1, degree + 1]
    current = degree
    ii = 2
    while True:
        current *= degree - 1
        moore.append(moore[ii - 1] + current)
        if moore[ii] >= mn:
            moore[ii] = mn
            break
        ii += 1
    maxhop = max((m + n - 2 + (length - 1)) // length, ii)
    moore = np.array((moore + [mn] * (maxhop + 1))[:maxhop + 1], dtype=np.int64)
    cols, rows = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    cols, rows = cols.ravel(), rows.ravel()
    total = 0.0
    mh = None
    for i in range(m):
        for j in range(n):
            hop = (np.abs(cols - i) + np.abs(rows - j) + length - 1) // length
            hist = np.cumsum(np.bincount(hop, minlength=maxhop + 1)[:maxhop + 1])
            mh = np.minimum(hist, moore)
            for k in range(1, maxhop + 1):
                total += float(mh[k] - mh[k - 1]) * k
    dboth = int(np.argmax(mh == mn))
    return dboth, total / (float(mn) * (mn - 1))


def ref_sa_grid(edges, width, height, degree, length, groups, seed, ncalcs, max_temp=100.0,
                min_temp=0.217147, cooling_cycle=1, hill_climbing=False, fixed_temp=0.0):
    """Run the reference's grid sa() (grid/sa.c:283-361) from srandom(seed) on
    a copy of `edges`.  Returns (best_edges, best_diam, best_aspl, steps)."""
    import math

    L = ref_lib("grid")
    e = np.ascontiguousarray(edges, dtype=np.int32).copy()
    lines = e.shape[0]
    nodes = width * height
    low_diam, low_aspl = grid_lower_bound(width, height, degree, length)
    if hill_climbing:
        fixed_temp = max_temp = min_temp = 0.0
        rate = 1.0
    else:
        rate = math.pow(min_temp / max_temp, cooling_cycle / ncalcs)
    bw = width // 2 if groups == 4 else width
    bh = height // 2 if groups > 1 else height
    rh = ref_rotate_hash(nodes, height, width, groups)
    ip = ctypes.POINTER(ctypes.c_int)
    L.sa.restype = ctypes.c_longlong
    L.sa.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_longlong, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_bool, ctypes.c_bool,
                                            ctypes.c_bool, ctypes.POINTER(ctypes.c_double),
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ip, ip, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_longlong)]
                     + [ctypes.c_int] * 6 + [ip, ctypes.c_bool])
    L.ref_srandom.argtypes = [ctypes.c_uint]
    mde = ctypes.c_double(0)
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(0)
    tot = ctypes.c_longlong(0)
    L.ref_srandom(seed)
    steps = L.sa(nodes, lines, degree, nodes // groups, ncalcs, rate, low_diam, low_aspl, False,
                 bool(hill_climbing), False, ctypes.byref(mde), max_temp, min_temp, fixed_temp,
                 _ip(e), ctypes.byref(diam), ctypes.byref(aspl), cooling_cycle,
                 ctypes.byref(tot), width, bw, height, bh, length, groups, _ip(rh),
                 bool(fixed_temp > 0 and not hill_climbing))
    return e, diam.value, aspl.value, int(steps)


def nonregular_padded(adj, lens) -> np.ndarray:
    """The non-regular variants read only the first lens[v] entries of row v
    (variants/not_regular_general_graph/evaluation.c:180, :291;
    variants/not_regular_grid_graph/evaluation.c:25, :50); filling the rest
    with v itself gives a table the regular evaluators read the same way (a
    self-loop adds nothing to a reach set)."""
    a = np.array(adj, dtype=np.int32, copy=True)
    n, w = a.shape
    cols = np.arange(w)[None, :]
    mask = cols >= np.asarray(lens, dtype=np.int64)[:, None]
    a[mask] = np.broadcast_to(np.arange(n, dtype=np.int32)[:, None], (n, w))[mask]
    return a


def ref_nonregular_evaluation(adj, lens, based_nodes=None, groups=1, added_centers=0,
                              algo=1) -> Eval:
    """The reference's non-regular general evaluation()
    (variants/not_regular_general_graph/evaluation.c:378)."""
    L = ref_lib("nrgeneral")
    adj = _c32(adj)
    ln = _c32(lens)
    n, d = adj.shape
    b = n - added_centers if based_nodes is None else based_nodes
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(0)
    L.evaluation.restype = ctypes.c_bool
    L.evaluation.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)] * 2 + \
        [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    ok = L.evaluation(n, b, groups, int(ln.sum()) // 2, d, _ip(adj), ctypes.byref(diam),
                      ctypes.byref(aspl), added_centers, _ip(ln), algo)
    S = round(aspl.value * n * (n - 1)) if ok else 0
    return Eval(bool(ok), diam.value, aspl.value if ok else None, int(S), diam.value)


def ref_nonregular_grid_evaluation(adj, lens, groups, based_nodes, height, based_height,
                                   rotate_hash, enable_bfs=False) -> Eval:
    """The reference's non-regular grid evaluation()
    (variants/not_regular_grid_graph/evaluation.c:194)."""
    L = ref_lib("nrgrid")
    adj = _c32(adj)
    ln = _c32(lens)
    rh = _c32(rotate_hash)
    n, d = adj.shape
    diam = ctypes.c_int(-7)
    aspl = ctypes.c_double(-7.0)
    ip = ctypes.POINTER(ctypes.c_int)
    L.evaluation.restype = ctypes.c_bool
    L.evaluation.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, ctypes.c_int,
                             ctypes.c_int, ctypes.c_int, ip, ctypes.POINTER(ctypes.c_double),
                             ctypes.c_bool, ip]
    ok = L.evaluation(n, d, _ip(ln), groups, _ip(adj), based_nodes, height, based_height,
                      ctypes.byref(diam), ctypes.byref(aspl), enable_bfs, _ip(rh))
    S = round(aspl.value * n * (n - 1)) if ok else 0
    return Eval(bool(ok), diam.value, aspl.value if ok else None, int(S), diam.value)
