"""GPU parity of the HBM-resident kernels (gg_wide.hip): graphs above 65,536
vertices (matrix_op, general/evaluation.c:153-373, and the grid evaluation,
grid/evaluation.c:122-190) and the reference's bfs() semantics
(general/evaluation.c:55-150) on tables where it differs from matrix_op.

The reference's own values for these inputs are in the golden CSV
(test_gpu_eval.py's golden tests run every case there); this file adds the
reference ABI for those cases, the CPU oracle on further seeded inputs, and
the error paths."""
import csv
import ctypes

import numpy as np
import pytest

import oracle
from cases import GOLDEN_CSV, build, general_cases, grid_cases
from graphgolf_amd import _lib as L
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G

pytestmark = pytest.mark.gpu

with open(GOLDEN_CSV) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}


def _case(cid, pool=None):
    return next(c for c in (pool or general_cases()) if c["id"] == cid)


def _shim_general():
    so = ctypes.CDLL(L.GENERAL_PATH)
    so.evaluation.restype = ctypes.c_bool
    so.evaluation.argtypes = [ctypes.c_int] * 5 + [L.i32p, L.i32p, ctypes.POINTER(ctypes.c_double),
                                                   ctypes.c_int, ctypes.c_int]
    return so


@pytest.mark.parametrize("cid", ["dir:g4c16:a0", "dir:g4c16:a1", "dir:g4c16:a2", "dir:raw:a0",
                                 "chain:n500:a0", "chain:n500:a1", "gasym:g4:a0", "sym:g5:a0",
                                 "ctr:big:a0", "big:131072", "big:131073", "big:ctr"])
def test_gpu_general_shim_bfs_and_large(cid):
    """The reference's evaluation() symbol (general/common.h:112-113) with
    its own algo values gives the reference's (ok, diameter, ASPL)."""
    so = _shim_general()
    case = _case(cid)
    adj, kw = build(case)
    n, d = adj.shape
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(0)
    ok = so.evaluation(n, kw["based_nodes"], kw["groups"], n * d // 2, d, L.ptr(adj),
                       ctypes.byref(diam), ctypes.byref(aspl), kw["added_centers"], case["algo"])
    row = GOLD[cid]
    assert int(ok) == int(row["ok"]) and diam.value == int(row["diameter"]), (cid, ok, diam.value)
    if ok:
        assert aspl.value == float.fromhex(row["aspl_hex"])
    else:
        assert aspl.value == E.DBL_MAX


@pytest.mark.parametrize("cid", ["grid:big:g1", "grid:big:g4"])
def test_gpu_grid_shim_large(cid):
    so = ctypes.CDLL(L.GRID_PATH)
    so.evaluation.restype = ctypes.c_bool
    so.evaluation.argtypes = [ctypes.c_int] * 3 + [L.i32p] + [ctypes.c_int] * 3 + \
        [L.i32p, ctypes.POINTER(ctypes.c_double), ctypes.c_bool, L.i32p]
    case = _case(cid, grid_cases())
    adj, kw = build(case)
    n, d = adj.shape
    diam = ctypes.c_int(0)
    aspl = ctypes.c_double(-1)
    ok = so.evaluation(n, d, kw["groups"], L.ptr(adj), kw["based_nodes"], kw["height"],
                       kw["based_height"], ctypes.byref(diam), ctypes.byref(aspl), False,
                       L.ptr(kw["rotate_hash"]))
    row = GOLD[cid]
    assert int(ok) == 1 and diam.value == int(row["diameter"])
    assert aspl.value == float.fromhex(row["aspl_hex"])


@pytest.mark.parametrize("n,d,g,c,seed", [(70000, 6, 1, 0, 3), (69632, 4, 4, 0, 4),
                                          (80000, 4, 1, 64, 5)])
def test_gpu_large_matches_oracle(n, d, g, c, seed):
    """matrix_op above 65,536 vertices against the oracle (which tiles the
    same way as the reference's matrix_op_mem_saving)."""
    e = G.random_symmetric(n, d, g, seed) if g > 1 else G.random_regular(n, d, seed)
    adj = G.edges_to_adj(n, d, e)
    B = (n - c) // g
    print(f"large n={n} d={d} g={g} c={c}", flush=True)
    r = E.apsp(adj, B, g, c)
    w = oracle.matrix_op(adj, B, g, c)
    assert r.connected == w.ok and (r.diameter, r.sum) == (w.diameter, w.S)
    assert r.word_bits == 64


def test_gpu_large_relabel_invariance():
    """Relabelling a large graph by a permutation leaves (diameter, S)
    unchanged (g = 1: every vertex is a source)."""
    n, d = 66000, 4
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 9))
    perm = np.random.default_rng(3).permutation(n).astype(np.int32)
    a = E.apsp(adj)
    b = E.apsp(adj, relabel=perm)
    assert a.connected and (a.diameter, a.sum) == (b.diameter, b.sum)


@pytest.mark.parametrize("n,d,seed,raw", [(3000, 3, 81, False), (5000, 5, 82, True),
                                          (66000, 3, 83, False)])
def test_gpu_directed_bfs_matches_oracle(n, d, seed, raw):
    """bfs() semantics on directed tables (in-neighbour CSR on the device)
    against the oracle's per-source BFS."""
    case = dict(kind="directed", n=n, d=d, seed=seed, raw=raw)
    adj, kw = build(case)
    print(f"directed bfs n={n} d={d} raw={raw}", flush=True)
    r = E.apsp(adj, bfs=True)
    w = oracle.bfs_general(adj)
    assert r.connected == w.ok
    if w.ok:
        assert (r.diameter, 2 * r.sum) == (w.diameter, w.S)


@pytest.mark.parametrize("g,c", [(1, 0), (2, 0), (4, 0), (4, 24)])
def test_gpu_bfs_semantics_undirected(g, c):
    """Undirected tables: rotation-invariant ones take the tile kernels
    (sum = S/2), the others the HBM path; both equal the oracle's bfs()."""
    n = 1200 + c
    for sym in (True, False):
        if sym and c == 0 and g > 1:
            adj = G.edges_to_adj(n, 6, G.random_symmetric(n, 6, g, 90 + g))
        else:
            adj = G.edges_to_adj(n, 6, G.random_regular(n, 6, 91 + g))
        B = (n - c) // g
        r = E.apsp(adj, B, g, c, bfs=True)
        w = oracle.bfs_general(adj, B, g, c)
        assert r.connected == w.ok and (r.diameter, 2 * r.sum) == (w.diameter, w.S), (g, c, sym)


def test_gpu_large_errors():
    """An entry outside [0,n) in a large table is an error, not a fault;
    a disconnected large graph gives the reference's sentinels."""
    n, d = 70000, 4
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 12))
    bad = adj.copy()
    bad[123, 2] = n
    with pytest.raises(L.GGError):
        E.apsp(bad)
    with pytest.raises(L.GGError):
        E.apsp(bad, bfs=True)
    half = n // 2
    e = np.concatenate([G.random_regular(half, d, 13), G.random_regular(half, d, 14) + half])
    disc = G.edges_to_adj(n, d, e)
    r = E.apsp(disc)
    assert not r.connected and r.diameter == n + 1
    assert E.evaluation(n, n, 1, n * d // 2, d, disc) == (False, E.INT_MAX, E.DBL_MAX)
    assert E.evaluation(n, n, 1, n * d // 2, d, disc, 0, E.BFS) == (False, E.INT_MAX, E.DBL_MAX)
    # a large relabel that is not a permutation
    rel = np.arange(n, dtype=np.int32)
    rel[7] = 8
    with pytest.raises(L.GGError):
        E.apsp(adj, relabel=rel)


def test_gpu_device_relabel_checked():
    """gg_apsp_device reads a device relabel table back and refuses one that
    is not a permutation (rows of the converted table would be missed).
    Device buffers come from plans (gg_plan_adj_device), uploaded with
    gg_plan_set_adj."""
    n, d = 512, 4
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 15))
    good = np.random.default_rng(1).permutation(n).astype(np.int32)
    bad = good.copy()
    bad[3] = bad[4]
    bufs = E.Plan([dict(nodes=n, degree=d), dict(nodes=n, degree=1), dict(nodes=n, degree=1)])
    for i, a in enumerate((adj, good.reshape(n, 1), bad.reshape(n, 1))):
        bufs.set_adj(i, a)
    E.synchronize()
    dev = [ctypes.cast(ctypes.c_void_p(bufs.device_adj(i)), L.i32p) for i in range(3)]
    lib = L.lib()
    out = L.GGResult()
    ref = E.apsp(adj)
    g = L.GGGraph(n, d, n, 1, 0, 0, dev[0], dev[1])
    assert lib.gg_apsp_device(ctypes.byref(g), ctypes.byref(out)) == L.GG_OK
    assert (out.diameter, out.sum) == (ref.diameter, ref.sum)
    g = L.GGGraph(n, d, n, 1, 0, 0, dev[0], dev[2])
    assert lib.gg_apsp_device(ctypes.byref(g), ctypes.byref(out)) == L.GG_ERR_INVALID
    assert b"permutation" in lib.gg_last_error()
    bufs.close()
