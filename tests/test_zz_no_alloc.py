"""The allocation-free evaluation (gg_apsp_device_noalloc) against
gg_apsp_device, through the C ABI, on device-resident inputs: consecutive
calls of shrinking and growing sizes share the library's scratch array, so
each one must clear what it reads."""
import csv
import ctypes

import numpy as np
import pytest

from cases import GOLDEN_CSV, build, general_cases, grid_cases
from graphgolf_amd import _lib as L
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G

pytestmark = pytest.mark.gpu

with open(GOLDEN_CSV) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}


def _lib():
    lib = L.lib()
    lib.gg_apsp_device_noalloc.restype = ctypes.c_int
    lib.gg_apsp_device_noalloc.argtypes = [ctypes.POINTER(L.GGGraph), ctypes.POINTER(L.GGResult)]
    return lib


class DeviceGraphs:
    """Device copies of adjacency (and relabel) tables, held in a plan's
    buffers (gg_plan_adj_device) and uploaded with gg_plan_set_adj."""

    def __init__(self, tables):
        shapes = [dict(nodes=t.shape[0], degree=t.shape[1]) for t in tables]
        self.plan = E.Plan(shapes)
        for i, t in enumerate(tables):
            self.plan.set_adj(i, t)
        E.synchronize()
        self.ptr = [ctypes.cast(ctypes.c_void_p(self.plan.device_adj(i)), L.i32p)
                    for i in range(len(tables))]

    def close(self):
        self.plan.close()


def _both(adj_ptr, n, d, B, g, c, rel_ptr=None, flags=0):
    lib = _lib()
    rel = rel_ptr if rel_ptr is not None else ctypes.cast(None, L.i32p)
    gr = L.GGGraph(n, d, B, g, c, flags, adj_ptr, rel)
    a, b = L.GGResult(), L.GGResult()
    ra = lib.gg_apsp_device_noalloc(ctypes.byref(gr), ctypes.byref(a))
    rb = lib.gg_apsp_device(ctypes.byref(gr), ctypes.byref(b))
    return ra, a, rb, b


def _same(a, b):
    return (a.connected, a.diameter, a.max_ecc, a.sum, a.word_bits) == \
        (b.connected, b.diameter, b.max_ecc, b.sum, b.word_bits)


def test_noalloc_matches_device_entry_point_across_sizes():
    """Large, then smaller, then the largest, then small again: every call
    equals gg_apsp_device on the same device table and the golden S."""
    ids = ["c2:11", "c1:1", "c5:1", "ctr:g4", "disc:1", "sym:g5"]
    built = [build(next(c for c in general_cases() if c["id"] == cid)) for cid in ids]
    dev = DeviceGraphs([adj for adj, _ in built])
    try:
        for k, (cid, (adj, kw)) in enumerate(zip(ids, built)):
            n, d = adj.shape
            print(f"noalloc {cid} n={n} d={d}", flush=True)
            for flags in (0, L.GG_GRAPH_SYMMETRIC):
                ra, a, rb, b = _both(dev.ptr[k], n, d, kw["based_nodes"], kw["groups"],
                                     kw["added_centers"], flags=flags)
                assert ra == L.GG_OK and rb == L.GG_OK, L.lib().gg_last_error()
                assert _same(a, b), (cid, flags)
                row = GOLD[cid]
                ok = a.connected and a.diameter < n
                assert int(ok) == int(row["ok"])
                if ok:
                    assert a.sum == int(row["S"]) and a.diameter == int(row["diameter"])
    finally:
        dev.close()


def test_noalloc_relabel_checked_on_device():
    """A relabelled grid matches; a relabel table with a repeated entry is
    refused; the next call (smaller, no relabel) is unaffected."""
    case = next(c for c in grid_cases() if c["id"] == "grid:g4:0")
    adj, kw = build(case)
    n, d = adj.shape
    bad = kw["rotate_hash"].copy()
    bad[5] = bad[6]
    small = G.edges_to_adj(128, 4, G.random_regular(128, 4, 3))
    dev = DeviceGraphs([adj, kw["rotate_hash"].reshape(n, 1), bad.reshape(n, 1), small])
    try:
        ra, a, rb, b = _both(dev.ptr[0], n, d, kw["based_nodes"], kw["groups"], 0, dev.ptr[1])
        assert ra == L.GG_OK and rb == L.GG_OK and _same(a, b)
        assert a.sum == int(GOLD["grid:g4:0"]["S"])
        lib = _lib()
        out = L.GGResult()
        gr = L.GGGraph(n, d, kw["based_nodes"], kw["groups"], 0, 0, dev.ptr[0], dev.ptr[2])
        assert lib.gg_apsp_device_noalloc(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_INVALID
        ra, a, rb, b = _both(dev.ptr[3], 128, 4, 128, 1, 0)
        assert ra == L.GG_OK and rb == L.GG_OK and _same(a, b)
    finally:
        dev.close()


def test_noalloc_refuses_beyond_its_scratch():
    lib = _lib()
    a = np.zeros((16, 4), dtype=np.int32)
    out = L.GGResult()
    for n, d, flags in ((70000, 4, 0), (256, 40, 0)):
        gr = L.GGGraph(n, d, n, 1, 0, flags, L.ptr(a), ctypes.cast(None, L.i32p))
        assert lib.gg_apsp_device_noalloc(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_TOO_LARGE
        assert b"scratch" in lib.gg_last_error() or b"device limit" in lib.gg_last_error() or \
            b"limit" in lib.gg_last_error()
    gr = L.GGGraph(16, 4, 16, 1, 0, L.GG_GRAPH_BFS, L.ptr(a), ctypes.cast(None, L.i32p))
    assert lib.gg_apsp_device_noalloc(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_INVALID
