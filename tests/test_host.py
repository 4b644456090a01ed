"""CPU-side checks of the product libraries: they load, export every symbol
the public headers declare, validate arguments without a device, and the
seeded graph builders produce what they promise."""
import ctypes
import os
import re

import numpy as np
import pytest

from graphgolf_amd import _lib as L
from graphgolf_amd import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text))
    return {n for n in names if n not in {"if", "for", "while", "sizeof", "return"}
            and not n.startswith("gg_eval_fn")}


def test_core_library_exports_every_declared_symbol():
    exported = L.exported_symbols(L.LIB_PATH)
    want = {n for n in declared("gg_eval.h") | declared("gg_host.h") | declared("gg_mlp.h")
            if n.startswith("gg_")}
    assert want, "no declarations parsed"
    missing = want - exported
    assert not missing, f"libggeval.so lacks {sorted(missing)}"


def test_shims_export_reference_abi():
    gen = L.exported_symbols(L.GENERAL_PATH)
    assert {"evaluation", "init_matrix_dev", "finalize_matrix_dev"} <= gen
    assert declared("gg_general.h") <= gen
    grid = L.exported_symbols(L.GRID_PATH)
    assert "evaluation" in grid and declared("gg_grid.h") <= grid
    # both shims load next to the core library
    ctypes.CDLL(L.GENERAL_PATH)
    ctypes.CDLL(L.GRID_PATH)


def _graph(n=16, d=4, B=None, g=1, c=0, adj=None, relabel=None):
    a = np.zeros((n, d), dtype=np.int32) if adj is None else adj
    gr = L.GGGraph(n, d, n if B is None else B, g, c, 0, L.ptr(a),
                   L.ptr(relabel) if relabel is not None else ctypes.cast(None, L.i32p))
    return gr, a


@pytest.mark.parametrize("kw", [dict(n=1), dict(d=0), dict(B=5), dict(g=0), dict(c=-1),
                                dict(n=16, B=4, g=4, c=1)])
def test_invalid_arguments_rejected_before_device(kw):
    gr, keep = _graph(**kw)
    out = L.GGResult()
    assert L.lib().gg_apsp(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_INVALID or \
        L.lib().gg_apsp(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_TOO_LARGE
    assert L.lib().gg_last_error()


def test_oversized_and_bfs_relabel_rejected_before_device():
    """n*d beyond 32-bit entry offsets, and the bfs() semantics with a
    relabel table, are refused before any device call (the sizes are read,
    not the tiny buffers)."""
    a = np.zeros((16, 4), dtype=np.int32)
    out = L.GGResult()
    big = L.GGGraph(1 << 29, 8, 1 << 29, 1, 0, 0, L.ptr(a), ctypes.cast(None, L.i32p))
    assert L.lib().gg_apsp(ctypes.byref(big), ctypes.byref(out)) == L.GG_ERR_TOO_LARGE
    assert b"entries" in L.lib().gg_last_error()
    rel = np.arange(16, dtype=np.int32)
    bfs = L.GGGraph(16, 4, 16, 1, 0, L.GG_GRAPH_BFS, L.ptr(a), L.ptr(rel))
    assert L.lib().gg_apsp(ctypes.byref(bfs), ctypes.byref(out)) == L.GG_ERR_INVALID
    assert b"relabel" in L.lib().gg_last_error()
    # an out-of-range host relabel entry is refused on the host as well
    rel[5] = 16
    g = L.GGGraph(16, 4, 16, 1, 0, 0, L.ptr(a), L.ptr(rel))
    assert L.lib().gg_apsp(ctypes.byref(g), ctypes.byref(out)) == L.GG_ERR_INVALID
    assert b"permutation" in L.lib().gg_last_error()


def test_row_order_mode_is_checked():
    assert L.lib().gg_set_row_order(3) == L.GG_ERR_INVALID
    assert L.lib().gg_set_row_order(-1) == L.GG_ERR_INVALID
    prev = L.lib().gg_set_row_order(2)
    assert L.lib().gg_set_row_order(prev) == 2


def test_row_order_and_features_reject_bad_arguments_before_device():
    adj = np.zeros((64, 4), dtype=np.int32)
    rows = np.zeros((64, 8), dtype=np.int32)
    cost = np.zeros(2, dtype=np.uint64)
    lib = L.lib()
    for n, d, bits in [(1, 4, 16), (64, 4, 12), (64, 40, 16), (70000, 4, 16)]:
        assert lib.gg_row_order(n, d, L.ptr(adj), bits, L.ptr(rows), cost.ctypes.data_as(L.u64p)) \
            == L.GG_ERR_INVALID
        assert lib.gg_last_error()
    ecc = np.zeros(64, dtype=np.int32)
    assert lib.gg_eval_features(64, 4, L.ptr(adj), 0, L.ptr(ecc), None, None) == L.GG_ERR_INVALID
    assert lib.gg_eval_features(1, 4, L.ptr(adj), 4, L.ptr(ecc), None, None) == L.GG_ERR_INVALID


def test_relabel_must_be_a_permutation():
    rel = np.arange(16, dtype=np.int32)
    rel[3] = 4
    gr, keep = _graph(relabel=rel)
    out = L.GGResult()
    assert L.lib().gg_apsp(ctypes.byref(gr), ctypes.byref(out)) == L.GG_ERR_INVALID
    assert b"permutation" in L.lib().gg_last_error()


def simple_regular(e, n, d):
    deg = np.bincount(e.ravel(), minlength=n)
    assert np.all(deg == d)
    assert np.all(e[:, 0] != e[:, 1])
    key = np.minimum(e[:, 0], e[:, 1]).astype(np.int64) * n + np.maximum(e[:, 0], e[:, 1])
    assert len(np.unique(key)) == len(key)


def connected(adj):
    n = adj.shape[0]
    seen = np.zeros(n, bool)
    seen[0] = True
    frontier = [0]
    while frontier:
        nxt = np.unique(adj[frontier].ravel())
        nxt = nxt[~seen[nxt]]
        seen[nxt] = True
        frontier = nxt.tolist()
    return seen.all()


@pytest.mark.parametrize("n,d", [(10, 3), (256, 8), (4096, 16), (1025, 12), (65536, 16)])
def test_random_regular(n, d):
    e = G.random_regular(n, d, 3)
    simple_regular(e, n, d)
    assert connected(G.edges_to_adj(n, d, e))
    assert np.array_equal(e, G.random_regular(n, d, 3))  # seeded
    assert not np.array_equal(e, G.random_regular(n, d, 4))


@pytest.mark.parametrize("n,d,g", [(4096, 16, 4), (600, 6, 5), (64, 3, 2), (1024, 8, 8)])
def test_random_symmetric_is_rotation_invariant(n, d, g):
    e = G.random_symmetric(n, d, g, 9)
    simple_regular(e, n, d)
    B = n // g
    key = lambda a, b: (min(a, b), max(a, b))
    s = {key(a, b) for a, b in e.tolist()}
    assert {key((a + B) % n, (b + B) % n) for a, b in e.tolist()} == s
    assert connected(G.edges_to_adj(n, d, e))


@pytest.mark.parametrize("w,h,d,r", [(100, 100, 4, 2), (10, 10, 6, 3), (12, 8, 4, 2)])
def test_random_grid(w, h, d, r):
    n = w * h
    e = G.random_grid(w, h, d, r, 1)
    simple_regular(e, n, d)
    dist = np.abs(e[:, 0] // h - e[:, 1] // h) + np.abs(e[:, 0] % h - e[:, 1] % h)
    assert dist.max() <= r
    assert connected(G.edges_to_adj(n, d, e))


def test_edges_to_adj_rejects_overfull_rows():
    e = np.array([[0, 1], [0, 2], [0, 3]], dtype=np.int32)
    with pytest.raises(L.GGError):
        G.edges_to_adj(4, 2, e)
    adj = G.edges_to_adj(4, 3, e)  # short rows padded with the vertex itself
    assert adj[1].tolist() == [0, 1, 1]


def test_file_formats_roundtrip(tmp_path):
    e = G.random_regular(64, 4, 1)
    p = tmp_path / "g.edges"
    G.write_edges(str(p), e)
    assert np.array_equal(G.read_edges(str(p)), e)
    ge = G.random_grid(6, 5, 4, 2, 2)
    q = tmp_path / "grid.edges"
    G.write_grid_edges(str(q), ge, 5)
    back, w, h = G.read_grid_edges(str(q))
    assert (w, h) == (6, 5) and np.array_equal(back, ge)


def test_rotate_hash_is_permutation():
    for (w, h, g) in [(10, 10, 2), (10, 10, 4), (8, 6, 2)]:
        rh = G.rotate_hash(w * h, h, w, g)
        assert sorted(rh.tolist()) == list(range(w * h))
