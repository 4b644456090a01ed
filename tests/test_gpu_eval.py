"""GPU parity of the evaluator: the HIP path against the reference's golden
vectors (bit-exact diameter, S and ASPL double) and against the CPU oracle on
seeded inputs, through the C ABI (libggeval.so and the reference-ABI shims)."""
import csv
import ctypes

import numpy as np
import pytest

import oracle
from cases import GOLDEN_CSV, build, c3_shapes, general_cases, grid_cases, two_opt
from graphgolf_amd import _lib as L
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G
from graphs_known import all_known

pytestmark = pytest.mark.gpu

with open(GOLDEN_CSV) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}


def same_as_row(row, ok, diameter, aspl, S):
    assert int(row["ok"]) == int(ok), row
    assert int(row["diameter"]) == diameter, row
    if int(row["ok"]):
        assert aspl == float.fromhex(row["aspl_hex"]), row
        assert S == int(row["S"]), row


@pytest.fixture(autouse=True)
def _auto_bits():
    E.set_word_bits(0)
    yield
    E.set_word_bits(0)


@pytest.mark.parametrize("case", general_cases(), ids=lambda c: c["id"])
def test_gpu_general_golden(case):
    adj, kw = build(case)
    n, d = adj.shape
    ok, diam, aspl = E.evaluation(n, kw["based_nodes"], kw["groups"], n * d // 2, d, adj,
                                  kw["added_centers"], case["algo"])
    if case["algo"] == E.BFS:  # bfs(): S = 2 x its sum over pairs s < i
        r = E.apsp(adj, kw["based_nodes"], kw["groups"], kw["added_centers"], bfs=True)
        same_as_row(GOLD[case["id"]], ok, diam, aspl, 2 * r.sum)
    else:
        r = E.apsp(adj, kw["based_nodes"], kw["groups"], kw["added_centers"])
        same_as_row(GOLD[case["id"]], ok, diam, aspl, r.sum)


@pytest.mark.parametrize("case", grid_cases(), ids=lambda c: c["id"])
def test_gpu_grid_golden(case):
    adj, kw = build(case)
    n, d = adj.shape
    ok, diam, aspl = E.grid_evaluation(n, d, kw["groups"], adj, kw["based_nodes"], kw["height"],
                                       kw["based_height"], case["bfs"], kw["rotate_hash"])
    r = E.apsp(adj, kw["based_nodes"], kw["groups"], 0, relabel=kw["rotate_hash"])
    same_as_row(GOLD[case["id"]], ok, diam, aspl, r.sum)


@pytest.mark.parametrize("name,adj,diam,S", all_known(), ids=[k[0] for k in all_known()])
def test_gpu_closed_forms(name, adj, diam, S):
    r = E.apsp(adj)
    assert r.connected and r.diameter == diam and r.sum == S


@pytest.mark.parametrize("bits", [8, 16, 32, 64])
def test_gpu_word_width_invariance(bits):
    """Every LDS word width gives the same integers (tiling invariance,
    general/evaluation.c:153-263 vs :265-373)."""
    E.set_word_bits(bits)
    for case_id in ("c1:1", "c2:11", "ctr:g4", "ctr:big", "sym:g5", "disc:1"):
        case = next(c for c in general_cases() if c["id"] == case_id)
        adj, kw = build(case)
        n, d = adj.shape
        ok, diam, aspl = E.evaluation(n, kw["based_nodes"], kw["groups"], n * d // 2, d, adj,
                                      kw["added_centers"])
        r = E.apsp(adj, kw["based_nodes"], kw["groups"], kw["added_centers"])
        same_as_row(GOLD[case_id], ok, diam, aspl, r.sum)
        if ok:
            assert r.word_bits == min(bits, r.word_bits)


@pytest.mark.parametrize("seed", range(24))
def test_gpu_random_vs_oracle(seed):
    rng = np.random.default_rng(1000 + seed)
    d = int(rng.integers(1, 33))
    n = int(rng.integers(max(d + 1, 4), 3000))
    if (n * d) % 2:
        n += 1
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed)) if d > 1 else None
    if adj is None:  # degree 1: a perfect matching plus one joining edge is not regular; use n=2
        n, adj = 2, np.array([[1], [0]], dtype=np.int32)
    g = int(rng.choice([1, 2, 3, 4]))
    c = int(rng.integers(0, 8)) if n > 64 else 0
    while (n - c) % g:
        c += 1
    B = (n - c) // g
    if c > B:
        c, g, B = 0, 1, n
    want = oracle.matrix_op(adj, B, g, c)
    got = E.apsp(adj, B, g, c)
    assert (got.connected, got.diameter if got.connected else None, got.sum if got.connected else None) == \
        (want.ok, want.raw_diameter if want.ok else None, want.S if want.ok else None)


def test_gpu_multi_edges_and_self_loops():
    """Adjacency rows with repeated entries and self-loops (both legal in the
    reference's table) evaluate like the oracle; push levels are disabled by
    the asymmetry check when rows disagree."""
    rng = np.random.default_rng(5)
    n, d = 300, 6
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 8))
    adj[::7, 0] = np.arange(0, n, 7)          # self-loops (keeps the rest symmetric? no)
    adj[3, 1] = adj[3, 2]                     # a repeated neighbour
    want = oracle.matrix_op(adj)
    got = E.apsp(adj)
    assert got.connected == want.ok
    if want.ok:
        assert (got.diameter, got.sum) == (want.diameter, want.S)
    # a directed (asymmetric) table
    adj2 = adj.copy()
    adj2[10, :] = rng.integers(0, n, d)
    want = oracle.matrix_op(adj2)
    got = E.apsp(adj2)
    assert got.connected == want.ok and (not want.ok or (got.diameter, got.sum) == (want.diameter, want.S))


def test_gpu_out_of_range_entry_is_an_error():
    adj = G.edges_to_adj(64, 4, G.random_regular(64, 4, 1))
    adj[5, 2] = 64
    with pytest.raises(L.GGError) as e:
        E.apsp(adj)
    assert e.value.code == L.GG_ERR_INVALID
    adj[5, 2] = -3
    with pytest.raises(L.GGError):
        E.apsp(adj)
    # the library stays usable afterwards
    ok = G.edges_to_adj(64, 4, G.random_regular(64, 4, 1))
    assert E.apsp(ok).connected


def test_gpu_disconnected_sentinels():
    case = next(c for c in general_cases() if c["id"] == "disc:1")
    adj, kw = build(case)
    n, d = adj.shape
    assert E.evaluation(n, n, 1, n * d // 2, d, adj) == (False, E.INT_MAX, E.DBL_MAX)
    r = E.apsp(adj)
    assert not r.connected and r.diameter == n + 1


def test_gpu_general_shim_abi():
    """The reference's evaluation() symbol in libgg_general.so, called with
    the reference's argument list (general/common.h:112-113)."""
    so = ctypes.CDLL(L.GENERAL_PATH)
    so.evaluation.restype = ctypes.c_bool
    so.evaluation.argtypes = [ctypes.c_int] * 5 + [L.i32p, L.i32p, ctypes.POINTER(ctypes.c_double),
                                                   ctypes.c_int, ctypes.c_int]
    so.init_matrix_dev(256, 8, 256, 1)
    for cid in ("c1:1", "ctr:g4", "c2:11", "disc:1"):
        case = next(c for c in general_cases() if c["id"] == cid)
        adj, kw = build(case)
        n, d = adj.shape
        diam = ctypes.c_int(0)
        aspl = ctypes.c_double(0)
        ok = so.evaluation(n, kw["based_nodes"], kw["groups"], n * d // 2, d, L.ptr(adj),
                           ctypes.byref(diam), ctypes.byref(aspl), kw["added_centers"], case["algo"])
        row = GOLD[cid]
        assert int(ok) == int(row["ok"]) and diam.value == int(row["diameter"])
        assert aspl.value == float.fromhex(row["aspl_hex"]) if ok else aspl.value == E.DBL_MAX


def test_gpu_grid_shim_abi():
    so = ctypes.CDLL(L.GRID_PATH)
    so.evaluation.restype = ctypes.c_bool
    so.evaluation.argtypes = [ctypes.c_int] * 3 + [L.i32p] + [ctypes.c_int] * 3 + \
        [L.i32p, ctypes.POINTER(ctypes.c_double), ctypes.c_bool, L.i32p]
    for cid in ("c4:1", "grid:g2:0", "grid:g4:1", "gfile:grid_w10h10d4.random.edges:1"):
        case = next(c for c in grid_cases() if c["id"] == cid)
        adj, kw = build(case)
        n, d = adj.shape
        diam = ctypes.c_int(0)
        aspl = ctypes.c_double(-1)
        ok = so.evaluation(n, d, kw["groups"], L.ptr(adj), kw["based_nodes"], kw["height"],
                           kw["based_height"], ctypes.byref(diam), ctypes.byref(aspl), case["bfs"],
                           L.ptr(kw["rotate_hash"]))
        same_as_row(GOLD[cid], ok, diam.value, aspl.value, int(GOLD[cid]["S"]))


def test_gpu_grid_disconnected_semantics():
    """grid/evaluation.c:182-185: matrix -> diameter n+1; bfs -> the largest
    eccentricity (grid/evaluation.c:95); ASPL untouched either way."""
    w, h = 8, 8
    n = w * h
    e = G.random_grid(w, h, 4, 2, 3)
    adj = G.edges_to_adj(n, 4, e)
    # cut the graph: make vertex 0 a self-looped island
    for u in adj[0]:
        row = adj[u]
        row[row == 0] = u
    adj[0, :] = 0
    ident = np.arange(n, dtype=np.int32)
    for bfs in (False, True):
        want = oracle.grid_evaluation(adj, 1, n, h, h, ident, bfs)
        ok, diam, aspl = E.grid_evaluation(n, 4, 1, adj, n, h, h, bfs, ident)
        assert not ok and not want.ok and aspl is None
        if not bfs:
            assert diam == n + 1 == want.diameter
        else:
            assert diam == want.diameter  # max eccentricity over the reached parts


def test_gpu_batch_uneven_c3():
    """512 graphs of uneven order/degree in one launch set: 16 checked
    against the golden vectors / oracle, all against single evaluations."""
    shapes = c3_shapes(512, 7)
    items, adjs = [], []
    for k, (n, d, s) in enumerate(shapes):
        if k < 12:
            e = two_opt(G.random_regular(n, d, s), n, np.random.default_rng(s), 1)
        else:
            e = G.random_regular(n, d, s)
        a = G.edges_to_adj(n, d, e)
        adjs.append(a)
        items.append(dict(adj=a, symmetric=True))
    res = E.eval_batch(items)
    for k, (n, d, s) in enumerate(shapes[:12]):
        row = GOLD[f"c3:{n}:{d}:{s}"]
        same_as_row(row, res[k].connected, res[k].diameter, res[k].aspl, res[k].sum)
    for k in (12, 100, 311, 511):
        want = oracle.matrix_op(adjs[k])
        assert (res[k].diameter, res[k].sum) == (want.diameter, want.S)
    # batch-order invariance
    perm = np.random.default_rng(1).permutation(512)[:64]
    res2 = E.eval_batch([items[i] for i in perm])
    for j, i in enumerate(perm):
        assert (res2[j].diameter, res2[j].sum) == (res[i].diameter, res[i].sum)


def test_gpu_plan_replays():
    """A captured plan replayed several times with adjacency updates and
    eager work between the replays; every replay matches the oracle."""
    shapes = [(1024, 8, 1), (2048, 12, 2), (777 * 2, 6, 3), (4096, 16, 4)]
    adjs = [G.edges_to_adj(n, d, G.random_regular(n, d, s)) for n, d, s in shapes]
    plan = E.Plan([dict(nodes=n, degree=d, symmetric=True) for n, d, s in shapes])
    rng = np.random.default_rng(3)
    for rep in range(4):
        cur = []
        for i, (n, d, s) in enumerate(shapes):
            e = G.random_regular(n, d, s + 17 * rep)
            a = G.edges_to_adj(n, d, e)
            plan.set_adj(i, a)
            cur.append(a)
        res = plan.run()
        # eager work between replays
        E.apsp(adjs[int(rng.integers(0, len(adjs)))])
        for a, r in zip(cur, res):
            want = oracle.matrix_op(a)
            assert (r.connected, r.diameter, r.sum) == (want.ok, want.diameter, want.S)
    plan.close()


def test_gpu_plan_timed_and_graph_launches_interleave():
    """Graph replays and eager timed launches of one plan alternate: kernel
    times exist only after timed launches (and reading them after a replay
    leaves no error for the next launch); a replay that meets an
    out-of-range entry reports it, and the next replay of a valid table is
    clean (the prep flags start from zero every run)."""
    n, d = 2048, 12
    a = G.edges_to_adj(n, d, G.random_regular(n, d, 8))
    want = oracle.matrix_op(a)
    plan = E.Plan([dict(nodes=n, degree=d, symmetric=True)])
    plan.set_adj(0, a)
    for rep in range(3):
        r = plan.run()[0]
        assert (r.diameter, r.sum) == (want.diameter, want.S)
        assert plan.kernel_ms() == -1.0
        plan.launch_timed()
        r = plan.wait()[0]
        assert (r.diameter, r.sum) == (want.diameter, want.S)
        assert plan.kernel_ms() > 0.0
    bad = a.copy()
    bad[7, 3] = n
    plan.set_adj(0, bad)
    with pytest.raises(L.GGError) as e:
        plan.run()
    assert e.value.code == L.GG_ERR_INVALID
    plan.set_adj(0, a)
    for launch in ("graph", "timed"):
        if launch == "graph":
            r = plan.run()[0]
        else:
            plan.launch_timed()
            r = plan.wait()[0]
        assert (r.connected, r.diameter, r.sum) == (True, want.diameter, want.S)
    plan.close()


def test_gpu_c5_full_size():
    """n=65536 d=16: golden vector (reference evaluator), plus size-independent
    properties: word-width invariance and relabelling invariance."""
    case = next(c for c in general_cases() if c["id"] == "c5:1")
    adj, kw = build(case)
    row = GOLD["c5:1"]
    r16 = E.apsp(adj)
    same_as_row(row, r16.connected, r16.diameter, r16.aspl, r16.sum)
    E.set_word_bits(8)
    r8 = E.apsp(adj)
    assert (r8.diameter, r8.sum, r8.word_bits) == (r16.diameter, r16.sum, 8)
    E.set_word_bits(0)
    perm = np.random.default_rng(2).permutation(adj.shape[0]).astype(np.int32)
    rp = E.apsp(adj, relabel=perm)
    assert (rp.diameter, rp.sum) == (r16.diameter, r16.sum)


def test_gpu_c5_symmetric_identity():
    """g=4 symmetric graph at n=65536: groups=4 formula == groups=1 formula."""
    n, d, g = 65536, 16, 4
    adj = G.edges_to_adj(n, d, G.random_symmetric(n, d, g, 3))
    a = E.apsp(adj, n // g, g)
    b = E.apsp(adj)
    assert a.connected and (a.diameter, a.sum) == (b.diameter, b.sum)


def test_gpu_distance_histogram():
    adj = G.edges_to_adj(1500, 8, G.random_regular(1500, 8, 4))
    maxd = 5
    got = E.distance_histogram(adj, maxd)
    _, want = oracle.distance_histogram(adj, maxd)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,d,maxd,seed", [(500, 6, 16, 1), (1200, 4, 5, 2), (300, 8, 3, 3)])
def test_gpu_eval_features(n, d, maxd, seed):
    """gg_eval_features: the histogram of gg_distance_histogram, eccentricities
    from per-source BFS (capped at maxd), and the evaluation's result."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    ecc, hist, res = E.features(adj, maxd)
    dist = shortest_path(csr_matrix((np.ones(n * d), (np.repeat(np.arange(n), d), adj.ravel())),
                                    shape=(n, n)), unweighted=True)
    assert (ecc == np.minimum(dist.max(axis=1), maxd).astype(np.int32)).all()
    _, want_hist = oracle.distance_histogram(adj, maxd)
    assert (hist == want_hist).all()
    want = oracle.matrix_op(adj)
    assert (res.diameter, res.sum) == (want.raw_diameter, want.S)


def test_gpu_eval_features_disconnected():
    a = G.edges_to_adj(200, 4, G.random_regular(200, 4, 4))
    k5 = np.array([[200 + j for j in range(5) if j != i] for i in range(5)], dtype=np.int32)
    adj = np.concatenate([a, k5]).astype(np.int32)
    ecc, hist, res = E.features(adj, 32)
    assert not res.connected
    assert (ecc == -1).all()  # no vertex reaches both components
    assert (hist[:200].sum(axis=1) == 200).all() and (hist[200:].sum(axis=1) == 5).all()
