"""Batched / surrogate-ranked SA on the GPU: every accepted graph's exact
(diameter, S) is re-checked against the CPU oracle by replaying the logged
moves, and the trained network beats the same batched method without it."""
import numpy as np
import pytest

import oracle
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G
from graphgolf_amd import sa as SA

pytestmark = pytest.mark.gpu


def replay_and_check(e0, n, d, res):
    """Replays the logged moves in order; every move that ends an accepted
    candidate (diameter >= 0 in the log) is re-scored by the oracle."""
    e = e0.copy()
    checked = 0
    for k, (t0, t1, r) in enumerate(res.accepted):
        SA.apply_move(e, t0, t1, r)
        if res.accepted_diam[k] < 0:
            assert res.accepted_sum[k] == 0
            continue
        o = oracle.matrix_op(G.edges_to_adj(n, d, e))
        assert o.ok and o.diameter == res.accepted_diam[k] and o.S == int(res.accepted_sum[k]), k
        checked += 1
    assert checked == res.stats["accepts"]
    assert len(res.accepted) == res.stats["moves_accepted"]
    best = oracle.matrix_op(G.edges_to_adj(n, d, res.edges))
    assert best.diameter == res.stats["best_diameter"] and best.aspl == res.stats["best_aspl"]
    assert np.all(np.bincount(res.edges.ravel(), minlength=n) == d)


@pytest.mark.parametrize("surrogate,exact,combine", [(True, 16, 1), (True, 1, 4), (True, 4, 8),
                                                     (False, 16, 1)])
def test_gpu_batched_sa_every_accept_exact(surrogate, exact, combine):
    n, d = 1024, 4
    e0 = G.random_regular(n, d, 3)
    res = SA.sa_batched(e0, n, d, max_steps=120, proposals=512, exact=exact, combine=combine,
                        warmup=48, explore_every=4, use_surrogate=surrogate, seed=5,
                        train_steps=2, batch_train=128, feature_refresh=5)
    assert res.stats["accepts"] > 0
    if surrogate and combine > 1:
        assert res.stats["moves_accepted"] > res.stats["accepts"]  # merged candidates were taken
    replay_and_check(e0, n, d, res)


def test_gpu_plain_device_sa():
    n, d = 256, 8
    e0 = G.random_regular(n, d, 2)
    res = SA.sa_batched(e0, n, d, max_steps=300, proposals=1, exact=1, use_surrogate=False, seed=1)
    assert res.stats["evaluations"] == 301
    replay_and_check(e0, n, d, res)


def _balls(adj, v):
    b1 = {v} | set(adj[v].tolist())
    b2 = set(b1)
    for u in adj[v]:
        b2 |= set(adj[u].tolist())
    return b1, b2


def features_reference(adj, props, vstat):
    """Set-based restatement of the move features (float32 arithmetic)."""
    n, d = adj.shape
    f32 = np.float32
    out = np.zeros((len(props), 64), dtype=np.float32)
    for i, (a, b, c, e, r) in enumerate(props):
        for k, v in enumerate((a, b, c, e)):
            out[i, 10 * k:10 * k + 10] = vstat[v]
        pairs = [(a, b), (c, e), (a, c if r else e), (b if r else c, e if r else b)]
        for q, (x, y) in enumerate(pairs):
            b1x, b2x = _balls(adj, x)
            b1y, b2y = _balls(adj, y)
            out[i, 40 + 4 * q] = f32(len(b1x & b1y)) / f32(d)
            out[i, 41 + 4 * q] = f32(len(b1x & b2y)) / f32(d)
            out[i, 42 + 4 * q] = f32(len(b2x & b2y)) / (f32(d) * f32(d))
            out[i, 43 + 4 * q] = f32(0.5) * (vstat[x, 7] + vstat[y, 7])
        out[i, 56] = r
    return out


@pytest.mark.parametrize("n,d", [(600, 6), (4096, 16), (65536, 16)])
def test_gpu_move_features_match_set_reference(n, d):
    rng = np.random.default_rng(n + d)
    e = G.random_regular(n, d, 9)
    adj = G.edges_to_adj(n, d, e)
    props = []
    while len(props) < 24:
        t0, t1 = rng.integers(0, len(e), 2)
        a, b = e[t0]
        c, dd = e[t1]
        if len({a, b, c, dd}) == 4:
            props.append((a, b, c, dd, int(rng.integers(0, 2))))
    props = np.array(props, dtype=np.int32)
    vstat = rng.standard_normal((n, 10)).astype(np.float32)
    got = SA.sa_features(adj, props, vstat)
    want = features_reference(adj, props, vstat)
    np.testing.assert_array_equal(got, want)


def _best_sum(n, d, res):
    return E.apsp(G.edges_to_adj(n, d, res.edges)).sum


def test_gpu_surrogate_gain_over_plain_and_random_ranking():
    """Same number of SA steps (one exact evaluation each): MLP-ranked merged
    candidates lower S more than plain SA and more than the same pipeline
    with the ranking replaced by random scores."""
    n, d, steps = 4096, 16, 500
    e0 = G.random_regular(n, d, 21)
    s0 = E.apsp(G.edges_to_adj(n, d, e0)).sum
    kw = dict(max_steps=steps, exact=1, seed=2)
    sur = dict(proposals=2048, combine=8, warmup=100, explore_every=8)
    plain = SA.sa_batched(e0, n, d, use_surrogate=False, **kw)
    mlp = SA.sa_batched(e0, n, d, **sur, **kw)
    rnd = SA.sa_batched(e0, n, d, random_ranking=True, **sur, **kw)
    gains = {k: s0 - _best_sum(n, d, r) for k, r in (("plain", plain), ("mlp", mlp), ("random", rnd))}
    evals = {k: r.stats["evaluations"] for k, r in (("plain", plain), ("mlp", mlp), ("random", rnd))}
    print("S gains", gains, "evaluations", evals, "rank corr", mlp.stats["surrogate_rank_corr"])
    assert evals["mlp"] <= evals["plain"] and evals["random"] <= evals["plain"]
    assert gains["mlp"] > 2 * gains["plain"], gains
    assert gains["mlp"] > 2 * gains["random"], gains


@pytest.mark.parametrize("proposals,combine", [(4096, 16), (12288, 128)])
def test_gpu_surrogate_sa_full_size_accepts_are_exact(proposals, combine):
    """n = 65536, d = 16: every accepted candidate re-scored by the unpatched
    evaluator after replaying the logged moves; the best graph by the oracle
    (candidates of up to 16 and up to 128 merged moves)."""
    n, d = 65536, 16
    print(f"full-size surrogate SA proposals={proposals} combine={combine}", flush=True)
    e0 = G.random_regular(n, d, 5)
    res = SA.sa_batched(e0, n, d, max_steps=100, proposals=proposals, combine=combine, warmup=30,
                        explore_every=8, seed=4)
    assert res.stats["moves_accepted"] > res.stats["accepts"] > 0
    e = e0.copy()
    for k, (t0, t1, r) in enumerate(res.accepted):
        SA.apply_move(e, t0, t1, r)
        if res.accepted_diam[k] >= 0:
            g = E.apsp(G.edges_to_adj(n, d, e))
            assert (g.diameter, g.sum) == (res.accepted_diam[k], int(res.accepted_sum[k])), k
    o = oracle.matrix_op(G.edges_to_adj(n, d, res.edges))
    assert o.diameter == res.stats["best_diameter"] and o.aspl == res.stats["best_aspl"]


@pytest.mark.parametrize("surrogate", [False, True])
def test_gpu_batched_grid_sa_respects_length_and_is_exact(surrogate):
    """Grid graphs (30x30, d=4, r=2): every scored move keeps each edge within
    the length limit; every accepted candidate re-scored by the oracle."""
    w, h, d, r = 30, 30, 4, 2
    n = w * h
    e0 = G.random_grid(w, h, d, r, 3)
    res = SA.sa_batched(e0, n, d, max_steps=60, proposals=512, exact=32, combine=4, warmup=20,
                        explore_every=4, use_surrogate=surrogate, seed=6, train_steps=2,
                        batch_train=128, feature_refresh=5, grid_height=h, grid_length=r)
    assert res.stats["accepts"] > 0
    replay_and_check(e0, n, d, res)
    e = e0.copy()
    for t0, t1, v in res.accepted:
        SA.apply_move(e, t0, t1, v)
        x, y = e // h, e % h
        assert np.all(np.abs(x[:, 0] - x[:, 1]) + np.abs(y[:, 0] - y[:, 1]) <= r)


def test_gpu_grid_batch_of_256_proposals():
    """The grid workload at full size: 256 length-respecting 2-opt
    proposals of a seeded 100x100 d=4 r=2 grid scored in one plan; all 256
    equal single evaluations, 8 of them the CPU oracle."""
    from graphgolf_amd import workloads as WL

    print("grid batch 100x100 x256", flush=True)
    _, variants = WL.grid_batch(100, 100, 4, 2, 1, 256, 4)
    adjs = [G.edges_to_adj(10000, 4, e) for e in variants]
    for e in variants:
        x, y = e // 100, e % 100
        assert np.all(np.abs(x[:, 0] - x[:, 1]) + np.abs(y[:, 0] - y[:, 1]) <= 2)
    plan = E.Plan([dict(nodes=10000, degree=4, symmetric=True) for _ in adjs])
    for i, a in enumerate(adjs):
        plan.set_adj(i, a)
    res = plan.run()
    plan.close()
    for a, r in zip(adjs, res):
        s = E.apsp(a)
        assert (r.connected, r.diameter, r.sum) == (s.connected, s.diameter, s.sum)
    for k in range(0, 256, 32):
        o = oracle.matrix_op(adjs[k])
        assert res[k].connected == o.ok and (res[k].diameter, res[k].sum) == (o.diameter, o.S)


@pytest.mark.parametrize("n,d,exact,grid", [(4096, 16, 1, 0), (2048, 8, 4, 0), (900, 4, 1, 30)])
def test_gpu_plain_sa_incremental_is_the_same_run(n, d, exact, grid):
    """Plain batched SA scored by the incremental scorer makes the same run:
    the same accepted moves with the same exact (diameter, S), the same best
    graph -- general graphs (one and four candidates per step) and a grid."""
    print(f"plain sa incremental n={n} d={d} exact={exact} grid={grid}", flush=True)
    e0 = G.random_grid(30, 30, 4, 2, 3) if grid else G.random_regular(n, d, 7)
    kw = dict(max_steps=200, exact=exact, use_surrogate=False, seed=4, max_temp=20.0,
              min_temp=0.5, grid_height=grid, grid_length=2 if grid else 0)
    full = SA.sa_batched(e0, n, d, **kw)
    inc = SA.sa_batched(e0, n, d, incremental=True, **kw)
    assert full.stats["accepts"] > 0
    assert np.array_equal(full.accepted, inc.accepted)
    assert np.array_equal(full.accepted_diam, inc.accepted_diam)
    assert np.array_equal(full.accepted_sum, inc.accepted_sum)
    assert np.array_equal(full.edges, inc.edges)
    assert full.stats["best_aspl"] == inc.stats["best_aspl"]
    assert full.stats["evaluations"] == inc.stats["evaluations"]


@pytest.mark.parametrize("n,d", [(64, 3), (4096, 16), (65536, 16)])
def test_gpu_device_proposals_are_valid_two_opt_moves(n, d):
    """The device draw: (a,b) and (c,d) are the drawn lines, the four
    endpoints distinct, `valid` exactly when neither added edge exists, lines
    and variants spread evenly; the same (seed, step) gives the same draw."""
    print(f"device proposals n={n} d={d}", flush=True)
    e = G.random_regular(n, d, 9)
    count = 20000
    props, lines = SA.sa_proposals(e, n, d, count, seed=3, step=7)
    t0, t1, valid = lines[:, 0], lines[:, 1], lines[:, 2]
    a, b, c, dd, r = props.T
    assert np.array_equal(e[t0, 0], a) and np.array_equal(e[t0, 1], b)
    assert np.array_equal(e[t1, 0], c) and np.array_equal(e[t1, 1], dd)
    assert np.all(t0 != t1)
    assert np.all((a != c) & (a != dd) & (b != c) & (b != dd))
    keys = set(map(tuple, np.sort(e, axis=1).tolist()))
    q1 = np.where(r == 1, c, dd)
    p2 = np.where(r == 1, b, c)
    q2 = np.where(r == 1, dd, b)
    exists = np.array([(min(x, y), max(x, y)) in keys or (min(u, v), max(u, v)) in keys
                       for x, y, u, v in zip(a, q1, p2, q2)])
    assert np.array_equal(valid == 1, ~exists)
    assert 0.45 < r.mean() < 0.55
    lines_n = len(e)
    hist = np.bincount(np.concatenate([t0, t1]) * 16 // lines_n, minlength=16)
    assert hist.min() > 0.8 * hist.mean()
    props2, lines2 = SA.sa_proposals(e, n, d, count, seed=3, step=7)
    assert np.array_equal(props, props2) and np.array_equal(lines, lines2)
    props3, _ = SA.sa_proposals(e, n, d, count, seed=3, step=8)
    assert not np.array_equal(props, props3)


def test_gpu_device_grid_proposals_are_local_valid_moves():
    """Grid draws: line t1 has an endpoint within the length of line t0's
    first endpoint, and `valid` is exactly "both added edges are new and
    within the length"."""
    w, h, d, r = 30, 30, 4, 2
    n = w * h
    e = G.random_grid(w, h, d, r, 3)
    props, lines = SA.sa_proposals(e, n, d, 20000, seed=5, step=2, grid_height=h, grid_length=r)
    t0, t1, valid = lines.T
    a, b, c, dd, v = props.T
    assert np.array_equal(e[t0, 0], a) and np.array_equal(e[t0, 1], b)
    assert np.array_equal(e[t1, 0], c) and np.array_equal(e[t1, 1], dd)

    def dist(x, y):
        return np.abs(x // h - y // h) + np.abs(x % h - y % h)
    assert np.all((dist(a, c) <= r) | (dist(a, dd) <= r))
    keys = set(map(tuple, np.sort(e, axis=1).tolist()))
    q1 = np.where(v == 1, c, dd)
    p2 = np.where(v == 1, b, c)
    q2 = np.where(v == 1, dd, b)
    fresh = np.array([(min(x, y), max(x, y)) not in keys and (min(u, z), max(u, z)) not in keys
                      for x, y, u, z in zip(a, q1, p2, q2)])
    want = fresh & (dist(a, q1) <= r) & (dist(p2, q2) <= r)
    assert np.array_equal(valid == 1, want)
    assert valid.mean() > 0.2


def test_gpu_surrogate_sa_incremental_warmup_is_the_same_run():
    """The surrogate run with its explore steps scored incrementally (until
    the first ranked step) makes the same run as with full evaluations."""
    n, d = 4096, 16
    e0 = G.random_regular(n, d, 12)
    kw = dict(max_steps=80, proposals=2048, combine=32, warmup=40, explore_every=8, seed=3,
              train_steps=1, batch_train=128)
    full = SA.sa_batched(e0, n, d, **kw)
    inc = SA.sa_batched(e0, n, d, incremental=True, **kw)
    assert full.stats["accepts"] > 0
    assert np.array_equal(full.accepted, inc.accepted)
    assert np.array_equal(full.accepted_diam, inc.accepted_diam)
    assert np.array_equal(full.accepted_sum, inc.accepted_sum)
    assert np.array_equal(full.edges, inc.edges)


def _rot_tables(n, g, h=0):
    """Vertex rotations of a symmetric graph: general v -> v + n/g, grids
    ROTATE (graphs.rotate)."""
    if not h:
        return [(np.arange(n) + q * (n // g)) % n for q in range(g)]
    w = n // h
    return [np.arange(n)] + [np.array([G.rotate(v, h, w, g, 90 * q) for v in range(n)])
                             for q in range(1, g)]


def _check_symmetric_graph(e, n, d, g, h=0, r=0):
    keys = np.sort(e, axis=1)
    assert np.all(keys[:, 0] != keys[:, 1])
    k = keys[:, 0].astype(np.int64) * n + keys[:, 1]
    assert len(np.unique(k)) == len(k)
    assert np.all(np.bincount(e.ravel(), minlength=n) == d)
    ks = set(k.tolist())
    for rot in _rot_tables(n, g, h)[1:]:
        re = np.sort(rot[e], axis=1)
        assert set((re[:, 0].astype(np.int64) * n + re[:, 1]).tolist()) == ks
    if h:
        x, y = e // h, e % h
        assert np.all(np.abs(x[:, 0] - x[:, 1]) + np.abs(y[:, 0] - y[:, 1]) <= r)


def _block_layout(e, g):
    """The reference's line layout (copies of line t at t + i*lines/g) from
    the generator's interleaved one (copies at consecutive lines)."""
    L = len(e)
    out = np.empty_like(e)
    for orbit in range(L // g):
        for i in range(g):
            out[i * (L // g) + orbit] = e[orbit * g + i]
    return out


@pytest.mark.parametrize("kind,g,surrogate", [("general", 2, True), ("general", 4, True),
                                              ("general", 4, False), ("grid", 2, True),
                                              ("grid", 4, True), ("grid", 4, False)])
def test_gpu_batched_symmetric_sa_keeps_symmetry_and_is_exact(kind, g, surrogate):
    """Symmetric graphs in the batched SA (2g-edge moves drawn on the device):
    every accepted graph, replayed from the log, stays simple, regular,
    rotation invariant (and within the grid length), and its (diameter, S)
    equals the oracle's groups-weighted evaluation."""
    if kind == "general":
        n, d, h, r = 1024, 8, 0, 0
        e0 = G.random_symmetric(n, d, g, 5)
        if g == 4:
            e0 = _block_layout(e0, g)
        kw = {}
    else:
        w = h = 12
        n, d, r = w * h, 4, 3
        e0 = G.random_grid_symmetric(w, h, d, r, g, 5)
        kw = dict(grid_height=h, grid_length=r)
    print(f"symmetric batched sa {kind} g={g} surrogate={surrogate}", flush=True)
    _check_symmetric_graph(e0, n, d, g, h, r)
    res = SA.sa_batched(e0, n, d, groups=g, max_steps=60, proposals=512, exact=4, combine=4,
                        warmup=10, explore_every=4, use_surrogate=surrogate, seed=3, train_steps=2,
                        batch_train=128, feature_refresh=5, max_temp=10.0, min_temp=0.5, **kw)
    assert res.stats["accepts"] > 0
    rh = G.rotate_hash(n, h, n // h, g) if h else None
    e = e0.copy()
    checked = 0
    for k, (t0, t1, v) in enumerate(res.accepted):
        SA.apply_move(e, t0, t1, v, nodes=n, degree=d, groups=g, grid_height=h)
        if res.accepted_diam[k] < 0:
            continue
        _check_symmetric_graph(e, n, d, g, h, r)
        adj = G.edges_to_adj(n, d, e)
        if h:
            o = oracle.grid_evaluation(adj, g, n // g, h, h // 2, rh)
        else:
            o = oracle.matrix_op(adj, n // g, g)
        assert (o.diameter, o.S) == (res.accepted_diam[k], int(res.accepted_sum[k])), k
        checked += 1
    assert checked > 0
    adj = G.edges_to_adj(n, d, res.edges)
    o = oracle.grid_evaluation(adj, g, n // g, h, h // 2, rh) if h else oracle.matrix_op(adj, n // g, g)
    assert o.diameter == res.stats["best_diameter"] and o.aspl == res.stats["best_aspl"]


def test_gpu_symmetric_device_proposals_keep_symmetry():
    """The symmetric device draw at C2's shape (n=4096, d=16, g=4): a valid
    proposal applied with its rotated copies keeps the graph simple,
    regular and rotation invariant; the same (seed, step) draws the same."""
    n, d, g = 4096, 16, 4
    e0 = G.random_symmetric(n, d, g, 11)
    res = SA.sa_batched(e0, n, d, groups=g, max_steps=8, proposals=4096, exact=1, combine=8,
                        warmup=2, explore_every=0, use_surrogate=True, seed=9)
    e = e0.copy()
    for (t0, t1, v), dm in zip(res.accepted, res.accepted_diam):
        SA.apply_move(e, t0, t1, v, nodes=n, degree=d, groups=g)
    _check_symmetric_graph(e, n, d, g)
    o = oracle.matrix_op(G.edges_to_adj(n, d, res.edges), n // g, g)
    assert o.diameter == res.stats["best_diameter"] and o.aspl == res.stats["best_aspl"]
