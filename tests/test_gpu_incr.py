"""Incremental re-scoring (include/gg_incr.h) against full evaluation: for
random 2-opt moves on random regular graphs, every scored move's (connected,
diameter, S) equals gg_apsp() on the moved adjacency, through random
accept/reject sequences, at n from 64 to 65536 and degrees 3 to 32 (one to
three probe units per endpoint); disconnecting moves are
reported disconnected; the scorer's own graph and state follow the accepted
moves.  The oracle checks the small cases directly."""
import numpy as np
import pytest

import oracle
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G
from graphgolf_amd.incr import IncrementalScorer

pytestmark = pytest.mark.gpu


def random_move(adj, rng):
    """A random 2-opt of the graph adj: (rem, add) with distinct endpoints,
    no added edge already present."""
    n, d = adj.shape
    while True:
        a = int(rng.integers(n))
        b = int(adj[a, rng.integers(d)])
        c = int(rng.integers(n))
        dd = int(adj[c, rng.integers(d)])
        if len({a, b, c, dd}) < 4:
            continue
        add = (a, dd, c, b) if rng.integers(2) == 0 else (a, c, b, dd)
        if add[1] in adj[add[0]] or add[3] in adj[add[2]]:
            continue
        return (a, b, c, dd), add


def apply_move(adj, rem, add):
    out = adj.copy()
    a, b, c, d = rem
    new_nb = {}
    for x, y in ((add[0], add[1]), (add[2], add[3])):
        new_nb[x] = y
        new_nb[y] = x
    for p, q in ((a, b), (b, a), (c, d), (d, c)):
        j = int(np.nonzero(out[p] == q)[0][0])
        out[p, j] = new_nb[p]
    return out


def check_run(n, d, seed, moves, p_accept, use_oracle=False, oracle_every=0):
    rng = np.random.default_rng(seed)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    full = E.apsp(adj)
    with IncrementalScorer(adj) as sc:
        st = sc.state()
        assert (st.diameter, st.sum) == (full.diameter, full.sum)
        affected = []
        for _ in range(moves):
            rem, add = random_move(adj, rng)
            moved = apply_move(adj, rem, add)
            want = E.apsp(moved)
            got, aff = sc.score(rem, add)
            assert got.connected == want.connected, (n, d, rem, add)
            if want.connected:
                assert (got.diameter, got.sum) == (want.diameter, want.sum), (n, d, rem, add)
                if use_oracle:
                    o = oracle.apsp_plain(moved)
                    assert (got.diameter, got.sum) == (o.diameter, o.S)
                elif oracle_every and len(affected) % oracle_every == 0:
                    o = oracle.matrix_op(moved)
                    assert (got.diameter, got.sum) == (o.diameter, o.S), (n, d, rem, add)
            affected.append(aff)
            if want.connected and rng.random() < p_accept:
                sc.accept()
                adj = moved
                st = sc.state()
                assert (st.diameter, st.sum) == (want.diameter, want.sum)
                assert np.array_equal(sc.adjacency(), adj)
            else:
                sc.reject()
        return np.array(affected)


@pytest.mark.parametrize("n,d,seed", [(64, 3, 1), (256, 8, 2), (500, 5, 3), (1024, 16, 4),
                                      (2048, 12, 5), (512, 17, 11), (1024, 24, 9), (1024, 31, 13),
                                      (2048, 32, 10)])
def test_incremental_matches_full_small(n, d, seed):
    print(f"incr n={n} d={d} seed={seed}", flush=True)
    check_run(n, d, seed, moves=60, p_accept=0.5, use_oracle=n <= 1024)


@pytest.mark.parametrize("n,d,seed", [(4096, 16, 6), (16384, 8, 7), (8192, 24, 12), (16384, 32, 14)])
def test_incremental_matches_full_medium(n, d, seed):
    """Every move against the device's full evaluation, every 5th against
    the CPU oracle too."""
    print(f"incr n={n} d={d} seed={seed}", flush=True)
    check_run(n, d, seed, moves=30, p_accept=0.4, oracle_every=5)


def test_incremental_matches_full_c5():
    print("incr n=65536 d=16", flush=True)
    aff = check_run(65536, 16, 8, moves=12, p_accept=0.5)
    # A 2-opt of a random 16-regular graph touches about 15 % of the sources.
    assert 0.05 * 65536 < np.median(aff) < 0.3 * 65536


def test_disconnecting_move_is_reported():
    # Two 4-regular blocks joined by two edges (0-8 and 1-9); removing both
    # and re-pairing inside the blocks disconnects the graph.
    rng = np.random.default_rng(0)
    print("incr disconnecting move", flush=True)
    for _ in range(20):
        e1 = G.random_regular(8, 4, int(rng.integers(1 << 30)))
        e2 = G.random_regular(8, 4, int(rng.integers(1 << 30))) + 8
        # break one edge in each block to attach the joins: (a,b) in block 1,
        # (c,d) in block 2 become (a,c), (b,d)
        a, b = e1[0]
        c, dd = e2[0]
        e = np.concatenate([e1[1:], e2[1:], [[a, c], [b, dd]]]).astype(np.int32)
        adj = G.edges_to_adj(16, 4, e)
        if not E.apsp(adj).connected:
            continue
        with IncrementalScorer(adj) as sc:
            got, aff = sc.score((a, c, b, dd), (a, b, c, dd))
            want = E.apsp(apply_move(adj, (a, c, b, dd), (a, b, c, dd)))
            assert not want.connected and not got.connected
            sc.reject()
            st = sc.state()
            assert st.sum == E.apsp(adj).sum
        return
    pytest.fail("no connected start graph was built")


def test_invalid_moves_are_refused():
    adj = G.edges_to_adj(64, 4, G.random_regular(64, 4, 3))
    with IncrementalScorer(adj) as sc:
        a, b = 0, int(adj[0, 0])
        with pytest.raises(Exception, match="distinct"):
            sc.score((a, b, a, b), (a, b, a, b))
        nb = set(adj[0])
        c = next(v for v in range(64) if v not in nb and v != 0 and v != b)
        with pytest.raises(Exception, match="does not exist|already exists|distinct"):
            sc.score((0, c, b, int(adj[b, 1])), (0, b, c, int(adj[b, 1])))
        with pytest.raises(Exception, match="pending|no pending"):
            sc.accept()
