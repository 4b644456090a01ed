"""The source test the incremental scorer (include/gg_incr.h) uses, checked
against the mathematics on the CPU: for random 2-opt moves of random regular
graphs, every source whose distance row changes (all-pairs BFS before and
after the move) is flagged -- including moves that disconnect -- and most
flagged sources do change.  The device kernels that evaluate the same test
are checked against full evaluation in tests/test_gpu_incr.py."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

from graphgolf_amd import graphs as G


def apsp(adj):
    n, d = adj.shape
    rows = np.repeat(np.arange(n), d)
    A = csr_matrix((np.ones(n * d), (rows, adj.ravel())), shape=(n, n))
    return shortest_path(A, unweighted=True)


def flagged(adj, D, rem, add):
    """The test: removal of (u,v), d(s,u) < d(s,v), flags s unless v has
    another neighbour one level up; an added (x,y) flags s if the levels of
    x and y differ by 2 or more."""
    n = adj.shape[0]
    aff = np.zeros(n, bool)
    for u, v in ((rem[0], rem[1]), (rem[2], rem[3])):
        for near, far in ((u, v), (v, u)):
            closer = D[:, near] < D[:, far]
            others = [w for w in adj[far] if w != near]
            other_parent = (D[:, others] == (D[:, far] - 1)[:, None]).any(axis=1)
            aff |= closer & ~other_parent
    for x, y in ((add[0], add[1]), (add[2], add[3])):
        aff |= np.abs(D[:, x] - D[:, y]) >= 2
    return aff


def moved(adj, rem, add):
    out = adj.copy()
    nb = {add[0]: add[1], add[1]: add[0], add[2]: add[3], add[3]: add[2]}
    for p, q in ((rem[0], rem[1]), (rem[1], rem[0]), (rem[2], rem[3]), (rem[3], rem[2])):
        out[p, int(np.nonzero(out[p] == q)[0][0])] = nb[p]
    return out


@pytest.mark.parametrize("n,d,seed", [(128, 3, 1), (256, 4, 2), (300, 5, 3), (512, 8, 4),
                                      (1024, 16, 5)])
def test_every_changed_source_is_flagged(n, d, seed):
    rng = np.random.default_rng(seed)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    tot_flag = tot_chg = 0
    for _ in range(25):
        D = apsp(adj)
        while True:
            a = int(rng.integers(n))
            b = int(adj[a, rng.integers(d)])
            c = int(rng.integers(n))
            e = int(adj[c, rng.integers(d)])
            add = (a, e, c, b) if rng.integers(2) == 0 else (a, c, b, e)
            if len({a, b, c, e}) == 4 and add[1] not in adj[add[0]] and add[3] not in adj[add[2]]:
                break
        rem = (a, b, c, e)
        adj2 = moved(adj, rem, add)
        D2 = apsp(adj2)
        changed = (D != D2).any(axis=1)
        aff = flagged(adj, D, rem, add)
        assert not (changed & ~aff).any(), (n, d, rem, add)
        tot_flag += aff.sum()
        tot_chg += changed.sum()
        if np.isfinite(D2).all():
            adj = adj2
    assert tot_chg >= 0.8 * tot_flag


def test_disconnecting_move_flags_every_source():
    # Two K5 blocks joined by (0,5) and (1,6): the move re-pairs them inside
    # the blocks and disconnects the graph; every distance row changes.
    e = [(i, j) for blk in (0, 5) for i in range(blk, blk + 5) for j in range(i + 1, blk + 5)]
    e = [x for x in e if x not in ((0, 1), (5, 6))] + [(0, 5), (1, 6)]
    adj = G.edges_to_adj(10, 4, np.array(e, dtype=np.int32))
    D = apsp(adj)
    rem, add = (0, 5, 1, 6), (0, 1, 5, 6)
    D2 = apsp(moved(adj, rem, add))
    assert not np.isfinite(D2).all()
    assert flagged(adj, D, rem, add).all()
