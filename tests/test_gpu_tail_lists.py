"""Tail lists of the tile BFS.

A pull level may record the vertices that are still not full, and the next
pull level then visits only those (when the unreached pairs fit the queue or
collapse to under half of the previous level's).  These graphs exercise:

* expanders at every word width, whose last levels run from a list;
* a long grid, where the collapse never comes and no list is built until
  the very end;
* a disconnected graph whose small second component stays unreached, so a
  list level adds nothing and must end the tile as disconnected;
* centre vertices (list increments split between base and centre rows);
* the per-vertex distance histogram, whose counts a list level must keep.

Every result is compared with the oracle.
"""
import numpy as np
import pytest

import oracle
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G

pytestmark = pytest.mark.gpu


def _same(got, want):
    assert got.connected == want.ok
    if want.ok:
        assert (got.diameter, got.sum) == (want.raw_diameter, want.S)


@pytest.mark.parametrize("n,d,bits,seed", [
    (8192, 8, 0, 1), (6000, 12, 16, 2), (4096, 16, 8, 3), (5000, 6, 32, 4), (3000, 24, 64, 5),
])
def test_expanders_match_oracle(n, d, bits, seed):
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    E.set_word_bits(bits)
    try:
        got = E.apsp(adj, symmetric=True)
    finally:
        E.set_word_bits(0)
    _same(got, oracle.matrix_op(adj))


def test_long_grid_matches_oracle():
    E.init(0)
    adj = G.edges_to_adj(3600, 4, G.random_grid(60, 60, 4, 2, 7))
    _same(E.apsp(adj, symmetric=True), oracle.matrix_op(adj))


@pytest.mark.parametrize("bits", [0, 16, 64])
def test_disconnected_small_component(bits):
    """A 4000-vertex 8-regular expander plus a separate K9: sources in the
    expander leave K9's 9 vertices unreached (a list of 9 that never fills)."""
    E.init(0)
    a = G.edges_to_adj(4000, 8, G.random_regular(4000, 8, 9))
    k9 = np.array([[4000 + j for j in range(9) if j != i] for i in range(9)], dtype=np.int32)
    adj = np.concatenate([a, k9]).astype(np.int32)
    E.set_word_bits(bits)
    try:
        got = E.apsp(adj, symmetric=True)
    finally:
        E.set_word_bits(0)
    want = oracle.matrix_op(adj)
    assert not want.ok and not got.connected
    assert got.diameter == adj.shape[0] + 1


@pytest.mark.parametrize("n,d,g,c", [(6005, 10, 4, 5), (4104, 12, 2, 8)])
def test_centres_match_oracle(n, d, g, c):
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 11))
    B = (n - c) // g
    _same(E.apsp(adj, B, g, c), oracle.matrix_op(adj, B, g, c))


def test_histogram_with_tail_lists():
    E.init(0)
    adj = G.edges_to_adj(2048, 8, G.random_regular(2048, 8, 12))
    got = E.distance_histogram(adj, 8)
    _, want = oracle.distance_histogram(adj, 8)
    assert (got == want).all()
