"""Bank-balanced layout (gg_bank_layout): a renumbering plus an order of each
row's entries, computed on the host.

CPU tests (the layout is host code): the relabel is a permutation that keeps
every vertex's LDS bank key; each returned row is the given row permuted;
the renumbered graph is the same graph up to isomorphism, so the oracle's
matrix_op gives the same diameter and S; the costs equal an independent
numpy recomputation of the bank cost; the layout's cost is well below the
given numbering's for random graphs; bad arguments are refused.
GPU tests: evaluations of the layout equal those of the given table at
every word width and row-order mode, and the device's cost of the laid-out
table equals the host's.
"""
import numpy as np
import pytest

import oracle
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G

SHIFT = {8: 0, 16: 1, 32: 2, 64: 2}  # LDS word index = (id << SHIFT) >> 2; bank = index mod 32


def bank_key(ids, word_bits):
    return ((np.asarray(ids, dtype=np.int64) << SHIFT[word_bits]) >> 2) & 31


def renumbered(adj, relabel):
    """Table of the renumbered graph: row relabel[v] lists relabel[adj[v]],
    in adj's order."""
    new = np.empty_like(adj)
    new[relabel] = relabel[adj]
    return new


def table_cost(rows, word_bits):
    """Sum over 32-row groups and entry slots of the most distinct LDS words
    on one bank (rows padded to a multiple of 8 with the row's own id)."""
    n, d = rows.shape
    dp = (d + 7) // 8 * 8
    full = np.empty((n, dp), dtype=np.int64)
    full[:, :d] = rows
    full[:, d:] = np.arange(n)[:, None]
    words = (full << SHIFT[word_bits]) >> 2
    total = 0
    for g0 in range(0, n, 32):
        blk = words[g0:g0 + 32]
        for s in range(dp):
            total += int(np.bincount(np.unique(blk[:, s]) & 31, minlength=32).max())
    return total


CASES = [(1024, 16, 16, 3), (2048, 8, 16, 4), (1000, 3, 16, 5), (4096, 16, 8, 6),
         (2048, 12, 32, 7), (1536, 6, 64, 8), (777, 4, 16, 9)]


@pytest.mark.parametrize("n,d,bits,seed", CASES)
def test_layout_is_a_bank_preserving_renumbering(n, d, bits, seed):
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    relabel, out, cost = E.bank_layout(adj, bits, iterations=200 * n, seed=seed, threads=2)
    assert sorted(relabel.tolist()) == list(range(n))
    assert np.array_equal(bank_key(relabel, bits), bank_key(np.arange(n), bits))
    # Each row keeps its neighbours; only their order changes.
    assert np.array_equal(np.sort(out, axis=1), np.sort(adj, axis=1))
    assert cost[0] == table_cost(adj, bits)
    assert cost[1] == table_cost(renumbered(out, relabel), bits)


@pytest.mark.parametrize("n,d,bits,seed", [(1024, 16, 16, 3), (2048, 8, 16, 4), (1000, 3, 16, 5)])
def test_layout_keeps_diameter_and_distance_sum(n, d, bits, seed):
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    relabel, out, _ = E.bank_layout(adj, bits, seed=seed, threads=2)
    a = oracle.matrix_op(adj)
    b = oracle.matrix_op(renumbered(out, relabel))
    assert (a.ok, a.diameter, a.S) == (b.ok, b.diameter, b.S)


def test_layout_lowers_bank_cost_of_random_graphs():
    # A random 16-regular graph: the busiest bank of a 32-row group holds
    # about 25 of its 512 entries, so the given numbering costs well above
    # 16 per group and slot-sum; the layout brings it close to that floor.
    n, d = 8192, 16
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 11))
    _, _, (c0, c1) = E.bank_layout(adj, 16, seed=1, threads=2)
    groups = n // 32
    assert c0 / groups > 45
    assert c1 / groups < 25


def test_layout_is_deterministic_for_a_seed_and_thread_count():
    adj = G.edges_to_adj(2048, 8, G.random_regular(2048, 8, 2))
    r1 = E.bank_layout(adj, 16, seed=5, threads=3)
    r2 = E.bank_layout(adj, 16, seed=5, threads=3)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1]) and r1[2] == r2[2]


def test_layout_refuses_bad_arguments():
    adj = G.edges_to_adj(256, 8, G.random_regular(256, 8, 1))
    with pytest.raises(Exception):
        E.bank_layout(adj, 12)
    bad = adj.copy()
    bad[3, 2] = 256
    with pytest.raises(Exception):
        E.bank_layout(bad, 16)
    with pytest.raises(Exception):
        E.bank_layout(np.zeros((256, 40), dtype=np.int32), 16)


@pytest.mark.gpu
@pytest.mark.parametrize("n,d,bits,seed", CASES)
def test_gpu_layout_evaluates_like_the_given_table(n, d, bits, seed):
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    relabel, out, cost = E.bank_layout(adj, bits, seed=seed, threads=2)
    want = E.apsp(adj)
    E.set_word_bits(bits)
    try:
        for mode in (0, 2):
            prev = E.set_row_order(mode)
            try:
                got = E.apsp(out, relabel=relabel)
            finally:
                E.set_row_order(prev)
            assert (got.connected, got.diameter, got.sum) == (want.connected, want.diameter, want.sum)
            assert got.word_bits == bits
    finally:
        E.set_word_bits(0)
    # The device's cost of the laid-out table (given order) equals the host's.
    _, dev_cost = E.row_order(renumbered(out, relabel), bits)
    assert dev_cost[0] == cost[1]


@pytest.mark.gpu
def test_gpu_layout_c5_shape_plan():
    n, d = 65536, 16
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 5))
    relabel, out, (c0, c1) = E.bank_layout(adj, 16, seed=1)
    assert c1 < 0.45 * c0
    prev = E.set_row_order(0)
    try:
        p = E.Plan([dict(nodes=n, degree=d, based_nodes=n, groups=1, symmetric=True,
                         relabel=relabel)])
        p.set_adj(0, out)
        got = p.run()[0]
        p.close()
    finally:
        E.set_row_order(prev)
    # The reference's own value for this graph (golden vector c5:1).
    assert (got.connected, got.diameter, got.sum) == (True, 6, 18691966716)
