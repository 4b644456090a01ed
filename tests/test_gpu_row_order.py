"""Bank-aware row order of the converted adjacency (gg_order_rows_kernel).

Every returned row must be a permutation of the adjacency row plus its
padding (the row's own id), so no reach set can change; the returned costs
must equal the LDS bank cost recomputed here from the rows (the largest
number of distinct 4-byte words one of 32 banks serves, summed over 32-row
groups and entry slots); the kept order never costs more than the given
one, and an order that is already conflict-free is kept as it is.
Evaluations with the order on and off must agree exactly.
"""
import numpy as np
import pytest

import oracle

from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G

pytestmark = pytest.mark.gpu

SHIFT = {8: 0, 16: 1, 32: 2, 64: 2}  # word index = (id << SHIFT) >> 2; bank = index mod 32


def padded(adj):
    n, d = adj.shape
    dp = (d + 7) // 8 * 8
    out = np.empty((n, dp), dtype=np.int64)
    out[:, :d] = adj
    out[:, d:] = np.arange(n)[:, None]
    return out


def host_cost(rows, word_bits):
    words = (np.asarray(rows, dtype=np.int64) << SHIFT[word_bits]) >> 2
    n, dp = words.shape
    total = 0
    for g0 in range(0, n, 32):
        blk = words[g0:g0 + 32]
        for s in range(dp):
            distinct = np.unique(blk[:, s])
            total += int(np.bincount(distinct & 31, minlength=32).max())
    return total


@pytest.mark.parametrize("n,d,bits,seed", [
    (65536, 16, 16, 5),
    (8192, 20, 8, 1),     # three chunks per row, byte words
    (10000, 5, 32, 2),    # one chunk with padding, a partial last group
    (9000, 32, 16, 3),    # four chunks
    (8192, 16, 64, 4),    # 64-bit words: keyed by the id
    (3000, 12, 64, 6),    # small graph, padded rows
])
def test_rows_permuted_and_cost_exact(n, d, bits, seed):
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, seed))
    rows, (c_old, c_new) = E.row_order(adj, bits)
    ref = padded(adj)
    assert rows.shape == ref.shape
    assert (np.sort(rows, axis=1) == np.sort(ref, axis=1)).all()
    assert host_cost(ref, bits) == c_old
    assert host_cost(rows, bits) == c_new
    assert c_new <= c_old
    if (n, d) == (65536, 16):
        # random ids put ~3.5 words on a slot's busiest bank, the order ~2.1
        assert c_new < 0.65 * c_old, (c_old, c_new)


def test_conflict_free_order_is_kept():
    E.init(0)
    n, offs = 16384, [1, 2, 3, 5, 8, 13, 21, 34]
    adj = np.array([[(v + o) % n for o in offs] for v in range(n)], dtype=np.int32)
    rows, (c_old, c_new) = E.row_order(adj, 16)
    assert c_old == (n // 32) * 8  # one word per bank in every slot
    assert c_new == c_old
    assert (rows == padded(adj)).all()


def test_rows_wider_than_four_chunks_are_refused():
    E.init(0)
    adj = G.edges_to_adj(1024, 40, G.random_regular(1024, 40, 4))
    with pytest.raises(Exception):
        E.row_order(adj, 16)


@pytest.mark.parametrize("n,d,bits", [
    (65536, 16, 0), (8192, 16, 8), (12288, 16, 32), (10000, 5, 16), (8200, 24, 16), (9000, 32, 32),
])
def test_evaluation_same_with_order_on_and_off(n, d, bits):
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 11))
    E.set_word_bits(bits)
    try:
        prev = E.set_row_order(0)
        off = E.apsp(adj, n, 1)
        E.set_row_order(2)
        on = E.apsp(adj, n, 1)
        E.set_row_order(prev)
    finally:
        E.set_word_bits(0)
    assert (on.diameter, on.sum) == (off.diameter, off.sum)
    assert on.word_bits == off.word_bits


def test_symmetric_groups_same_with_order_on_and_off():
    E.init(0)
    n, d, g = 16384, 12, 4
    adj = G.edges_to_adj(n, d, G.random_symmetric(n, d, g, 7))
    prev = E.set_row_order(0)
    off = E.apsp(adj, n // g, g)
    E.set_row_order(2)
    on = E.apsp(adj, n // g, g)
    E.set_row_order(prev)
    assert (on.diameter, on.sum) == (off.diameter, off.sum)


@pytest.mark.parametrize("n,d,bits", [(1024, 8, 64), (3000, 24, 32), (5000, 16, 0), (4096, 32, 64)])
def test_small_graphs_same_with_order_everywhere(n, d, bits):
    """Mode 2 orders every graph, also below the size threshold of mode 1."""
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 13))
    E.set_word_bits(bits)
    try:
        prev = E.set_row_order(0)
        off = E.apsp(adj, n, 1)
        E.set_row_order(2)
        on = E.apsp(adj, n, 1)
        E.set_row_order(prev)
    finally:
        E.set_word_bits(0)
    assert (on.diameter, on.sum, on.word_bits) == (off.diameter, off.sum, off.word_bits)


@pytest.mark.parametrize("bits", [8, 16, 32, 64])
@pytest.mark.parametrize("n,d", [(2000, 3), (3000, 4), (1500, 4)])
def test_padded_one_chunk_rows_with_and_without_order(n, d, bits):
    """Rows of at most 4 entries carry 4 or more padding entries (the row's
    own id) in their one 16-byte chunk: without the device row order the
    pull levels skip the last 4; with it, padding is permuted into the row
    and every entry is gathered.  Both agree with the oracle."""
    E.init(0)
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 17))
    want = oracle.matrix_op(adj)
    E.set_word_bits(bits)
    try:
        for mode in (0, 2):
            prev = E.set_row_order(mode)
            try:
                got = E.apsp(adj, n, 1)
            finally:
                E.set_row_order(prev)
            assert (got.connected, got.diameter, got.sum) == (want.ok, want.diameter, want.S)
    finally:
        E.set_word_bits(0)


def test_batch_same_with_order_everywhere():
    E.init(0)
    adjs = [G.edges_to_adj(n, d, G.random_regular(n, d, 20 + n)) for n, d in
            [(1500, 8), (2048, 12), (9000, 16), (700, 32), (4000, 24)]]
    prev = E.set_row_order(0)
    off = E.eval_batch([dict(adj=a) for a in adjs])
    E.set_row_order(2)
    on = E.eval_batch([dict(adj=a) for a in adjs])
    E.set_row_order(prev)
    assert [(r.diameter, r.sum) for r in on] == [(r.diameter, r.sum) for r in off]


def test_captured_plan_replays_with_order():
    """The order kernel runs inside the plan's captured graph: every replay,
    with eager evaluations in between, equals the eager result."""
    E.init(0)
    n, d = 16384, 16  # 1024 tiles of 16 sources: ordered by default
    adjs = [G.edges_to_adj(n, d, G.random_regular(n, d, s)) for s in (21, 22)]
    want = [E.apsp(a, symmetric=True) for a in adjs]
    plan = E.Plan([dict(nodes=n, degree=d, based_nodes=n, groups=1, relabel=None, symmetric=True)])
    try:
        for rep in range(4):
            k = rep % 2
            plan.set_adj(0, adjs[k])
            got = plan.run()[0]
            assert (got.diameter, got.sum) == (want[k].diameter, want[k].sum), rep
            eager = E.apsp(adjs[1 - k], symmetric=True)  # eager work between replays
            assert (eager.diameter, eager.sum) == (want[1 - k].diameter, want[1 - k].sum)
    finally:
        plan.close()


def test_relabelled_graph_same_with_order_everywhere():
    E.init(0)
    n, d = 9000, 12
    adj = G.edges_to_adj(n, d, G.random_regular(n, d, 23))
    perm = np.random.default_rng(3).permutation(n).astype(np.int32)
    prev = E.set_row_order(0)
    off = E.apsp(adj, relabel=perm, symmetric=True)
    E.set_row_order(2)
    on = E.apsp(adj, relabel=perm, symmetric=True)
    E.set_row_order(prev)
    assert (on.diameter, on.sum) == (off.diameter, off.sum)
