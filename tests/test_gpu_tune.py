"""The autotuner and the persistent batch launch on the GPU: every tuned
configuration gives the same (connected, diameter, S) as the oracle, the
tuning table round-trips through its file, plans and evaluations made after
tuning (or after loading a table) run the tuned configuration, and the
uneven 512-graph batch runs as one queue-driven launch."""
import csv

import numpy as np
import pytest

import oracle
from cases import GOLDEN_CSV, build, general_cases
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G
from graphgolf_amd import workloads as WL

pytestmark = pytest.mark.gpu

with open(GOLDEN_CSV) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}


@pytest.fixture(autouse=True)
def _clean_table():
    E.tune_table_clear()
    E.set_word_bits(0)
    yield
    E.tune_table_clear()
    E.set_word_bits(0)


def _batch(count=24, seed=7):
    adjs = WL.c3_batch(count, seed)
    return [dict(adj=a, based_nodes=a.shape[0], groups=1, symmetric=True) for a in adjs]


def _plan_results(items):
    plan = E.Plan([dict(nodes=it["adj"].shape[0], degree=it["adj"].shape[1],
                        based_nodes=it["based_nodes"], groups=it["groups"],
                        symmetric=it["symmetric"]) for it in items])
    for i, it in enumerate(items):
        plan.set_adj(i, it["adj"])
    res = plan.run()
    cfg = plan.config()
    plan.close()
    return res, cfg


def test_autotune_single_records_bucket_and_round_trips(tmp_path):
    bits, ms = E.autotune(4096, 16, 1024, 4, 1, 3)
    assert bits in (8, 16, 32, 64) and ms > 0
    cfg = E.tune_lookup(4096, 16, 1024, 4)
    assert cfg["word_bits"] == bits and cfg["persistent"] in (0, 1)
    # the same bucket covers other sizes of ceil(log2 n) = 12
    assert E.tune_lookup(3000, 16, 750, 4)["word_bits"] == min(bits, 64)
    path = str(tmp_path / "tune.txt")
    E.tune_table_save(path)
    rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
    assert len(rows) == 1 and rows[0][:5] == ["12", "16", "4", "0", "0"] and int(rows[0][5]) == bits
    E.tune_table_clear()
    assert E.tune_table_load(path) == 1
    assert E.tune_lookup(4096, 16, 1024, 4) == cfg
    # the tuned choice is what an evaluation of that shape runs, and its
    # result is the golden one
    case = next(c for c in general_cases() if c["id"] == "c2:11")
    adj, kw = build(case)
    r = E.apsp(adj, kw["based_nodes"], kw["groups"])
    assert r.word_bits == bits and r.sum == int(GOLD["c2:11"]["S"])


@pytest.mark.parametrize("bits,block,order,pers", [(64, 0, 0, 1), (64, 0, 2, 0), (32, 0, 0, 1),
                                                   (16, 0, 2, 1), (8, 0, 0, 0)])
def test_loaded_configurations_agree_with_the_oracle(tmp_path, bits, block, order, pers):
    """A batch of uneven graphs under several configurations loaded from a
    table file: the plan reports the configuration and every graph matches
    the oracle."""
    items = _batch(24)
    big = max(items, key=lambda it: it["adj"].shape[0])["adj"]
    nb = int(np.ceil(np.log2(big.shape[0])))
    path = tmp_path / "t.txt"
    path.write_text(f"{nb} {big.shape[1]} 1 0 1 {bits} {block} {order} {pers}\n")
    assert E.tune_table_load(str(path)) == 1
    res, cfg = _plan_results(items)
    assert cfg["persistent"] == pers and cfg["row_order"] == order
    assert cfg["word_bits"] == bits
    for it, r in zip(items, res):
        w = oracle.matrix_op(it["adj"])
        assert (r.connected, r.diameter, r.sum) == (w.ok, w.diameter, w.S)


def test_c3_batch_tuned_and_persistent():
    """The uneven 512-graph batch: tuned on itself, then run by a plan with
    the tuned configuration; all 512 results equal single evaluations, and
    8 of them the oracle."""
    items = _batch(512)
    cfg, ms = E.autotune_batch(items, reps=2)
    assert ms > 0 and cfg["word_bits"] in (8, 16, 32, 64)
    res, pcfg = _plan_results(items)
    assert pcfg["persistent"] == cfg["persistent"] and pcfg["word_bits"] == cfg["word_bits"]
    single = E.eval_batch(items)
    for r, s in zip(res, single):
        assert (r.connected, r.diameter, r.sum) == (s.connected, s.diameter, s.sum)
    for k in range(0, 512, 64):
        w = oracle.matrix_op(items[k]["adj"])
        assert (res[k].diameter, res[k].sum) == (w.diameter, w.S)


def test_default_batch_is_one_persistent_launch():
    items = _batch(8)
    res, cfg = _plan_results(items)
    assert cfg["persistent"] == 1
    for it, r in zip(items, res):
        w = oracle.matrix_op(it["adj"])
        assert (r.diameter, r.sum) == (w.diameter, w.S)


def test_tuner_covers_symmetric_and_centre_shapes():
    """Shapes with groups and centres: the tuner's candidates all agree
    (it fails otherwise) and the recorded choice gives the golden S."""
    for cid in ("ctr:g4", "sym:g5"):
        case = next(c for c in general_cases() if c["id"] == cid)
        adj, kw = build(case)
        n, d = adj.shape
        E.autotune(n, d, kw["based_nodes"], kw["groups"], 1, 2) if kw["added_centers"] == 0 else \
            E.autotune_batch([dict(adj=adj, based_nodes=kw["based_nodes"], groups=kw["groups"],
                                   added_centers=kw["added_centers"])], reps=2)
        r = E.apsp(adj, kw["based_nodes"], kw["groups"], kw["added_centers"])
        assert r.sum == int(GOLD[cid]["S"])
