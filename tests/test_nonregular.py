"""The non-regular variants' evaluation() (variants/not_regular_general_graph,
variants/not_regular_grid_graph): golden vectors written by the variants' own
evaluation() (tests/golden/make_golden.py --nonregular) pin the oracle's
restatement (row v read up to its degree = the row padded with v itself);
CPU only."""
import csv

import numpy as np
import pytest

import oracle
from cases import GOLDEN_NONREGULAR, build_nonregular, nonregular_cases
from graphgolf_amd import _lib as L

with open(GOLDEN_NONREGULAR) as _f:
    GOLD = {r["id"]: r for r in csv.DictReader(_f)}


def oracle_eval(case):
    adj, lens, kw = build_nonregular(case)
    padded = oracle.nonregular_padded(adj, lens)
    if case["kind"] == "nr_general":
        return oracle.matrix_op(padded, kw["based_nodes"], kw["groups"], kw["added_centers"])
    return oracle.grid_evaluation(padded, kw["groups"], kw["based_nodes"], kw["height"],
                                  kw["based_height"], kw["rotate_hash"], case["bfs"])


@pytest.mark.parametrize("case", nonregular_cases(), ids=lambda c: c["id"])
def test_oracle_matches_nonregular_reference(case):
    row = GOLD[case["id"]]
    r = oracle_eval(case)
    assert int(r.ok) == int(row["ok"])
    if case["kind"] == "nr_general" and not r.ok:
        return  # INT_MAX / DBL_MAX in the reference
    assert r.diameter == int(row["diameter"])
    if r.ok:
        assert r.S == int(row["S"])
        assert r.aspl == float.fromhex(row["aspl_hex"])


def test_unused_slots_are_ignored():
    case = nonregular_cases()[7]
    adj, lens, kw = build_nonregular(case)
    a = oracle.nonregular_padded(adj, lens)
    other = adj.copy()
    other[np.arange(adj.shape[1])[None, :] >= lens[:, None]] = 0  # different garbage
    assert np.array_equal(a, oracle.nonregular_padded(other, lens))


def test_live_nonregular_reference_agrees():
    if not (oracle.ref_available("nrgeneral") and oracle.ref_available("nrgrid")):
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(3)
    for case in nonregular_cases()[::5]:
        adj, lens, kw = build_nonregular(case)
        if case["kind"] == "nr_general":
            ref = oracle.ref_nonregular_evaluation(adj, lens, kw["based_nodes"], kw["groups"],
                                                   kw["added_centers"], case["algo"])
        else:
            ref = oracle.ref_nonregular_grid_evaluation(adj, lens, kw["groups"], kw["based_nodes"],
                                                        kw["height"], kw["based_height"],
                                                        kw["rotate_hash"], case["bfs"])
        o = oracle_eval(case)
        assert ref.ok == o.ok and (not o.ok or (ref.diameter, ref.S) == (o.diameter, o.S))
    del rng


def test_nonregular_shims_export_the_variant_abi():
    import ctypes

    for path in (L.NONREGULAR_GENERAL_PATH, L.NONREGULAR_GRID_PATH):
        assert "evaluation" in L.exported_symbols(path)
        ctypes.CDLL(path)
