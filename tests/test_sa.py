"""The plain SA driver against the reference's sa() (general/sa.c:413-520):
same seed and start graph -> same final graph, bit-identical best ASPL.
CPU only: the driver runs with the oracle plugged in as its evaluator."""
import hashlib
import json

import numpy as np
import pytest

import oracle
from cases import GOLDEN_SA, sa_cases
from graphgolf_amd import graphs as G
from graphgolf_amd import sa as SA

with open(GOLDEN_SA) as _f:
    GOLD = {r["id"]: r for r in json.load(_f)}


def oracle_eval(adj):
    r = oracle.matrix_op(adj)
    return r.ok, r.raw_diameter, r.S


@pytest.mark.parametrize("case", sa_cases(), ids=lambda c: c["id"])
def test_sa_driver_matches_reference_trajectory(case):
    e0 = G.random_regular(case["n"], case["d"], case["graph_seed"])
    got = SA.sa_general(e0, case["n"], case["d"], case["seed"], case["ncalcs"],
                        cooling_cycle=case.get("cooling_cycle", 1),
                        hill_climbing=case.get("hill", False), evaluator=oracle_eval)
    row = GOLD[case["id"]]
    assert got.steps == row["steps"]
    assert got.best_diameter == row["best_diameter"]
    assert got.best_aspl == float.fromhex(row["best_aspl_hex"])
    assert hashlib.sha256(got.edges.tobytes()).hexdigest() == row["edges_sha256"]


def test_sa_driver_matches_live_reference():
    if not oracle.ref_available("general"):
        pytest.skip("reference build (oracle/_ref) not present")
    n, d = 96, 4
    e0 = G.random_regular(n, d, 21)
    ref_e, ref_d, ref_a, ref_steps = oracle.ref_sa_general(e0, n, d, 3, 1500)
    got = SA.sa_general(e0, n, d, 3, 1500, evaluator=oracle_eval)
    assert np.array_equal(got.edges, ref_e) and got.best_aspl == ref_a and got.steps == ref_steps


def test_sa_result_is_a_valid_graph():
    n, d = 128, 6
    e0 = G.random_regular(n, d, 8)
    got = SA.sa_general(e0, n, d, 4, 500, evaluator=oracle_eval, log=2000)
    deg = np.bincount(got.edges.ravel(), minlength=n)
    assert np.all(deg == d)
    r = oracle.matrix_op(G.edges_to_adj(n, d, got.edges))
    assert r.ok and r.aspl == got.best_aspl and r.diameter == got.best_diameter
    # the log holds one (diameter, S) per evaluation, starting with the initial graph
    assert len(got.log_sum) == got.evaluations
    r0 = oracle.matrix_op(G.edges_to_adj(n, d, e0))
    assert got.log_sum[0] == r0.S and got.log_diameter[0] == r0.diameter


def test_sa_evaluator_error_propagates():
    def bad(adj):
        raise RuntimeError("boom")
    with pytest.raises(Exception):
        SA.sa_general(G.random_regular(64, 4, 1), 64, 4, 1, 10, evaluator=bad)
