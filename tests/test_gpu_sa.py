"""The plain SA driver on the device evaluator: every evaluation of a run is
checked against the CPU oracle, and full C1 runs (n=256, d=8, 10k steps)
reproduce the reference's sa() trajectory."""
import hashlib
import json

import numpy as np
import pytest

import oracle
from cases import GOLDEN_SA, sa_cases
from graphgolf_amd import evaluation as E
from graphgolf_amd import graphs as G
from graphgolf_amd import sa as SA

pytestmark = pytest.mark.gpu

with open(GOLDEN_SA) as _f:
    GOLD = {r["id"]: r for r in json.load(_f)}


@pytest.mark.parametrize("case", sa_cases(), ids=lambda c: c["id"])
def test_gpu_sa_matches_reference_trajectory(case):
    e0 = G.random_regular(case["n"], case["d"], case["graph_seed"])
    got = SA.sa_general(e0, case["n"], case["d"], case["seed"], case["ncalcs"],
                        cooling_cycle=case.get("cooling_cycle", 1),
                        hill_climbing=case.get("hill", False))
    row = GOLD[case["id"]]
    assert got.steps == row["steps"]
    assert got.best_diameter == row["best_diameter"]
    assert got.best_aspl == float.fromhex(row["best_aspl_hex"])
    assert hashlib.sha256(got.edges.tobytes()).hexdigest() == row["edges_sha256"]


@pytest.mark.parametrize("case", [c for c in sa_cases() if c["d"] <= 16], ids=lambda c: c["id"])
def test_gpu_sa_incremental_matches_reference_trajectory(case):
    """The same runs with every proposal scored by the incremental scorer."""
    e0 = G.random_regular(case["n"], case["d"], case["graph_seed"])
    got = SA.sa_general(e0, case["n"], case["d"], case["seed"], case["ncalcs"],
                        cooling_cycle=case.get("cooling_cycle", 1),
                        hill_climbing=case.get("hill", False), incremental=True)
    row = GOLD[case["id"]]
    assert got.steps == row["steps"]
    assert got.best_diameter == row["best_diameter"]
    assert got.best_aspl == float.fromhex(row["best_aspl_hex"])
    assert hashlib.sha256(got.edges.tobytes()).hexdigest() == row["edges_sha256"]


@pytest.mark.parametrize("n,d,steps", [(4096, 16, 300), (65536, 16, 60)])
def test_gpu_sa_incremental_same_log_as_full(n, d, steps):
    """Every evaluation of the run, (diameter, S), is the same with and
    without the incremental scorer, at the C2 and C5 sizes."""
    print(f"sa incremental vs full n={n}", flush=True)
    e0 = G.random_regular(n, d, 11)
    full = SA.sa_general(e0, n, d, 5, steps, max_temp=50.0, min_temp=1.0, log=4 * steps)
    inc = SA.sa_general(e0, n, d, 5, steps, max_temp=50.0, min_temp=1.0, log=4 * steps,
                        incremental=True)
    assert full.evaluations == inc.evaluations
    assert np.array_equal(full.log_diameter, inc.log_diameter)
    assert np.array_equal(full.log_sum, inc.log_sum)
    assert np.array_equal(full.edges, inc.edges)
    assert full.accepts == inc.accepts and full.accepts > 0


def test_gpu_sa_every_evaluation_checked_against_oracle():
    """C1 shape: each proposal is evaluated on the GPU and by the oracle."""
    n, d = 256, 8
    checked = [0]

    def both(adj):
        g = E.apsp(adj, symmetric=True)
        o = oracle.matrix_op(adj)
        assert g.connected == o.ok
        if o.ok:
            assert (g.diameter, g.sum) == (o.diameter, o.S)
        checked[0] += 1
        return g.connected, g.diameter, g.sum

    got = SA.sa_general(G.random_regular(n, d, 1), n, d, 7, 2000, evaluator=both)
    assert checked[0] == got.evaluations >= 2001


@pytest.mark.parametrize("case_id", ["gsa:c4", "gsa:w10h10d4r3g2", "gsa:w10h10d4r3g4",
                                     "gsa:w8h8d6r3g4:hill"])
def test_gpu_grid_sa_matches_reference_trajectory(case_id):
    """The grid driver on the device evaluator (rotate_hash relabelling in the
    prep kernel) visits the reference grid sa()'s graphs."""
    from cases import GOLDEN_GRID_SA, grid_sa_cases

    with open(GOLDEN_GRID_SA) as f:
        row = {r["id"]: r for r in json.load(f)}[case_id]
    case = {c["id"]: c for c in grid_sa_cases()}[case_id]
    w, h, d, r, g = case["width"], case["height"], case["degree"], case["length"], case["groups"]
    e0 = G.random_grid_symmetric(w, h, d, r, g, case["graph_seed"])
    got = SA.sa_grid(e0, w, h, d, r, case["seed"], case["ncalcs"], groups=g,
                     cooling_cycle=case.get("cooling_cycle", 1),
                     hill_climbing=case.get("hill", False), fixed_temp=case.get("fixed_temp", 0.0))
    assert got.steps == row["steps"]
    assert got.best_diameter == row["best_diameter"]
    assert got.best_aspl == float.fromhex(row["best_aspl_hex"])
    assert hashlib.sha256(got.edges.tobytes()).hexdigest() == row["edges_sha256"]
