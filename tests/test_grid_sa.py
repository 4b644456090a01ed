"""The grid SA driver against the reference's grid sa() (grid/sa.c:283-361):
same seed and start graph -> same final graph and bit-identical best ASPL,
including 2- and 4-fold symmetric grids (1g moves), hill climbing, cooling
cycles and a fixed temperature.  CPU only: the oracle is the evaluator."""
import hashlib
import json

import numpy as np
import pytest

import oracle
from cases import GOLDEN_GRID_SA, grid_sa_cases
from graphgolf_amd import graphs as G
from graphgolf_amd import sa as SA

with open(GOLDEN_GRID_SA) as _f:
    GOLD = {r["id"]: r for r in json.load(_f)}


def oracle_grid_eval(adj, groups, based_nodes, rotate_hash):
    r = oracle.grid_evaluation(adj, groups, based_nodes, 0, 0, rotate_hash)
    return r.ok, r.diameter, r.S


def run_case(case, evaluator):
    w, h, d, r, g = case["width"], case["height"], case["degree"], case["length"], case["groups"]
    e0 = G.random_grid_symmetric(w, h, d, r, g, case["graph_seed"])
    row = GOLD[case["id"]]
    assert hashlib.sha256(e0.tobytes()).hexdigest() == row["start_sha256"]
    got = SA.sa_grid(e0, w, h, d, r, case["seed"], case["ncalcs"], groups=g,
                     cooling_cycle=case.get("cooling_cycle", 1),
                     hill_climbing=case.get("hill", False),
                     fixed_temp=case.get("fixed_temp", 0.0), evaluator=evaluator)
    assert got.steps == row["steps"]
    assert got.best_diameter == row["best_diameter"]
    assert got.best_aspl == float.fromhex(row["best_aspl_hex"])
    assert hashlib.sha256(got.edges.tobytes()).hexdigest() == row["edges_sha256"]
    return got


@pytest.mark.parametrize("case", grid_sa_cases(), ids=lambda c: c["id"])
def test_grid_sa_matches_reference_trajectory(case):
    run_case(case, oracle_grid_eval)


def test_grid_sa_matches_live_reference():
    if not oracle.ref_available("grid"):
        pytest.skip("reference build (oracle/_ref) not present")
    w, h, d, r, g = 10, 10, 4, 3, 4
    e0 = G.random_grid_symmetric(w, h, d, r, g, 21)
    ref_e, ref_d, ref_a, ref_steps = oracle.ref_sa_grid(e0, w, h, d, r, g, 3, 800)
    got = SA.sa_grid(e0, w, h, d, r, 3, 800, groups=g, evaluator=oracle_grid_eval)
    assert np.array_equal(got.edges, ref_e) and got.best_aspl == ref_a and got.steps == ref_steps


@pytest.mark.parametrize("w,h,d,r,want", [
    (2, 2, 3, 2, (1, 1.0)),       # K4 fits: every pair adjacent
    (3, 1, 2, 1, (2, 8 / 6)),     # the path 0-1-2 is the only graph
    (1, 5, 2, 1, (4, 40 / 20)),   # the path of 5
])
def test_grid_lower_bound_hand_cases(w, h, d, r, want):
    got = G.grid_lower_bound(w, h, d, r)
    assert got[0] == want[0] and got[1] == pytest.approx(want[1], rel=0, abs=1e-15)
    assert got == oracle.grid_lower_bound(w, h, d, r)


@pytest.mark.parametrize("w,h,d,r", [(10, 10, 4, 2), (20, 20, 6, 3), (7, 9, 5, 4)])
def test_grid_lower_bound_below_real_graphs(w, h, d, r):
    lb_d, lb_a = G.grid_lower_bound(w, h, d, r)
    assert (lb_d, lb_a) == oracle.grid_lower_bound(w, h, d, r)
    if (w * h * d) % 2:
        return
    for seed in range(3):
        e = G.random_grid(w, h, d, r, seed)
        o = oracle.matrix_op(G.edges_to_adj(w * h, d, e))
        assert o.ok and o.diameter >= lb_d and o.aspl >= lb_a


@pytest.mark.parametrize("w,h,d,r,g", [(10, 10, 4, 3, 2), (10, 10, 4, 3, 4), (8, 8, 6, 3, 4),
                                       (12, 6, 4, 2, 2)])
def test_random_grid_symmetric_is_valid(w, h, d, r, g):
    n = w * h
    e = G.random_grid_symmetric(w, h, d, r, g, 5)
    assert np.all(np.bincount(e.ravel(), minlength=n) == d)
    x, y = e // h, e % h
    assert np.all(np.abs(x[:, 0] - x[:, 1]) + np.abs(y[:, 0] - y[:, 1]) <= r)
    keys = {(min(a, b), max(a, b)) for a, b in e.tolist()}
    assert len(keys) == len(e) and all(a != b for a, b in keys)  # simple graph
    degs = (180,) if g == 2 else (90, 180, 270)
    for deg in degs:
        rot = {tuple(sorted((G.rotate(a, h, w, g, deg), G.rotate(b, h, w, g, deg)))) for a, b in keys}
        assert rot == keys
    rh = G.rotate_hash(n, h, w, g)
    o = oracle.grid_evaluation(G.edges_to_adj(n, d, e), g, n // g, h, h // 2, rh)
    full = oracle.matrix_op(G.edges_to_adj(n, d, e))
    assert o.ok and (o.diameter, o.S) == (full.diameter, full.S)
