"""Writes tests/golden/golden_eval.csv from the reference's own evaluator.

Runs in the development container only: it calls the reference's
unmodified evaluation() (general/evaluation.c:375, grid/evaluation.c:192)
compiled by oracle/Makefile into oracle/_ref/.  The CSV holds, per case,
the reference's return value, diameter, ASPL (as a float hex string) and
S = round(ASPL * n * (n-1)); golden_sa.json, the outcome of the
reference's sa() (general/sa.c:413-520) from fixed seeds; and
golden_programs.json, the output files of the reference's whole programs
(main.c included) run on start files this build writes.

    python tests/golden/make_golden.py            # everything
    python tests/golden/make_golden.py --append   # only evaluation cases not in the CSV yet
"""
from __future__ import annotations

import csv
import hashlib
import json
import os
import sys

# The reference's OpenMP paths, loaded into this process next to MPICH, have
# corrupted the heap here when run with several threads; the outputs do not
# depend on the thread count, so the vectors are written single-threaded.
os.environ["OMP_NUM_THREADS"] = "1"

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import oracle  # noqa: E402
from cases import (GOLDEN_CSV, GOLDEN_GRID_SA, GOLDEN_NONREGULAR, GOLDEN_PROGRAMS,  # noqa: E402
                   GOLDEN_SA, build, build_nonregular, general_cases, grid_cases, grid_sa_cases,
                   nonregular_cases, program_cases, program_result_lines, sa_cases,
                   write_program_input)
from graphgolf_amd import graphs as G  # noqa: E402


def grid_sa() -> None:
    """Outcomes of the reference's grid sa() (grid/sa.c:283-361)."""
    rows = []
    for case in grid_sa_cases():
        w, h, d, r, g = case["width"], case["height"], case["degree"], case["length"], case["groups"]
        e0 = G.random_grid_symmetric(w, h, d, r, g, case["graph_seed"])
        e, diam, aspl, steps = oracle.ref_sa_grid(
            e0, w, h, d, r, g, case["seed"], case["ncalcs"],
            cooling_cycle=case.get("cooling_cycle", 1), hill_climbing=case.get("hill", False),
            fixed_temp=case.get("fixed_temp", 0.0))
        rows.append(dict(case, start_sha256=hashlib.sha256(e0.tobytes()).hexdigest(),
                         best_diameter=diam, best_aspl_hex=float(aspl).hex(), steps=steps,
                         edges_sha256=hashlib.sha256(e.tobytes()).hexdigest()))
        print(rows[-1], flush=True)
    with open(GOLDEN_GRID_SA, "w") as f:
        json.dump(rows, f, indent=1)


def nonregular() -> None:
    """Outputs of the non-regular variants' own evaluation()."""
    rows = []
    for case in nonregular_cases():
        adj, lens, kw = build_nonregular(case)
        if case["kind"] == "nr_general":
            r = oracle.ref_nonregular_evaluation(adj, lens, kw["based_nodes"], kw["groups"],
                                                 kw["added_centers"], case["algo"])
        else:
            r = oracle.ref_nonregular_grid_evaluation(adj, lens, kw["groups"], kw["based_nodes"],
                                                      kw["height"], kw["based_height"],
                                                      kw["rotate_hash"], case["bfs"])
        aspl = float(r.aspl).hex() if r.ok else ""
        rows.append((case["id"], int(r.ok), r.diameter, aspl, r.S if r.ok else 0))
        print(rows[-1], flush=True)
    with open(GOLDEN_NONREGULAR, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["id", "ok", "diameter", "aspl_hex", "S"])
        w.writerows(rows)


def programs() -> None:
    """Output files of the reference's whole programs (general/main.c:317-497,
    grid/main.c:586-777), built by oracle/Makefile into oracle/_ref/."""
    import subprocess
    import tempfile
    rows = []
    outputs = {}
    for case in program_cases():
        with tempfile.TemporaryDirectory() as tmp:
            inp, out = os.path.join(tmp, "in.txt"), os.path.join(tmp, "out.txt")
            args = []
            in_sha = ""
            if case["input"] is not None:
                write_program_input(case, inp, outputs)
                with open(inp, "rb") as f:
                    in_sha = hashlib.sha256(f.read()).hexdigest()
                args = ["-f", inp]
            exe = os.path.join(ROOT, "oracle", "_ref", case["kind"] + "_program")
            res = subprocess.run([exe] + args + ["-o", out] + case["args"].split(), cwd=tmp,
                                 capture_output=True, text=True, timeout=600,
                                 env=dict(os.environ, OMP_NUM_THREADS="1"))
            if res.returncode != 0:
                raise SystemExit(f"{case['id']}: {res.stdout[-500:]} {res.stderr[-500:]}")
            with open(out, "rb") as f:
                outputs[case["id"]] = f.read()
                out_sha = hashlib.sha256(outputs[case["id"]]).hexdigest()
            diam_line, steps = program_result_lines(res.stdout)
            rows.append(dict(case, input_sha256=in_sha, output_sha256=out_sha,
                             result=diam_line, steps=steps))
            print(rows[-1], flush=True)
    with open(GOLDEN_PROGRAMS, "w") as f:
        json.dump(rows, f, indent=1)


def main() -> None:
    if not (oracle.ref_available("general") and oracle.ref_available("grid")):
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    if "--grid-sa" in sys.argv:
        grid_sa()
        return
    if "--nonregular" in sys.argv:
        nonregular()
        return
    if "--programs" in sys.argv:
        programs()
        return
    # --append: keep the rows already in the CSV and evaluate only new cases.
    keep = {}
    if "--append" in sys.argv and os.path.exists(GOLDEN_CSV):
        with open(GOLDEN_CSV) as f:
            keep = {r["id"]: (r["id"], r["ok"], r["diameter"], r["aspl_hex"], r["S"])
                    for r in csv.DictReader(f)}
    rows = []
    for case in general_cases():
        if case["id"] in keep:
            rows.append(keep[case["id"]])
            continue
        adj, kw = build(case)
        r = oracle.ref_evaluation(adj, kw["based_nodes"], kw["groups"], kw["added_centers"],
                                  case["algo"])
        rows.append((case["id"], int(r.ok), r.diameter, float(r.aspl).hex(), r.S))
        print(rows[-1], flush=True)
    for case in grid_cases():
        if case["id"] in keep:
            rows.append(keep[case["id"]])
            continue
        adj, kw = build(case)
        r = oracle.ref_grid_evaluation(adj, kw["groups"], kw["based_nodes"], kw["height"],
                                       kw["based_height"], kw["rotate_hash"], case["bfs"])
        aspl = float(r.aspl).hex() if r.ok else ""
        rows.append((case["id"], int(r.ok), r.diameter, aspl, r.S))
        print(rows[-1], flush=True)
    with open(GOLDEN_CSV, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["id", "ok", "diameter", "aspl_hex", "S"])
        w.writerows(rows)
    if "--append" in sys.argv:
        return
    # Simulated-annealing trajectories of the reference's sa() (general/sa.c:413).
    sa_rows = []
    for case in sa_cases():
        e0 = G.random_regular(case["n"], case["d"], case["graph_seed"])
        e, diam, aspl, steps = oracle.ref_sa_general(
            e0, case["n"], case["d"], case["seed"], case["ncalcs"],
            cooling_cycle=case.get("cooling_cycle", 1), hill_climbing=case.get("hill", False))
        sa_rows.append(dict(case, best_diameter=diam, best_aspl_hex=float(aspl).hex(), steps=steps,
                            edges_sha256=hashlib.sha256(e.tobytes()).hexdigest()))
        print(sa_rows[-1], flush=True)
    with open(GOLDEN_SA, "w") as f:
        json.dump(sa_rows, f, indent=1)
    grid_sa()
    nonregular()
    programs()


if __name__ == "__main__":
    main()
