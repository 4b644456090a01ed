"""The command-line programs (graphgolf_amd/cli.py) against the reference's
whole programs (general/main.c, grid/main.c): from the same start file, seed
and flags, the same output file byte for byte, the same result line and step
count -- covering the pre-run moves, -M resume, -H/-Y, -w/-c, -C, -F, the
reference's own lattice (no -f) and symmetric copies of a base file (-g).
CPU only: the oracle is the evaluator; tests/test_gpu_cli.py runs the same
cases on the device evaluator."""
import hashlib
import io
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from cases import GOLDEN_PROGRAMS, program_cases, program_result_lines, write_program_input
from graphgolf_amd import cli
from graphgolf_amd import graphs as G

with open(GOLDEN_PROGRAMS) as _f:
    GOLD = {r["id"]: r for r in json.load(_f)}


def oracle_eval(adj, based_nodes=None, groups=1, added_centers=0):
    r = oracle.evaluation(adj, based_nodes, groups, added_centers)
    return r.ok, r.diameter, r.S


def oracle_grid_eval(adj, groups, based_nodes, rotate_hash):
    r = oracle.grid_evaluation(adj, groups, based_nodes, 0, 0, rotate_hash)
    return r.ok, r.diameter, r.S


CASES = {c["id"]: c for c in program_cases()}


def earlier_outputs(case, tmp_path, device):
    """The output files a case's input is taken from, produced by running
    the earlier cases through the CLI (each checked against its golden)."""
    outs = {}
    spec = case["input"]
    if spec is not None and spec[0] == "output_of":
        prev = CASES[spec[1]]
        d = tmp_path / ("prev_" + prev["id"])
        d.mkdir()
        outs[prev["id"]] = run_program_case(prev, d, device)
    return outs


def run_program_case(case, tmp_path, device: bool, incremental=None):
    """Runs one golden program case through the CLI; checks it against the
    reference program's recorded output.  Returns the output file's bytes."""
    row = GOLD[case["id"]]
    args = []
    if case["input"] is not None:
        inp = str(tmp_path / "in.txt")
        write_program_input(case, inp, earlier_outputs(case, tmp_path, device))
        with open(inp, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == row["input_sha256"]
        args = ["-f", inp]
    out = str(tmp_path / "out.txt")
    args += ["-o", out] + case["args"].split()
    buf = io.StringIO()
    if case["kind"] == "general":
        cli.run_general(args, evaluator=None if device else oracle_eval, out=buf,
                        incremental=incremental)
    else:
        cli.run_grid(args, evaluator=None if device else oracle_grid_eval, out=buf)
    with open(out, "rb") as f:
        data = f.read()
    assert hashlib.sha256(data).hexdigest() == row["output_sha256"]
    diam_line, steps = program_result_lines(buf.getvalue())
    assert diam_line == row["result"]
    assert steps == row["steps"]
    return data


@pytest.mark.parametrize("case", program_cases(), ids=lambda c: c["id"])
def test_program_matches_reference_program(case, tmp_path):
    run_program_case(case, tmp_path, device=False)


def test_lattice_matches_reference_program(tmp_path):
    """gg_grid_lattice is create_lattice: the reference program run for zero
    annealing steps is not possible (-n >= 1), so compare through a -Y run of
    one step whose start graph is the lattice: the step count and the output
    agree only if the lattice and the stream position agree."""
    exe = os.path.join(oracle.__path__[0], "_ref", "grid_program")
    if not os.path.exists(exe):
        pytest.skip("reference build (oracle/_ref) not present")
    for w, h, d, r, s in [(9, 7, 4, 2, 1), (12, 12, 6, 3, 2), (11, 6, 3, 2, 5)]:
        ref_out, my_out = str(tmp_path / f"r{s}.txt"), str(tmp_path / f"m{s}.txt")
        flags = f"-W {w} -H {h} -D {d} -R {r} -s {s} -n 30".split()
        subprocess.run([exe] + flags + ["-o", ref_out], cwd=tmp_path, capture_output=True,
                       check=True, timeout=300, env=dict(os.environ, OMP_NUM_THREADS="1"))
        cli.run_grid(flags + ["-o", my_out], evaluator=oracle_grid_eval, out=io.StringIO())
        with open(ref_out) as a, open(my_out) as b:
            assert a.read() == b.read(), (w, h, d, r, s)


def test_general_program_matches_live_reference(tmp_path):
    exe = os.path.join(oracle.__path__[0], "_ref", "general_program")
    if not os.path.exists(exe):
        pytest.skip("reference build (oracle/_ref) not present")
    inp = str(tmp_path / "in.txt")
    G.write_edges(inp, G.random_regular(128, 5, 41))
    for flags in ("-s 17 -n 500", "-M -s 4 -n 500 -w 10 -c 1", "-g 2 -s 5 -n 300",
                  "-g 4 -H -s 2 -n 300", "-g 3 -s 8 -n 300 -w 20 -c 2"):
        ref_out, my_out = str(tmp_path / "r.txt"), str(tmp_path / "m.txt")
        for p in (ref_out, my_out):
            if os.path.exists(p):
                os.remove(p)
        res = subprocess.run([exe, "-f", inp, "-o", ref_out] + flags.split(), cwd=tmp_path,
                             capture_output=True, text=True, check=True, timeout=300,
                             env=dict(os.environ, OMP_NUM_THREADS="1"))
        buf = io.StringIO()
        cli.run_general(["-f", inp, "-o", my_out] + flags.split(), evaluator=oracle_eval, out=buf)
        with open(ref_out) as a, open(my_out) as b:
            assert a.read() == b.read(), flags
        assert program_result_lines(buf.getvalue()) == program_result_lines(res.stdout)


def test_grid_lattice_properties():
    """The lattice is d-regular (a self-loop counts twice) with every edge
    within the length and no repeated edge.  Self-loops can remain: the
    reference's last pass, over repeated edges, keeps any swap that does not
    add a repeat, loops included (grid/main.c:305-325)."""
    for w, h, d, r, s in [(6, 6, 4, 3, 5), (7, 5, 4, 2, 3), (5, 6, 3, 2, 2), (16, 16, 6, 2, 9)]:
        e = G.grid_lattice(w, h, d, r, s)
        n = w * h
        assert np.array_equal(np.bincount(e.ravel(), minlength=n), np.full(n, d))
        dist = np.abs(e[:, 0] // h - e[:, 1] // h) + np.abs(e[:, 0] % h - e[:, 1] % h)
        assert dist.max() <= r
        assert len(np.unique(np.sort(e, axis=1), axis=0)) == len(e)


# ---- Argument handling: the reference's checks and messages ----------------

@pytest.mark.parametrize("argv,msg", [
    (["-f", "x", "-H", "-w", "5"], "Both -H and -w cannot be used."),
    (["-f", "x", "-H", "-c", "5"], "Both -H and -c cannot be used."),
    (["-f", "x", "-s", "-1"], "-s value >= 0"),
    (["-f", "x", "-n", "0"], "-n value >= 1"),
    (["-f", "x", "-w", "0"], "-w value > 0"),
    (["-f", "x", "-g", "2", "-D"], "When using -D option, -g must be 1."),
    (["-f", "x", "-g", "0"], "-g value >= 1"),
    (["-f", "x", "-v", "0"], "-v value > 0"),
    (["-f", "/nonexistent/file"], "File not found"),
])
def test_general_argument_errors(argv, msg):
    with pytest.raises(cli.CliError, match=msg.replace("(", r"\(").replace(")", r"\)")
                       .replace("+", r"\+")):
        cli.run_general(argv, evaluator=oracle_eval, out=io.StringIO())


@pytest.mark.parametrize("argv,msg", [
    (["-W", "4", "-H", "4", "-D", "3"], "Must need -R"),
    (["-R", "2"], 'Must set -f or "-W and -H and -D"'),
    (["-W", "4", "-H", "4", "-D", "3", "-R", "2", "-M"], "Must set both -M and -f"),
    (["-W", "4", "-H", "4", "-D", "3", "-R", "2", "-Y", "-w", "3"], "Both -Y and -w"),
    (["-W", "4", "-H", "4", "-D", "3", "-R", "2", "-w", "3", "-c", "3"], "The same values"),
    (["-W", "4", "-H", "4", "-D", "3", "-R", "2", "-g", "3"], "-g value == 1 or 2 or 4"),
    (["-W", "6", "-H", "4", "-D", "4", "-R", "2", "-g", "4"], "When g = 4, width"),
    (["-W", "4", "-H", "5", "-D", "4", "-R", "2", "-g", "2"], "When g = 2, height"),
    (["-W", "4", "-H", "4", "-D", "3", "-R", "2", "-d", "-Y"], "Both -Y and -d"),
])
def test_grid_argument_errors(argv, msg):
    with pytest.raises(cli.CliError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        cli.run_grid(argv, evaluator=oracle_grid_eval, out=io.StringIO())


def test_existing_output_file_is_refused(tmp_path):
    inp, out = str(tmp_path / "in.txt"), str(tmp_path / "out.txt")
    G.write_edges(inp, G.random_regular(32, 3, 1))
    open(out, "w").close()
    with pytest.raises(cli.CliError, match="exsits"):
        cli.run_general(["-f", inp, "-o", out], evaluator=oracle_eval, out=io.StringIO())


def test_unregular_input_is_refused(tmp_path):
    inp = str(tmp_path / "in.txt")
    e = G.random_regular(32, 3, 1).copy()
    e[0, 1] = e[0, 0]  # a self-loop: the degrees stay even, the check names it
    G.write_edges(inp, e)
    with pytest.raises(cli.CliError, match="NG"):
        cli.run_general(["-f", inp, "-n", "5"], evaluator=oracle_eval, out=io.StringIO())


def test_disconnected_grid_input_is_refused(tmp_path):
    # two 2x2 blocks with no edge between them: 4-regular multigraph, r = 1
    inp = str(tmp_path / "in.txt")
    e = []
    for base in (0, 4):  # a 2x4 grid, columns {0,1} and {2,3}
        a, b, c, d = base, base + 1, base + 2, base + 3
        e += [(a, b), (a, c), (b, d), (c, d)] * 2
    G.write_grid_edges(inp, np.array(e, dtype=np.int32), 2)
    with pytest.raises(Exception, match="never reached"):
        cli.run_grid(["-f", inp, "-M", "-R", "1", "-n", "5"], evaluator=oracle_grid_eval,
                     out=io.StringIO())


def test_main_reports_errors_with_status_1(capsys):
    assert cli.main(["general", "-f", "/nonexistent/file"]) == 1
    assert "File not found" in capsys.readouterr().err
    assert cli.main([]) == 2
