"""The drop-in from the reference's side: the reference's own general and
grid programs (main.c, sa.c, common.c, timer.c, unmodified) linked against
libgg_general.so / libgg_grid.so in place of their evaluation.c, as
oracle/Makefile builds them (oracle/_ref/*_program_gg; the gg_cuda precedent,
general/Makefile:51-52).  Every golden program case must give the reference
program's output file byte for byte, its result line and step count, and a
non-zero APSP time: the shim's TIMER_APSP brackets reached the reference's
timer.c (general/timer.c:23-38)."""
import hashlib
import json
import os
import re
import subprocess

import pytest

from cases import GOLDEN_PROGRAMS, program_cases, program_result_lines, write_program_input

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
with open(GOLDEN_PROGRAMS) as _f:
    GOLD = {r["id"]: r for r in json.load(_f)}
CASES = {c["id"]: c for c in program_cases()}


def _exe(kind):
    path = os.path.join(ROOT, "oracle", "_ref", f"{kind}_program_gg")
    if not os.path.exists(path):
        pytest.skip("the drop-in programs are built with the reference sources (make -C oracle)")
    return path


def run_dropin(case, tmp_path):
    """Runs one golden program case through the reference program linked to
    the shim; returns (output bytes, stdout)."""
    outputs = {}
    spec = case["input"]
    if spec is not None and spec[0] == "output_of":
        prev = CASES[spec[1]]
        d = tmp_path / ("prev_" + prev["id"])
        d.mkdir()
        outputs[prev["id"]], _ = run_dropin(prev, d)
    args = []
    if spec is not None:
        inp = str(tmp_path / "in.txt")
        write_program_input(case, inp, outputs)
        args = ["-f", inp]
    out = str(tmp_path / "out.txt")
    print(f"drop-in {case['kind']} {case['id']}: {' '.join(args + case['args'].split())}", flush=True)
    res = subprocess.run([_exe(case["kind"])] + args + ["-o", out] + case["args"].split(),
                         cwd=str(tmp_path), capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert res.returncode == 0, (res.stdout[-1500:], res.stderr[-1500:])
    with open(out, "rb") as f:
        return f.read(), res.stdout


@pytest.mark.parametrize("case", program_cases(), ids=lambda c: c["id"])
def test_reference_program_on_the_shim(case, tmp_path):
    data, stdout = run_dropin(case, tmp_path)
    row = GOLD[case["id"]]
    assert hashlib.sha256(data).hexdigest() == row["output_sha256"]
    diam_line, steps = program_result_lines(stdout)
    assert diam_line == row["result"]
    assert steps == row["steps"]
    if "-D" in case["args"].split() or "-d" in case["args"].split():
        return  # temperature detection ends before the time breakdown
    m = re.search(r"\(APSP: ([0-9.]+) sec\.", stdout)
    assert m, stdout[-800:]
    assert float(m.group(1)) > 0.0, "TIMER_APSP never accumulated: the shim's brackets missed timer.c"
