"""The command-line programs on the device evaluator: the same golden cases as
tests/test_cli.py (the reference programs' output files, byte for byte)."""
import pytest

from cases import program_cases
from test_cli import run_program_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", program_cases(), ids=lambda c: c["id"])
def test_program_on_device_matches_reference_program(case, tmp_path):
    run_program_case(case, tmp_path, device=True)


@pytest.mark.parametrize("case", [c for c in program_cases() if c["kind"] == "general"],
                         ids=lambda c: c["id"])
def test_general_program_incremental_matches_reference_program(case, tmp_path):
    """The same runs with every proposal scored by the incremental scorer."""
    run_program_case(case, tmp_path, device=True, incremental=True)



@pytest.mark.parametrize("cid", ["gen_n64d4", "gen_g2_v2", "grid_lat_6x6_g2"])
def test_program_in_a_fresh_process_matches_reference_program(cid, tmp_path):
    """`python -m graphgolf_amd ...` as a user runs it: a fresh process whose
    first GPU use is inside the run (the device starts before the program
    seeds libc random(), so its start-up cannot shift the draws)."""
    import hashlib
    import os
    import subprocess
    import sys

    from cases import write_program_input
    from test_cli import GOLD

    case = next(c for c in program_cases() if c["id"] == cid)
    args = []
    if case["input"] is not None:
        inp = str(tmp_path / "in.txt")
        write_program_input(case, inp)
        args = ["-f", inp]
    out = str(tmp_path / "out.txt")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    res = subprocess.run([sys.executable, "-m", "graphgolf_amd", case["kind"]] + args +
                         ["-o", out] + case["args"].split(), cwd=root, capture_output=True,
                         text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    with open(out, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == GOLD[cid]["output_sha256"]
