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

