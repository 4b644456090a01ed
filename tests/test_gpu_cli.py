"""The command-line programs on the device evaluator: the same golden cases as
tests/test_cli.py (the reference programs' output files, byte for byte)."""
import pytest

from cases import program_cases
from test_cli import run_program_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", program_cases(), ids=lambda c: c["id"])
def test_program_on_device_matches_reference_program(case, tmp_path):
    run_program_case(case, tmp_path, device=True)
