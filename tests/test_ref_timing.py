"""The benchmark's reference CPU-baseline leg (``oracle/ref_timing.py``).

The child process times the reference's own ``evaluation()``; these checks
pin its algorithm choice to the reference program's rule and its result to
the oracle on a small graph.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from oracle import ref_timing
from graphgolf_amd import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_algo_choice_follows_the_reference_program():
    # general/main.c:414-416: n * ceil(n/64) * 8 bytes against 2 GiB
    assert ref_timing.algo_for(65536, 65536, 1) == ref_timing.MATRIX_OP
    assert ref_timing.algo_for(131072, 131072, 1) == ref_timing.MATRIX_OP  # exactly 2 GiB
    assert ref_timing.algo_for(131073, 131073, 1) == ref_timing.MATRIX_OP_MEM_SAVING
    assert ref_timing.algo_for(262144, 65536, 4) == ref_timing.MATRIX_OP  # 2 GiB with groups
    assert ref_timing.algo_for(524288, 65536, 8) == ref_timing.MATRIX_OP_MEM_SAVING


@pytest.mark.skipif(not oracle.ref_available("general"), reason="reference build absent")
def test_child_reports_the_oracle_result(tmp_path):
    adj = G.edges_to_adj(256, 8, G.random_regular(256, 8, 3))
    path = str(tmp_path / "adj.npy")
    np.save(path, adj.astype(np.int32))
    p = subprocess.run([sys.executable, "-m", "oracle.ref_timing", path, "0.05"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    r = json.loads([ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1])
    want = oracle.matrix_op(adj)
    assert r["ok"] and (r["diameter"], r["S"]) == (want.raw_diameter, want.S)
    assert r["algo"] == ref_timing.MATRIX_OP and r["threads"] >= 1 and len(r["times"]) >= 1
