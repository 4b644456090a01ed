"""The benchmark's own arithmetic and rules (bench.py), on the CPU.

The canonical algorithmic bytes follow the single-tile formulation of the
reference's bit-parallel BFS (L levels, each reading the adjacency table of
n*d int32 and the reach words of n vertices and their d neighbours plus two
more words, w = ceil(B/64) words of 8 bytes); the C5 figure is pinned to
the value the measurement notes quote.  The roofline block divides the
committed rocprofv3 LDS bytes by the kernel time, and its fraction must be
a fraction of the peak.  A run asked for more than one GPU stops before any
device is touched.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_canonical_bytes_of_c5():
    # n = B = 65536, d = 16, diameter 6: w = 1024 words per vertex.
    assert bench.algorithmic_bytes(65536, 16, 65536, 6) == 58_007_224_320
    # The formula term by term on a small case: 3 * 10 * (4*4 + 8*1*(4+2)).
    assert bench.algorithmic_bytes(10, 4, 10, 3) == 3 * 10 * (16 + 48)
    # w rounds B up to whole 64-bit words.
    assert bench.algorithmic_bytes(100, 4, 65, 1) == 100 * (16 + 8 * 2 * 6)


@pytest.mark.parametrize("workload", ["c5", "c3", "c4b"])
def test_roofline_from_committed_profiles(workload):
    lds = bench.read_profile_value(f"lds_{workload}.md", "lds_bytes_per_launch")
    busy = bench.read_profile_value(f"lds_{workload}.md", "lds_busy_frac")
    traffic = bench.read_traffic(workload)
    assert lds is not None and lds > 0 and 0.0 < busy < 1.0 and traffic > 0
    kernel_ms = {"c5": 1.2332, "c3": 11.41, "c4b": 91.88}[workload]
    r = bench.lds_roofline(workload, kernel_ms, traffic, 1.0, 16 if workload == "c5" else 64,
                           "layout" if workload == "c5" else False)
    assert r["bound"] == "lds" and r["unit"] == "GB/s"
    assert r["achieved"] == pytest.approx(lds / (kernel_ms * 1e-3) / 1e9, rel=1e-4)
    assert r["frac"] == pytest.approx(r["achieved"] / r["peak"], abs=1e-4)
    assert 0.0 < r["frac"] < 1.0 and 0.0 < r["gather_ceiling_frac"] < 1.0
    # The peak is 256 CUs x 128 B/clk x 2.4 GHz.
    assert r["peak"] == pytest.approx(256 * 128 * 2.4, rel=1e-9)


def test_gather_ceilings():
    # 16-bit words on 32 distinct banks: 64 lanes x 2 B per 2.89 cycles per CU.
    assert bench.gather_ceiling_GBps(16, "layout") == pytest.approx(256 * 2.4 * 128 / 2.89, rel=1e-9)
    # Every ceiling stays within the LDS width (128 B/clk per CU).
    for bits in (8, 16, 32, 64):
        for ordered in (False, True, "layout"):
            assert bench.gather_ceiling_GBps(bits, ordered) <= 256 * 2.4 * 128 + 1e-6


def test_workload_shapes_and_layout_rule():
    c1 = bench.build_workload("c1")
    assert len(c1) == 1 and c1[0]["adj"].shape == (256, 8)
    assert not bench.layout_applies(c1)  # below 8192 vertices
    c2 = bench.build_workload("c2")
    assert c2[0]["adj"].shape == (4096, 16) and c2[0]["groups"] == 4
    assert not bench.layout_applies(c2)  # symmetric groups keep their numbering
    c4 = bench.build_workload("c4")
    assert c4[0]["adj"].shape == (10000, 4)
    assert not bench.layout_applies(c4)  # a grid carries its own relabel
    fake = [dict(adj=np.zeros((65536, 16), np.int32), groups=1, based_nodes=65536)]
    assert bench.layout_applies(fake)
    assert not bench.layout_applies(fake * 2)  # batches are not laid out


def test_rows_ordered_rule():
    items = [dict(based_nodes=65536)]
    assert bench._rows_ordered({"row_order": 2, "word_bits": 16}, items) is True
    assert bench._rows_ordered({"row_order": 0, "word_bits": 16}, items) is False
    assert bench._rows_ordered({"row_order": 1, "word_bits": 16}, items) is True  # 4096 tiles
    assert bench._rows_ordered({"row_order": 1, "word_bits": 64}, [dict(based_nodes=10000)]) is False
    assert bench._rows_ordered({"row_order": 1, "word_bits": 16}, [dict(based_nodes=65536, adj_given=0)]) == "layout"


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
