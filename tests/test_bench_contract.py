"""CPU checks of bench.py's contract: one GPU only, the algorithmic byte
counts quoted in the roofline, and the committed traffic measurement."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_algorithmic_bytes_per_element_cycle():
    npe = 125341 / 120000  # nodes per element of the 120 000-hex cylinder
    # FP64: 8 IPs x (14 doubles read + written + detJ0) + conn + forces + 6 doubles per node
    assert bench.algo_bytes_per_elem("fp64", npe) == pytest.approx(8 * (14 * 8 * 2 + 8) + 32 + 192 + npe * 48)
    # mixed: stress 6 + dep, sy, gamma in FP32 and ep, T, e_int, e_inel in FP64 = 68 B each way
    assert bench.algo_bytes_per_elem("mixed", npe) == pytest.approx(8 * (68 * 2 + 4) + 32 + 96 + npe * 48)
    assert round(bench.algo_bytes_per_elem("mixed", npe), 1) == 1298.1


def test_measured_traffic_only_for_its_workload():
    t = bench.measured_traffic("taylor3d", "mixed", 20, 100)
    with open(os.path.join(ROOT, "profiles", "k_elem_mixed_hex8_traffic.json")) as f:
        d = json.load(f)
    expect = (d["fetch_size_kib_per_launch"] * d["fetch_scale"] + d["write_size_kib_per_launch"] * d["write_scale"]) * 1024
    assert t == pytest.approx(expect, rel=1e-3)
    assert bench.measured_traffic("taylor3d", "fp64", 20, 100) is None
    assert bench.measured_traffic("necking", "mixed", 40, 208) is None
