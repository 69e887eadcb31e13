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


def test_element_kernel_names_the_launched_instantiation():
    sched = {"elem_block": 64, "waves": 0, "corr_batch": 4, "graph_cycles": 80}
    assert bench.element_kernel("mixed", sched) == "k_elem_mixed<8, 3, 64, false, true, 0, false>"
    assert bench.element_kernel("mixed-rr32", sched) == "k_elem_mixed<8, 3, 64, false, false, 0, false>"
    assert bench.element_kernel("fp64", dict(sched, waves=3)) == "k_elem_fp64<8, 3, 64, false, false, 3, false>"


def test_measured_traffic_only_for_its_kernel_and_workload():
    with open(os.path.join(ROOT, "profiles", "element_kernels_hbm_traffic.json")) as f:
        d = json.load(f)
    assert d["kernels"]
    for name, k in d["kernels"].items():
        expect = (k["fetch_size_kib_per_launch"] * d["fetch_scale"]
                  + k["write_size_kib_per_launch"] * d["write_scale"]) * 1024
        assert bench.measured_traffic("taylor3d", name, 20, 100) == pytest.approx(expect, rel=1e-3)
        assert bench.measured_traffic("necking", name, 40, 208) is None
        assert bench.measured_traffic("taylor3d", name, 10, 50) is None
    assert bench.measured_traffic("taylor3d", "k_elem_mixed<8, 3, 32, false, true, 0>", 20, 100) is None
