"""bench.py's host-side pieces: the per-kernel summary of a rocprofv3 run
(launches per step, algorithmic bytes, counter scaling by the copy probe,
steady-state dispatches only) and the GPU-count guard."""

from __future__ import annotations

import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_summary_per_step_bytes_and_scaled_traffic(tmp_path):
    stats = tmp_path / "stats.csv"
    _write(stats, ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs",
                   "MaxNs", "StdDev"],
           [["void edm::k_force_tile<float, true>(edm::Cells<float>)", 100, 0, 50000, 0, 0, 0, 0],
            ["void edm::k_cat_tile<float>(edm::Cells<float>)", 200, 0, 40000, 0, 0, 0, 0],
            ["edm::k_frame_noise(long)", 1, 0, 400000, 0, 0, 0, 0]])
    pmc_h = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    # the copy probe: 512 MiB read reads as half, written as is
    _write(fetch, pmc_h, [["__amd_rocclr_copyBuffer", "FETCH_SIZE", 262144.0]] +
           [["void edm::k_force_tile<float, true>(edm::Cells<float>)", "FETCH_SIZE", v]
            for v in (90000.0, 90000.0, 10000.0, 10000.0)])
    _write(write, pmc_h, [["__amd_rocclr_copyBuffer", "WRITE_SIZE", 524288.0]] +
           [["void edm::k_force_tile<float, true>(edm::Cells<float>)", "WRITE_SIZE", v]
            for v in (5000.0, 5000.0, 8000.0, 8000.0)])
    w = dict(slots=1000, alive=900, cells=100, dye=10, walk=5, grid=16, steps=100, pass_=0,
             profiled_steps=100, probe_mib=512)
    w["pass"] = 20
    d = bench.summarise_profile(str(stats), str(fetch), str(write), w)
    ks = {k["kernel"]: k for k in d["kernels"]}
    assert "k_frame_noise" not in ks  # once per frame, not per step
    f = ks["k_force_tile<float, true>"]
    assert f["launches_per_step"] == 1.0 and f["avg_us"] == 50.0
    assert f["algorithmic_bytes"] == 18 * 1000
    # steady state: the second half of the dispatches; FETCH x 2, WRITE x 1
    assert f["traffic_bytes"] == round(10000.0 * 1024 * 2 + 8000.0 * 1024)
    assert ks["k_cat_tile<float>"]["launches_per_step"] == 2.0
    assert d["kernel_us_per_step"] == pytest.approx(50.0 + 2 * 40.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
