"""The parts of bench.py that need no GPU: the one-GPU guard, kernel-name
keys, the counter summary with its copy-probe calibration, the committed
traffic summary behind the bench line's `stages` and `roofline.traffic`,
and the serving backlog rule."""

import csv
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_kernel_key_drops_parameters_and_namespaces():
    k = bench._kernel_key
    assert k("void mv::(anonymous namespace)::herk_split_kernel<2, 1, false, "
             "128>(mv::(anonymous namespace)::SplitProblem, mv::(anonymous "
             "namespace)::SliceCfg)") == "mv::herk_split_kernel<2, 1, false, 128>"
    assert k("mv::(anonymous namespace)::band_zero_kernel(int*, int)") == \
        "mv::band_zero_kernel"
    assert k("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def _counter_csv(path, counter, rows):
    path.mkdir(parents=True)
    with open(path / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Counter_Name", "Counter_Value",
                                "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, kib, t0, t1 in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": kib, "Start_Timestamp": t0,
                        "End_Timestamp": t1})


def test_summarize_counters_scales_by_the_copy_probe(tmp_path):
    """The 1 GiB copy reads as FETCH_SIZE x 0.5 and WRITE_SIZE x 1 of its
    bytes here: kernel figures are scaled by 2 and 1, durations are the
    dispatches' means over both runs."""
    gib_kib = float(1 << 20)
    k = "void mv::(anonymous namespace)::zmat_fused_kernel<2, 2, true>(int)"
    _counter_csv(tmp_path / "f", "FETCH_SIZE",
                 [("__amd_rocclr_copyBuffer", gib_kib / 2, 0, 100),
                  (k, 1000.0, 0, 100_000), (k, 3000.0, 0, 300_000)])
    _counter_csv(tmp_path / "w", "WRITE_SIZE",
                 [("__amd_rocclr_copyBuffer", gib_kib, 0, 100),
                  (k, 500.0, 0, 200_000), (k, 500.0, 0, 200_000)])
    out = bench.summarize_counters(str(tmp_path / "f"), str(tmp_path / "w"))
    assert out["_calibration"]["fetch_scale"] == pytest.approx(2.0)
    assert out["_calibration"]["write_scale"] == pytest.approx(1.0)
    z = out["mv::zmat_fused_kernel<2, 2, true>"]
    assert z["dispatches"] == 2
    assert z["read_bytes"] == pytest.approx(2.0 * 2000.0 * 1024)
    assert z["write_bytes"] == pytest.approx(500.0 * 1024)
    assert z["bytes_per_launch"] == pytest.approx(z["read_bytes"]
                                                  + z["write_bytes"])
    assert z["mean_us_counter_runs"] == pytest.approx(200.0)
    assert "__amd_rocclr_copyBuffer" not in out


def test_committed_traffic_feeds_stages_and_roofline():
    rec = json.loads(bench.TRAFFIC.read_text())
    assert rec["_workload"]["nant"] == bench.NANT
    assert rec["_workload"]["nsrc"] == bench.NSRC
    stages = bench.stage_traffic(bench.NANT, bench.NSRC, bench.NFREQ)
    assert [s for s, _, _ in bench.STAGE_KERNELS] == list(stages)
    herk = stages["HERK (split-BF16)"]
    assert herk["launches_per_step"] == bench.NFREQ
    # the HERK reads at least its Z terms once: 700 rows x 2 x 59.7k
    # columns x 2 terms x 4 B
    assert herk["read_bytes"] >= 0.6e9
    t = bench.measured_traffic("mv::herk_split_kernel<2, 1, false, 128>",
                               bench.NANT, bench.NSRC)
    assert t == pytest.approx(rec["mv::herk_split_kernel<2, 1, false, 128>"]
                              ["bytes_per_launch"])
    # another workload has no committed figures
    assert bench.stage_traffic(64, 1000, 4) is None
    assert bench.measured_traffic("mv::expand_kernel<double>", 64, 1000) is None


def test_backlog_slope_over_the_last_two_thirds():
    steady = [(t, 5.0 + 0.01 * np.sin(t)) for t in np.arange(0, 60, 0.5)]
    assert abs(bench.backlog_slope(steady, 60.0)) < 0.01
    growing = [(t, 3.0 * t) for t in np.arange(0, 60, 0.5)]
    assert bench.backlog_slope(growing, 60.0) == pytest.approx(3.0)
    # the first third (filling up) is not counted
    ramp = [(t, 10.0 * t if t < 20 else 200.0) for t in np.arange(0, 60, 0.5)]
    assert abs(bench.backlog_slope(ramp, 60.0)) < 1e-9
    assert bench.backlog_slope([(1.0, 2.0)], 60.0) == 0.0
