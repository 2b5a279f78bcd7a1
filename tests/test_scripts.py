"""The command-line scripts beside the library parse their arguments and
print their usage without a GPU (their work needs one); on a GPU, bench.py
prints its one JSON line."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPTS = ["n-body-simulation_amd/ensemble.py", "n-body-simulation_amd/precision_study.py",
           "n-body-simulation_amd/snapshot.py", "n-body-simulation_amd/farfield_train.py",
           "n-body-simulation_amd/tuning.py", "bench.py"]


@pytest.mark.parametrize("script", SCRIPTS)
def test_usage(script):
    r = subprocess.run([sys.executable, os.path.join(ROOT, script), "--help"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "usage:" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_json_line():
    """bench.py on a small disc (20k bodies, 5 timed steps, 2 reference CPU
    steps) prints exactly one JSON line with every field of its contract,
    its figures consistent with each other."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--n", "20000",
                        "--steps", "5", "--warmup", "1", "--cpu-steps", "2"],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [line for line in r.stdout.splitlines() if line.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 1
    assert d["value"] > 0 and d["higher_is_better"] is True
    assert abs(d["value"] * d["ms_per_step"] / 1e3 - 1.0) < 1e-2
    assert "workload" in d["config"] and d["config"]["n_bodies"] == 20000
    # The metric names what was timed: the body count and the step window.
    assert "20000 bodies" in d["metric"] and "steps 1..5" in d["metric"]
    acc = d["accuracy"]
    # Reference law (fastInvSqrt): its ~1.1e-3 RMS floor against FP64.
    assert acc["targets"] == 4096 and 0 < acc["rms"] < 5e-3 and acc["max_rel"] > 0
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert 0 < rf["frac"] < 1 and abs(rf["achieved"] / rf["peak"] - rf["frac"]) < 1e-3
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("reference", "port")
    assert cb["steps"] == 2 and cb["cpu_model"] and cb["gpu_value_same_window"] > 0
    eq = d["equal_error"]
    assert eq["learned"]["theta"] == 1.2 and eq["plain"]["rms"] <= eq["learned"]["rms"]
    assert eq["learned"]["steps_per_s"] > 0 and eq["plain"]["steps_per_s"] > 0
    assert eq["speedup_learned_over_plain"] == pytest.approx(
        eq["learned"]["steps_per_s"] / eq["plain"]["steps_per_s"], rel=1e-3)
