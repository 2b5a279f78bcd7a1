"""The command-line scripts beside the library parse their arguments and
print their usage without a GPU (their work needs one)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPTS = ["n-body-simulation_amd/ensemble.py", "n-body-simulation_amd/precision_study.py",
           "n-body-simulation_amd/snapshot.py", "bench.py"]


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
