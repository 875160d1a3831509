"""The cut of a batch of grid requests into runs (host logic, CPU only):
tests/cpp/grid_plan_check.cpp checks sdfx_grid_plan.hpp's plan_grid_runs on
random batches (every request once, in order, runs within the cap unless a
single request exceeds it, runs greedy)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_plan_grid_runs(tmp_path):
    exe = tmp_path / "grid_plan_check"
    src = os.path.join(ROOT, "tests", "cpp", "grid_plan_check.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
