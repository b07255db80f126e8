"""The binomial checkpoint schedule of the adjoint's reverse sweep (csrc/binomial.h), compiled
with g++: its dynamic programme's re-run count against Griewank and Walther's closed form for
T <= 120 steps and 1..8 slots, and simulated sweeps (optimal and equal-split) that must visit
every step's adjoint in reverse order from that step's own state without overwriting a held
slot."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "binomial_check.cpp")


def test_binomial_schedule(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "binomial_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "zoomy-dmplex_amd", "csrc"), SRC,
                    "-o", exe], check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "cases 1590 bad 0", r.stdout
