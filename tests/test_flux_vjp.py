"""The reverse-mode numerics hook numerical_flux_fluctuations_vjp (used by the adjoint's face
kernel) against the transposed Jacobian of numerical_flux / numerical_fluctuations from
forward-mode duals through the same model functions: random wet states, bed ties, lake at
rest, a face that dries after hydrostatic reconstruction, the Kurganov-Petrova shallow branch
and identical states.  Tolerance 1e-11 relative to the largest gradient entry (both sides
are exact derivatives of the same FP64 arithmetic, so they differ only by rounding)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "zoomy-dmplex_amd", "models")
SRC = os.path.join(ROOT, "tests", "cpp", "flux_vjp_check.cpp")


@pytest.mark.parametrize("model", ["sme0", "sme1", "sme1m", "sme2", "vam"])
def test_flux_vjp_matches_dual_jacobian(model, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / f"fvc_{model}")
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I", os.path.join(MODELS, model),
                    "-I", MODELS, "-I", os.path.join(ROOT, "include"), SRC, "-o", exe], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("cases 400 bad 0"), r.stdout
