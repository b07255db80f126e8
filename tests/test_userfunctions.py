"""include/zoomy/UserFunctions.H: the dense kernels a generated Model.H / Numerics.H calls
(eigenvalues, eigensystem, solve), checked through tests/cpp/userfunctions_check.cpp.

- The reference's own known answers (test_userfunctions.cpp:8-110): the shallow-water
  quasilinear matrix with u = 2, c = 3 has eigenvalues {-1, 5}, R diag(lambda) L = A and
  L R = I; diag(2, 3) after it; the 3 x 3 system with x = [1, 1, 3]; the identity system; a
  non-finite matrix gives +inf eigenvalues and R = L = I (so |A| = R |Lambda| L is inf, never
  NaN), and a finite matrix afterwards is exact again.
- Random matrices with a known real spectrum (S diag(d) S^-1, n = 2..8): eigenvalues against
  numpy, the eigenbasis reproduces A and L = R^-1; a complex pair gives its real part twice;
  a defective matrix still returns finite R and L (the pseudo-inverse path).
- Random linear systems against numpy.linalg.solve.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "userfunctions_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("uf") / "userfunctions_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), SRC, "-o", exe],
                   check=True)

    def run(lines):
        out = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
        return [[float(v) for v in ln.split()[1:]] for ln in out.stdout.strip().splitlines()]

    return run


def eig_line(A):
    A = np.asarray(A, float)
    return "E %d %s" % (A.shape[0], " ".join(repr(float(v)) for v in A.ravel()))


def solve_line(A, b):
    A = np.asarray(A, float)
    return "S %d %s %s" % (A.shape[0], " ".join(repr(float(v)) for v in A.ravel()),
                           " ".join(repr(float(v)) for v in b))


def split(row, n):
    w = np.array(row[:n])
    ew = np.array(row[n:2 * n])
    R = np.array(row[2 * n:2 * n + n * n]).reshape(n, n)
    L = np.array(row[2 * n + n * n:]).reshape(n, n)
    return w, ew, R, L


def test_reference_known_answers(checker):
    u, c = 2.0, 3.0
    swe = [[0.0, 1.0], [c * c - u * u, 2 * u]]
    inf, nan = float("inf"), float("nan")
    rows = checker([eig_line(swe), eig_line([[2.0, 0.0], [0.0, 3.0]]),
                    solve_line([[2, 1, 0], [1, 3, 1], [0, 1, 4]], [3, 7, 13]),
                    solve_line(np.eye(3), [7, 8, 9]),
                    eig_line([[0.0, 1.0], [inf, 4.0]]), eig_line([[0.0, 1.0], [nan, 4.0]]),
                    eig_line([[2.0, 0.0], [0.0, 3.0]])])
    w, ew, R, L = split(rows[0], 2)
    assert np.allclose(sorted(w), [u - c, u + c], atol=1e-10)
    assert np.allclose(sorted(ew), [u - c, u + c], atol=1e-10)
    assert np.max(np.abs(R @ np.diag(ew) @ L - np.array(swe))) <= 1e-10
    assert np.max(np.abs(L @ R - np.eye(2))) <= 1e-10
    absA = R @ np.diag(np.abs(ew)) @ L
    assert np.all(np.isfinite(absA))
    assert max(abs(x) for x in w) == pytest.approx(5.0, abs=1e-10)
    assert np.allclose(sorted(rows[1][:2]), [2.0, 3.0], atol=1e-10)
    assert np.allclose(rows[2], [1, 1, 3], atol=1e-10)
    assert np.allclose(rows[3], [7, 8, 9], atol=1e-15)
    for r in rows[4:6]:
        w, ew, R, L = split(r, 2)
        assert np.all(np.isinf(w)) and np.all(w > 0) and np.all(np.isinf(ew)) and np.all(ew > 0)
        assert np.array_equal(R, np.eye(2)) and np.array_equal(L, np.eye(2))
        assert not np.isnan(R[0, 0] * abs(ew[0]) * L[0, 0])
    assert np.allclose(sorted(rows[6][:2]), [2.0, 3.0], atol=1e-10)


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8])
def test_real_spectrum_and_eigenbasis(checker, n):
    rng = np.random.default_rng(100 + n)
    lines, mats = [], []
    for _ in range(6):
        d = rng.uniform(-5, 5, n)
        S = rng.standard_normal((n, n)) + 2 * np.eye(n)
        A = S @ np.diag(d) @ np.linalg.inv(S)
        mats.append((A, d))
        lines.append(eig_line(A))
    for (A, d), row in zip(mats, checker(lines)):
        w, ew, R, L = split(row, n)
        scale = max(1.0, np.max(np.abs(d)))
        assert np.allclose(np.sort(w), np.sort(d), atol=1e-9 * scale), (np.sort(w), np.sort(d))
        assert np.allclose(np.sort(ew), np.sort(d), atol=1e-9 * scale)
        assert np.max(np.abs(R @ np.diag(ew) @ L - A)) <= 1e-8 * np.max(np.abs(A))
        assert np.max(np.abs(L @ R - np.eye(n))) <= 1e-8
        assert np.allclose(np.linalg.norm(R, axis=0), 1.0, atol=1e-12)   # unit eigenvectors


def test_complex_pair_and_defective_matrix(checker):
    rot = [[1.0, -2.0, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, -3.0]]   # 1 +- 2i and -3
    jordan = [[2.0, 1.0], [0.0, 2.0]]
    rows = checker([eig_line(rot), eig_line(jordan)])
    assert np.allclose(sorted(rows[0][:3]), [-3.0, 1.0, 1.0], atol=1e-10)
    w, ew, R, L = split(rows[1], 2)
    assert np.allclose(w, [2.0, 2.0], atol=1e-7)
    assert np.all(np.isfinite(R)) and np.all(np.isfinite(L))


def test_sme_quasilinear_matrix_spectrum(checker):
    """The 1-D SME level 0 quasilinear matrix (Model.H:75-91 form): eigenvalues 0, u -+ sqrt(g h)."""
    g = 9.81
    rng = np.random.default_rng(7)
    lines, want = [], []
    for _ in range(8):
        h, u = rng.uniform(0.1, 2.0), rng.uniform(-2, 2)
        A = [[0, 0, 0], [0, 0, 1], [g * h, g * h - u * u, 2 * u]]
        lines.append(eig_line(A))
        want.append(sorted([0.0, u - np.sqrt(g * h), u + np.sqrt(g * h)]))
    for row, w in zip(checker(lines), want):
        assert np.allclose(sorted(row[:3]), w, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_solve_matches_numpy(checker, n):
    rng = np.random.default_rng(200 + n)
    cases = [(rng.standard_normal((n, n)) + n * np.eye(n), rng.standard_normal(n)) for _ in range(5)]
    rows = checker([solve_line(A, b) for A, b in cases])
    for (A, b), x in zip(cases, rows):
        ref = np.linalg.solve(A, b)
        assert np.allclose(x, ref, rtol=1e-12, atol=1e-13)
