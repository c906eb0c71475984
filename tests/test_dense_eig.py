"""The host eigensolver of the Lanczos projected problem (csrc/dense_eig.hpp) against
LAPACK (numpy.linalg.eigh) on the matrix shapes the restarted Lanczos produces."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "null--space-free-6d-projection-method_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("eig") / "dense_eig")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "dense_eig_driver.cpp"),
                    "-o", exe], check=True)
    return exe


def run(exe, A):
    m = A.shape[0]
    inp = "%d\n" % m + "\n".join("%.17g" % x for x in A.ravel(order="F"))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    vals = np.array([float(x) for x in out])
    return vals[:m], vals[m:].reshape((m, m), order="F")


def arrowhead_tridiagonal(rng, k, m):
    """H of a thick-restarted Lanczos cycle: diag(d) (k kept Ritz values) coupled to step k+1
    by c, then the tridiagonal of the new steps."""
    H = np.zeros((m, m))
    d = np.sort(rng.random(k))[::-1] * 200.0
    H[np.arange(k), np.arange(k)] = d
    alpha = rng.random(m - k) * 50
    beta = rng.random(m - k - 1) * 20
    H[np.arange(k, m), np.arange(k, m)] = alpha
    H[np.arange(k, m - 1), np.arange(k + 1, m)] = beta
    H[np.arange(k + 1, m), np.arange(k, m - 1)] = beta
    c = rng.standard_normal(k) * 1e-3
    H[:k, k] = c
    H[k, :k] = c
    return H


@pytest.mark.parametrize("case", ["tridiag40", "arrow188", "dense64", "clustered", "diag", "one"])
def test_sym_eig_matches_lapack(driver, case):
    rng = np.random.default_rng(sum(map(ord, case)))
    if case == "tridiag40":
        A = arrowhead_tridiagonal(rng, 0, 40)
    elif case == "arrow188":
        A = arrowhead_tridiagonal(rng, 140, 188)
    elif case == "dense64":
        B = rng.standard_normal((64, 64))
        A = B + B.T
    elif case == "clustered":  # pairs split by 1e-12 (the near-degenerate pairs of the spectrum)
        Q, _ = np.linalg.qr(rng.standard_normal((50, 50)))
        ev = np.repeat(rng.random(25), 2) + np.tile([0.0, 1e-12], 25)
        A = (Q * ev) @ Q.T
        A = 0.5 * (A + A.T)
    elif case == "diag":
        A = np.diag(rng.standard_normal(30))
    else:
        A = np.array([[3.5]])
    w, V = run(driver, A)
    wr = np.linalg.eigvalsh(A)
    scale = max(np.max(np.abs(wr)), 1e-300)
    assert np.max(np.abs(w - wr)) <= 1e-13 * scale
    assert np.max(np.abs(V.T @ V - np.eye(len(w)))) <= 1e-13
    assert np.max(np.abs(A @ V - V * w)) <= 1e-13 * scale
