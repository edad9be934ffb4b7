"""Eigen driver (latttmat_amd/csrc/tm_solver.cpp + tm_dense.cpp) on the CPU.

tests/native builds the driver with g++ against a host stand-in for the
few HIP runtime calls it makes and a dense-matrix operator, so the restart
logic, the conjugate-pair handling, the mixed-precision refinement and the
retry/left-vector semantics of pfunc_inner (TMat/pfunc.hpp:39-114) are
checked here without a GPU, against LAPACK (numpy)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
EXE = os.path.join(NATIVE, "driver_test")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-j3", "-C", NATIVE], check=True, capture_output=True)
    return EXE


def run(driver, A, sym, nev, ncv=0, prec=0, left=0, tmp=None, basis32=0):
    n = A.shape[0]
    path = os.path.join(tmp, "m.txt")
    with open(path, "w") as f:
        f.write(f"{n} {int(sym)} {nev} {ncv} {prec} {left} {basis32}\n")
        f.write(" ".join("%.17g" % v for v in A.flatten(order="F")))
    r = subprocess.run([driver, path], capture_output=True, text=True, timeout=120)
    out = {"evals": [], "res": []}
    for line in r.stdout.splitlines():
        p = line.split()
        if p[0] == "status":
            out["status"] = int(p[1])
            out["nconv"] = int(p[3])
            out["mv64"], out["mv32"] = int(p[5]), int(p[6])
        elif p[0] == "eval":
            out["evals"].append(complex(float(p[1]), float(p[2])))
            out["res"].append(float(p[3]))
        elif p[0] == "left":
            out["left"] = float(p[1])
    return out


def test_symmetric_xat(driver, oracle, tmp_path):
    D = oracle.OracleOp(oracle.XAT, 9, 0, 0, 0.5, 0.6, 0.0, 0.0).dense()
    D = 0.5 * (D + D.T)
    w = np.sort(np.linalg.eigvalsh(D))[::-1]
    for nev in (1, 2, 4):
        o = run(driver, D, True, nev, tmp=str(tmp_path))
        assert o["status"] == 0 and o["nconv"] == nev
        got = np.array([e.real for e in o["evals"]])
        assert np.all(np.abs(got / w[:nev] - 1) < 1e-12), (got, w[:nev])


def test_nonsymmetric_zbt_pairs_and_left(driver, oracle, tmp_path):
    D = oracle.OracleOp(oracle.ZBT, 10, 0, 0, 0.5, 0.5, 0.3, 0.0).dense()
    w = np.linalg.eigvals(D)
    w = w[np.argsort(-np.abs(w))]
    o = run(driver, D, False, 2, left=1, tmp=str(tmp_path))
    assert o["status"] == 0
    assert abs(o["evals"][0] / w[0] - 1) < 1e-12
    # lambda1 is one of a complex-conjugate pair (modulated phase)
    assert abs(w[1].imag) > 1
    assert min(abs(o["evals"][1] - w[1]), abs(o["evals"][1] - w[2])) / abs(w[1]) < 1e-9
    # left eigenvalue of T^T equals lambda0 (residual tol 1e-10)
    assert abs(o["left"] / w[0].real - 1) < 1e-10


@pytest.mark.parametrize("basis32", [0, 1])
def test_mixed_precision_refinement(driver, oracle, tmp_path, basis32):
    # operator results rounded to FP32 in phase 1 (basis stored in FP64 or
    # FP32); FP64 Rayleigh-Ritz after
    D = oracle.OracleOp(oracle.XAM2D, 12, 0, 2, 0.4, 0.5, 0.0, 0.0).dense()
    D = 0.5 * (D + D.T)
    w = np.sort(np.linalg.eigvalsh(D))[::-1]
    o = run(driver, D, True, 4, prec=1, tmp=str(tmp_path), basis32=basis32)
    assert o["mv32"] > 0 and o["mv64"] > 0
    got = np.array([e.real for e in o["evals"]])
    assert abs(got[0] / w[0] - 1) <= 1e-12
    assert np.all(np.abs(got / w[:4] - 1) <= 1e-11)


def test_small_matrix_clamps_ncv(driver, tmp_path):
    # ncv >= cols() is clamped to cols()-1 (TMat/pfunc.hpp:90)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6))
    A = A @ A.T
    o = run(driver, A, True, 2, tmp=str(tmp_path))
    w = np.sort(np.linalg.eigvalsh(A))[::-1]
    got = np.array([e.real for e in o["evals"]])
    assert np.all(np.abs(got / w[:2] - 1) < 1e-12)


def test_invariant_subspace_breakdown(driver, tmp_path):
    # A start vector that lives in a 3-dimensional invariant subspace cannot
    # fill an ncv=7 basis: the device-side step flags the breakdown and the
    # host redoes the cycle from there with a fresh orthogonal direction.
    n = 40
    evals = np.repeat([5.0, 3.0, 1.0], [1, 1, n - 2])
    Q = np.linalg.qr(np.random.default_rng(7).standard_normal((n, n)))[0]
    A = (Q * evals) @ Q.T
    o = run(driver, A, True, 2, tmp=str(tmp_path))
    got = np.array([e.real for e in o["evals"]])
    assert o["status"] == 0
    assert np.all(np.abs(got / np.array([5.0, 3.0]) - 1) < 1e-12), got


@pytest.mark.parametrize("ncv", [5, 12, 30])
def test_symmetric_lagged_steps(driver, oracle, tmp_path, ncv):
    # Lanczos steps whose correction is finished in the next step's sweep:
    # short cycles (one lagged step), the usual basis, and a basis beyond the
    # lagged kernel's 24 vectors (the steps past it finish v_j first)
    D = oracle.OracleOp(oracle.XAT, 10, 0, 0, 0.45, 0.55, 0.0, 0.0).dense()
    D = 0.5 * (D + D.T)
    w = np.sort(np.linalg.eigvalsh(D))[::-1]
    for prec in (0, 1):
        o = run(driver, D, True, 4, ncv=ncv, prec=prec, tmp=str(tmp_path), basis32=prec)
        assert o["status"] == 0 and o["nconv"] == 4
        got = np.array([e.real for e in o["evals"]])
        assert np.all(np.abs(got / w[:4] - 1) < 1e-11), (ncv, prec, got, w[:4])


@pytest.mark.parametrize("ncv", [5, 12, 30])
def test_nonsymmetric_lagged_steps(driver, oracle, tmp_path, ncv):
    # lagged Arnoldi steps (first projection from a sweep of the basis
    # against T r, correction and normalisation in the next step's sweep or
    # the restart); 30 vectors also run the steps past the lagged sweep's 24
    D = oracle.OracleOp(oracle.ZBT, 10, 0, 0, 0.5, 0.5, 0.3, 0.0).dense()
    w = np.linalg.eigvals(D)
    w = w[np.argsort(-np.abs(w))]
    o = run(driver, D, False, 2, ncv=ncv, tmp=str(tmp_path))
    assert o["status"] == 0 and o["nconv"] == 2
    assert abs(o["evals"][0] / w[0] - 1) < 1e-12, (ncv, o["evals"], w[:3])
    # lambda1 is one of a complex-conjugate pair
    assert min(abs(o["evals"][1] - w[1]), abs(o["evals"][1] - w[2])) / abs(w[1]) < 1e-9
