"""The operator concept as a Spectra-style consumer sees it.

tests/native/spectra_consumer.cpp is compiled with g++ against
include/latttmat.hpp alone and runs generic eigensolvers templated on the
operator type that call only rows(), cols(), perform_op(x, y) and
get_t_instance(), as SymEigsSolver / GenEigsSolver do with the reference's
operator classes (TMat/pfunc.hpp:52-54, 84-98).  Its eigenvalues must equal
tm_solve's (and, for the symmetric case, the oracle's dense spectrum)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "spectra_consumer")


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def _run(*args):
    assert os.path.exists(EXE), "tests/native/spectra_consumer is not built (run build())"
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return {ln.split()[0]: float(ln.split()[1]) for ln in r.stdout.splitlines()
            if ln.startswith("lambda0")}


def test_symmetric_consumer(lt, oracle):
    got = _run("sym", 12)["lambda0"]
    op = lt.TransferMatrix("A", 2, 1, 12, J=0.5, kappa=0.6)
    r = lt.solve(op, nev=1)
    assert abs(got / r.evals[0] - 1) < 1e-12
    D = oracle.OracleOp(oracle.XAT, 12, 0, 0, 0.5, 0.6, 0.0, 0.0).dense()
    assert abs(got / np.linalg.eigvalsh(0.5 * (D + D.T))[-1] - 1) < 1e-12


def test_transpose_instance_consumer(lt):
    got = _run("gen", 10)
    op = lt.TransferMatrix("B", 2, 3, 10, J=0.5, kappa=0.5, kappa2=0.3)
    r = lt.solve(op, nev=2, want_left=True, tol=1e-12)
    lam0 = r.evals[0].real
    assert abs(got["lambda0"] / lam0 - 1) < 1e-12
    assert abs(got["lambda0_T"] / lam0 - 1) < 1e-12
