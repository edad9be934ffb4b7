"""GPU eigensolves against dense LAPACK, the Kaufman closed form and the
full-vs-reduced identity.  Eigenvalue bar: 1e-12 relative."""
import numpy as np
import pytest

from tests import physics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def test_xat_l12_dense(lt):
    # config 1: A, axialdir 1, L=12, betaJ=0.5, kappa1=0.6, h=0, neigs=1
    op = lt.TransferMatrix("A", 2, 1, 12, J=0.5, kappa=0.6)
    r = lt.solve(op, nev=1, want_vectors=True)
    print("matvecs", r.matvecs_f64, "restarts", r.restarts, flush=True)
    from oracle import oracle as O
    oop = O.OracleOp(O.XAT, 12, 0, 0, 0.5, 0.6, 0.0, 0.0)
    D = oop.dense()
    w = np.linalg.eigvalsh(0.5 * (D + D.T))
    lam0 = w[-1]
    assert abs(r.evals[0] / lam0 - 1) < 1e-12
    v = r.vecs[:, 0]
    res = np.linalg.norm(D @ v - r.evals[0] * v) / abs(r.evals[0])
    assert res < 1e-9


@pytest.mark.parametrize("Lw,K", [(8, 0.5), (12, 0.3), (14, 0.44)])
def test_kaufman(lt, Lw, K):
    op = lt.TransferMatrix("A", 2, 1, Lw, J=K, kappa=0.0)
    r = lt.solve(op, nev=2)
    l0, l1 = physics.kaufman(Lw, K)
    nf = op.getnormfactor()
    got = np.sort(np.abs(r.evals))[::-1] / nf
    print(Lw, K, got, l0, l1, flush=True)
    assert abs(got[0] / l0 - 1) < 1e-12
    assert abs(got[1] / l1 - 1) < 1e-11


def test_reduced_equals_full(lt):
    full = lt.TransferMatrix("A", 2, 1, 16, J=0.4, kappa=0.5)
    red = lt.TransferMatrix("A", 2, 5, 16, J=0.4, kappa=0.5, neigs=1)
    rf = lt.solve(full, nev=1)
    rr = lt.solve(red, nev=1)
    assert abs(rr.evals[0] / rf.evals[0] - 1) < 1e-12


def test_zbt_nonsym_dense(lt):
    op = lt.TransferMatrix("B", 2, 3, 10, J=0.5, kappa=0.5, kappa2=0.3)
    r = lt.solve(op, nev=2, want_left=True, want_vectors=True)
    from oracle import oracle as O
    D = O.OracleOp(O.ZBT, 10, 0, 0, 0.5, 0.5, 0.3, 0.0).dense()
    w = np.linalg.eigvals(D)
    w = w[np.argsort(-np.abs(w))]
    print("gpu", r.evals, "dense", w[:3], flush=True)
    assert abs(r.evals[0] / w[0] - 1) < 1e-12
    assert abs(abs(r.evals[1]) / abs(w[1]) - 1) < 1e-10
    # the left solve stops at residual tol 1e-10 (pfunc.hpp:52 uses the same
    # defaults); for a non-symmetric matrix the eigenvalue error is first order
    assert abs(r.eval_left / w[0].real - 1) < 1e-10
    # left vector is the right Perron vector of T^T
    lvec = r.vecl
    assert np.linalg.norm(D.T @ lvec - r.eval_left * lvec) / abs(r.eval_left) < 1e-8


def test_mixed_precision_reduced(lt):
    op = lt.TransferMatrix("A", 2, 5, 20, J=0.4, kappa=0.5, neigs=4)
    r64 = lt.solve(op, nev=4, precision=lt.TM_PREC_F64)
    rmx = lt.solve(op, nev=4, precision=lt.TM_PREC_MIXED)
    print("f64", r64.evals, r64.matvecs_f64, "mixed", rmx.evals, rmx.matvecs_f32,
          rmx.matvecs_f64, flush=True)
    rel = np.abs(rmx.evals - r64.evals) / np.abs(r64.evals)
    assert rel[0] <= 1e-12
    assert np.all(rel <= 1e-11)


@pytest.mark.parametrize("extra,basis32", [(0, True), (2, True), (1, False)])
def test_mixed_refine_block_and_basis(lt, extra, basis32):
    # FP64 Rayleigh-Ritz block of nev + extra FP32 Ritz vectors, FP32- or
    # FP64-stored basis in the FP32 phase: same eigenvalues and vectors as FP64
    op = lt.TransferMatrix("A", 2, 5, 20, J=0.4, kappa=0.5, neigs=4)
    r64 = lt.solve(op, nev=4, precision=lt.TM_PREC_F64, tol=1e-13, want_vectors=True)
    rmx = lt.solve(op, nev=4, ncv=13, precision=lt.TM_PREC_MIXED, refine_extra=extra,
                   basis_f32=basis32, want_vectors=True)
    print("extra", extra, "basis32", basis32, "mv", rmx.matvecs_f32, rmx.matvecs_f64, flush=True)
    rel = np.abs(rmx.evals - r64.evals) / np.abs(r64.evals)
    assert np.all(rel <= 1e-12), rel
    for c in range(4):
        ov = abs(rmx.vecs[:, c] @ r64.vecs[:, c])
        assert abs(ov - 1) < 1e-9, (c, ov)


def test_xdt_dense(lt):
    op = lt.TransferMatrix("D", 2, 1, 10, J=0.6, kappa=0.7, h=0.1)
    r = lt.solve(op, nev=1)
    T = physics.xdt_dense(10, 0.6, 0.7, 0.1) * op.getnormfactor()
    lam0 = np.linalg.eigvalsh(T)[-1]
    assert abs(r.evals[0] / lam0 - 1) < 1e-12


@pytest.mark.parametrize("ncv", [6, 16, 30])
@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_lagged_lanczos_basis_sizes(lt, ncv, prec):
    # symmetric solves run lagged Lanczos steps (the residual's correction
    # and normalisation inside the next step's sweep, the last one inside the
    # restart); a 30-vector basis also runs the steps past the lagged sweep's
    # 24 vectors, which finish v_j first
    # (the symmetry-reduced operator: the full one's momentum pairs are
    # degenerate, and a single start vector finds one copy of a degenerate
    # pair or two depending on rounding)
    op = lt.TransferMatrix("A", 2, 5, 14, J=0.45, kappa=0.55, neigs=4)
    p = lt.TM_PREC_F64 if prec == "f64" else lt.TM_PREC_MIXED
    r = lt.solve(op, nev=4, ncv=ncv, precision=p, want_vectors=True)
    from oracle import oracle as O
    D = O.OracleOp(O.XAM2D, 14, 0, 2, 0.45, 0.55, 0.0, 0.0).dense()
    D = 0.5 * (D + D.T)
    w = np.sort(np.linalg.eigvalsh(D))[::-1]
    print(ncv, prec, "restarts", r.restarts, r.matvecs_f32, r.matvecs_f64, flush=True)
    assert np.all(np.abs(np.asarray(r.evals) / w[:4] - 1) < 1e-12), (r.evals, w[:4])
    for c in range(4):
        v = r.vecs[:, c]
        res = np.linalg.norm(D @ v - r.evals[c] * v) / abs(r.evals[c])
        assert res < 1e-9, (c, res)


@pytest.mark.parametrize("ncv", [6, 16, 30])
def test_lagged_arnoldi_basis_sizes(lt, ncv):
    # non-symmetric solves run lagged Arnoldi steps; a 30-vector basis also
    # runs the steps past the lagged sweep's 24 vectors
    op = lt.TransferMatrix("B", 2, 3, 10, J=0.5, kappa=0.5, kappa2=0.3)
    r = lt.solve(op, nev=2, ncv=ncv)
    from oracle import oracle as O
    D = O.OracleOp(O.ZBT, 10, 0, 0, 0.5, 0.5, 0.3, 0.0).dense()
    w = np.linalg.eigvals(D)
    w = w[np.argsort(-np.abs(w))]
    print(ncv, "restarts", r.restarts, r.matvecs_f64, r.evals, flush=True)
    assert abs(r.evals[0] / w[0] - 1) < 1e-12
    assert min(abs(r.evals[1] - w[1]), abs(r.evals[1] - w[2])) / abs(w[1]) < 1e-9
