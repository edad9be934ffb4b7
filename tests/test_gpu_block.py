"""Block Krylov-Schur (tm_solve_opts.block = b, b vectors per Krylov step)
on the GPU against the one-vector Krylov-Schur.

The block form's Gram matrices [V | W]^T W come from the FP64 MFMA kernel
(v_mfma_f64_16x16x4f64) and its Cholesky / coupling work from device kernels;
both paths end in the same FP64 convergence test and the explicit residual
check, so the leading eigenvalues must agree to the Krylov tolerance.

  reduced ANNNI L=20 and L=28 (TMat/tmat.cpp:280-375, the symmetric XAM),
      FP64 and mixed precision, b = 2 and 4: evals to 1e-12 relative;
  3D 5x5 reduced (TMat/tmat3d.cpp:203-321), FP64, b = 4;
  eigenvectors of the block solve: explicit residual of each pair."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def _lead(evals, k):
    return np.sort(np.abs(np.asarray(evals).real))[::-1][:k]


@pytest.mark.parametrize("L", [20, 28])
@pytest.mark.parametrize("b", [2, 4])
def test_block_reduced_xam(lt, L, b):
    op = lt.TransferMatrix("A", 2, 5, L, J=0.4, kappa=0.5, neigs=4)
    ref = lt.solve(op, nev=4, tol=1e-12)
    r64 = lt.solve(op, nev=4, ncv=12, tol=1e-12, block=b)
    rmx = lt.solve(op, nev=4, ncv=12, precision=lt.TM_PREC_MIXED, block=b)
    e64 = np.max(np.abs(_lead(r64.evals, 4) / _lead(ref.evals, 4) - 1))
    emx = np.max(np.abs(_lead(rmx.evals, 4) / _lead(ref.evals, 4) - 1))
    print(f"L={L} b={b}: FP64 {e64:.1e} ({r64.matvecs_f64} T.x)  mixed {emx:.1e} "
          f"({rmx.matvecs_f32}+{rmx.matvecs_f64} T.x)  one-vector {ref.matvecs_f64} T.x",
          flush=True)
    assert e64 < 1e-12 and emx < 1e-12
    assert r64.extra["resid_check"] <= 1e-6 and rmx.extra["resid_check"] <= 1e-6


def test_block_3d_reduced(lt):
    op = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4)
    ref = lt.solve(op, nev=2, tol=1e-12)
    r = lt.solve(op, nev=2, ncv=12, tol=1e-12, block=4)
    e = np.max(np.abs(_lead(r.evals, 2) / _lead(ref.evals, 2) - 1))
    print(f"3D 5x5 b=4: {e:.1e}", flush=True)
    assert e < 1e-12


def test_block_vectors_residual(lt):
    """Ritz vectors of the block solve: ||T x - lambda x|| / |lambda| per pair."""
    op = lt.TransferMatrix("A", 2, 5, 20, J=0.4, kappa=0.5, neigs=3)
    r = lt.solve(op, nev=3, ncv=12, tol=1e-12, block=4, want_vectors=True)
    X = r.vecs
    for k in range(3):
        x = X[:, k].real
        res = np.linalg.norm(op.apply(x) - r.evals[k].real * x) / (abs(r.evals[k]) * np.linalg.norm(x))
        print(k, r.evals[k], res, flush=True)
        assert res < 1e-9
