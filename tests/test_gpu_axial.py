"""GPU parity of the ANNNI axial transfer matrices over two-layer states:
YAT (A axialdir 2) with its transpose, and Y2AT (A axialdir 0).

The device path applies the L single-bit steps of the reference as one
Kronecker product over a layer, plus an LDS-tiled layer swap with the
popcount diagonal (YAT) or two diagonals (Y2AT).  The oracle restates the
reference's step loops (TMat/tmat.cpp:126-302) and is itself pinned to the
two-layer Hamiltonian matrices.  Tolerances: FP64 1e-13, FP32 2e-5 of
max|y|; eigenvalues 1e-12 relative (non-symmetric solves stop at the
residual tolerance, so their eigenvalue test uses 1e-10 with tol 1e-10).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 2e-5


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as m
    return m


def _vec(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("size", [4, 5, 6, 7, 8, 9, 10, 11])
@pytest.mark.parametrize("axialdir", [2, 0])
def test_axial_apply(lt, oracle, size, axialdir):
    kw = dict(J=0.45, kappa=0.35, h=0.1 if size % 2 else 0.0)
    op = lt.TransferMatrix("A", 2, axialdir, size, **kw)
    kind = oracle.YAT if axialdir == 2 else oracle.Y2AT
    oop = oracle.OracleOp(kind, size, 0, 0, kw["J"], kw["kappa"], 0.0, kw["h"])
    assert op.cols() == oop.n == 1 << (2 * size)
    assert op.kind() == ("YAT" if axialdir == 2 else "Y2AT") and not op.is_symmetric()
    assert abs(op.getnormfactor() / oop.normfactor - 1) < 1e-15
    for tr in (False, True):
        x = _vec(op.cols(), 3 + tr)
        yr = oop.apply(x, tr)
        e64 = _rel(op.apply(x, tr), yr)
        e32 = _rel(op.apply_prec(x, lt.TM_F32, tr), yr)
        print(f"  {op.kind()} L={size} tr={tr} fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
        assert e64 <= TOL64, e64
        assert e32 <= TOL32, e32


def test_axial_f32_io(lt, oracle):
    import torch
    for axialdir, kind in ((2, oracle.YAT), (0, oracle.Y2AT)):
        op = lt.TransferMatrix("A", 2, axialdir, 9, J=0.4, kappa=0.6, h=0.05)
        oop = oracle.OracleOp(kind, 9, 0, 0, 0.4, 0.6, 0.0, 0.05)
        n = op.cols()
        for tr in (0, 1):
            x32 = _vec(n, 5).astype(np.float32)
            xd = torch.from_numpy(x32).to("cuda")
            yd = torch.zeros(n, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            op.apply_dev(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F32, lt.TM_F32, tr)
            torch.cuda.synchronize()
            err = _rel(yd.cpu().numpy().astype(np.float64), oop.apply(x32.astype(np.float64), tr))
            print(f"  {op.kind()} FP32 io tr={tr} rel={err:.2e}", flush=True)
            assert err <= TOL32, err


@pytest.mark.parametrize("prec", [0, 1])
def test_axial_two_layers_per_step(lt, prec):
    """Y2AT advances two axial layers per application: lambda0(Y2AT)/nf equals
    (lambda0(YAT)/nf)^2; the YAT left vector is the Perron vector of YAT^T."""
    L = 10
    yat = lt.TransferMatrix("A", 2, 2, L, J=0.35, kappa=0.4)
    y2 = lt.TransferMatrix("A", 2, 0, L, J=0.35, kappa=0.4)
    r1 = lt.solve(yat, nev=1, tol=1e-12, want_left=True, want_vectors=True, precision=prec)
    r2 = lt.solve(y2, nev=1, tol=1e-12, precision=prec)
    l1 = r1.evals[0].real / yat.getnormfactor()
    l2 = r2.evals[0].real / y2.getnormfactor()
    print("YAT", l1, "Y2AT", l2, "left", r1.eval_left, flush=True)
    assert abs(l1 * l1 / l2 - 1) < 1e-11
    assert abs(r1.eval_left / r1.evals[0].real - 1) < 1e-11
    v = r1.vecl
    w = yat.apply(v, transpose=True)
    assert np.linalg.norm(w - r1.eval_left * v) / (abs(r1.eval_left) * np.linalg.norm(v)) < 1e-9


def test_axial_dense_small(lt, oracle):
    """Leading eigenvalue at L = 5 (1024 states) against dense LAPACK of the
    oracle matrix."""
    for axialdir, kind in ((2, oracle.YAT), (0, oracle.Y2AT)):
        op = lt.TransferMatrix("A", 2, axialdir, 5, J=0.5, kappa=0.3, h=0.05)
        r = lt.solve(op, nev=2, tol=1e-12)
        T = oracle.OracleOp(kind, 5, 0, 0, 0.5, 0.3, 0.0, 0.05).dense()
        w = np.linalg.eigvals(T)
        w = w[np.argsort(-np.abs(w))]
        print(op.kind(), r.evals, w[:3], flush=True)
        assert abs(r.evals[0] / w[0] - 1) < 1e-11


def test_axial_grad_unsupported(lt):
    op = lt.TransferMatrix("A", 2, 2, 4, J=0.5, kappa=0.3)
    with pytest.raises(lt.TMError) as e:
        op.grad_hf(np.ones(op.cols()), np.ones(op.cols()))
    assert e.value.status == 5


@pytest.mark.parametrize("kind,size,ext,k,k2,h", [("YAM", 4, 0, 0.35, 0.0, 0.0), ("YAM", 5, 2, 0.35, 0.0, 0.0),
                                                  ("YAM", 6, 0, 0.6, 0.0, 0.1), ("YAM", 8, 0, 0.3, 0.0, 0.0),
                                                  ("YAM", 9, 2, 0.3, 0.0, 0.0), ("YBM", 4, 0, 0.3, 0.2, 0.0),
                                                  ("YBM", 7, 0, 0.6, 0.4, 0.1), ("YBM", 8, 2, 0.2, 0.6, 0.0),
                                                  ("YAM", 10, 0, 0.45, 0.0, 0.0)])
def test_yam_ybm_apply(lt, oracle, kind, size, ext, k, k2, h):
    """Reduced axial operators: orbit tables bit-exact, T.x and T^T.x against
    the restated YAMMatVecProd2D(_T) (layertype 1 for YBM)."""
    model = "A" if kind == "YAM" else "B"
    op = lt.TransferMatrix(model, 2, 6, size, J=0.45, kappa=k, kappa2=k2, h=h,
                           neigs=4 if ext == 2 else 1)
    oop = oracle.OracleOp(getattr(oracle, kind), size, 0, ext, 0.45, k, k2, h)
    assert op.kind() == kind and not op.is_symmetric()
    reps, dups = op.orbits()
    np.testing.assert_array_equal(reps, oop.reps())
    np.testing.assert_array_equal(dups, oop.dups())
    assert abs(op.getnormfactor() / oop.normfactor - 1) < 1e-15
    for tr in (False, True):
        x = _vec(op.cols(), 7 + tr)
        yr = oop.apply(x, tr)
        e64 = _rel(op.apply(x, tr), yr)
        e32 = _rel(op.apply_prec(x, lt.TM_F32, tr), yr)
        print(f"  {kind} L={size} M={op.cols()} tr={tr} fp64 {e64:.2e} fp32 {e32:.2e}", flush=True)
        assert e64 <= TOL64, e64
        assert e32 <= TOL32, e32


@pytest.mark.parametrize("prec", [0, 1])
def test_yam_lambda0_equals_yat(lt, oracle, prec):
    L = 10
    yam = lt.TransferMatrix("A", 2, 6, L, J=0.35, kappa=0.4)
    yat = lt.TransferMatrix("A", 2, 2, L, J=0.35, kappa=0.4)
    r1 = lt.solve(yam, nev=1, tol=1e-12, want_left=True, want_vectors=True, precision=prec)
    r2 = lt.solve(yat, nev=1, tol=1e-12, precision=prec)
    print("YAM", yam.cols(), r1.evals, r1.eval_left, "YAT", r2.evals, flush=True)
    assert abs(r1.evals[0].real / r2.evals[0].real - 1) < 1e-11
    assert abs(r1.eval_left / r1.evals[0].real - 1) < 1e-11
    v = r1.vecl
    w = yam.apply(v, transpose=True)
    assert np.linalg.norm(w - r1.eval_left * v) / (abs(r1.eval_left) * np.linalg.norm(v)) < 1e-9
