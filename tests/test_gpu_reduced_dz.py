"""GPU parity of the reduced DNNI (XDM, D axialdir 5/6) and reduced
BNNNI/3NN diagonal (ZBM, B axialdir 7) operators.

The device path expands a reduced vector over the full layer, runs the
XDT / ZBT gate circuit and gathers at the orbit representatives.  The
oracle restates the reference's own algorithms
(XDMMatVecProd2D::perform_op, ZBMMatVecProd2D::perform_op_internal), so
agreement checks the one against the other.  Tolerances as for every
operator: FP64 1e-13 and FP32 2e-5 of max|y|; orbit tables bit-exact;
eigenvalues 1e-12 relative.
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


def _check(lt, op, oop, transpose=False):
    reps, dups = op.orbits()
    np.testing.assert_array_equal(reps, oop.reps())
    np.testing.assert_array_equal(dups, oop.dups())
    assert abs(op.getnormfactor() / oop.normfactor - 1) < 1e-15
    for sd in (1, 2):
        x = _vec(op.cols(), sd)
        e64 = _rel(op.apply(x, transpose), oop.apply(x, transpose))
        e32 = _rel(op.apply_prec(x, lt.TM_F32, transpose), oop.apply(x, transpose))
        print(f"  {op.kind()} M={op.cols()} tr={transpose} seed={sd} fp64 {e64:.2e} "
              f"fp32 {e32:.2e}", flush=True)
        assert e64 <= TOL64, e64
        assert e32 <= TOL32, e32


@pytest.mark.parametrize("size,neigs,h,axialdir", [(8, 1, 0.0, 5), (9, 1, 0.0, 6), (12, 4, 0.0, 5),
                                                   (14, 1, 0.1, 5), (16, 1, 0.0, 5),
                                                   (17, 4, 0.0, 6)])
def test_xdm_apply(lt, oracle, size, neigs, h, axialdir):
    ext = 2 if neigs > 1 else 0
    op = lt.TransferMatrix("D", 2, axialdir, size, J=0.45, kappa=0.35, h=h, neigs=neigs)
    assert op.kind() == "XDM" and op.is_symmetric()
    oop = oracle.OracleOp(oracle.XDM, size, 0, ext, 0.45, 0.35, 0.0, h)
    _check(lt, op, oop)


@pytest.mark.parametrize("size,neigs,k,k2,h", [(8, 1, 0.5, 0.3, 0.0), (10, 4, 0.2, 0.1, 0.0),
                                               (12, 1, 0.8, 0.4, 0.1), (14, 1, 0.3, 0.7, 0.0),
                                               (16, 4, 0.6, 0.0, 0.0)])
def test_zbm_apply_and_transpose(lt, oracle, size, neigs, k, k2, h):
    ext = 2 if neigs > 1 else 0
    op = lt.TransferMatrix("B", 2, 7, size, J=0.4, kappa=k, kappa2=k2, h=h, neigs=neigs)
    assert op.kind() == "ZBM" and not op.is_symmetric()
    oop = oracle.OracleOp(oracle.ZBM, size, 0, ext, 0.4, k, k2, h)
    _check(lt, op, oop)
    _check(lt, op, oop, transpose=True)


def test_xdm_zbm_f32_io(lt, oracle):
    """FP32 vectors in and out of the device apply (the FP32 phase of a mixed
    solve), against the oracle on the same rounded input."""
    import torch
    cases = [(lt.TransferMatrix("D", 2, 5, 14, J=0.4, kappa=0.3, h=0.1),
              oracle.OracleOp(oracle.XDM, 14, 0, 0, 0.4, 0.3, 0.0, 0.1), (0,)),
             (lt.TransferMatrix("B", 2, 7, 14, J=0.3, kappa=0.5, kappa2=0.2),
              oracle.OracleOp(oracle.ZBM, 14, 0, 0, 0.3, 0.5, 0.2, 0.0), (0, 1))]
    for op, oop, trs in cases:
        n = op.cols()
        for tr in trs:
            x32 = _vec(n, 11).astype(np.float32)
            xd = torch.from_numpy(x32).to("cuda")
            yd = torch.zeros(n, dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            op.apply_dev(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F32, lt.TM_F32, tr)
            torch.cuda.synchronize()
            err = _rel(yd.cpu().numpy().astype(np.float64), oop.apply(x32.astype(np.float64), tr))
            print(f"  {op.kind()} FP32 io tr={tr} rel={err:.2e}", flush=True)
            assert err <= TOL32, err


@pytest.mark.parametrize("prec", [0, 1])
def test_xdm_lambda0_equals_xdt(lt, prec):
    red = lt.TransferMatrix("D", 2, 5, 18, J=0.5, kappa=0.6, h=0.0)
    full = lt.TransferMatrix("D", 2, 1, 18, J=0.5, kappa=0.6, h=0.0)
    rr = lt.solve(red, nev=1, precision=prec)
    rf = lt.solve(full, nev=1)
    print("XDM", red.cols(), rr.evals, "XDT", rf.evals, flush=True)
    assert red.cols() == 3914  # D_18 x flip orbits (Burnside)
    assert abs(rr.evals[0] / rf.evals[0] - 1) < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
def test_zbm_lambda0_and_left_vector(lt, oracle, prec):
    red = lt.TransferMatrix("B", 2, 7, 16, J=0.4, kappa=0.6, kappa2=0.3)
    full = lt.TransferMatrix("B", 2, 3, 16, J=0.4, kappa=0.6, kappa2=0.3)
    rr = lt.solve(red, nev=2, want_left=True, want_vectors=True, precision=prec)
    rf = lt.solve(full, nev=2)
    print("ZBM", red.cols(), rr.evals, rr.eval_left, "ZBT", rf.evals, flush=True)
    assert abs(rr.evals[0] / rf.evals[0] - 1) < 1e-12
    assert abs(rr.eval_left / rf.evals[0].real - 1) < 1e-10
    # the left vector is the Perron vector of T_red^T (oracle transpose)
    oop = oracle.OracleOp(oracle.ZBM, 16, 0, 0, 0.4, 0.6, 0.3, 0.0)
    v = rr.vecl
    res = np.linalg.norm(oop.apply(v, True) - rr.eval_left * v) / abs(rr.eval_left)
    assert res < 1e-8


def test_xdm_zbm_gradient_equals_full(lt):
    """Hellmann-Feynman derivatives of lambda0 on the reduced operators equal
    those on the full ones (same lambda0 as a function of the parameters)."""
    for red_args, full_args, kw in ((("D", 2, 5, 14), ("D", 2, 1, 14),
                                     dict(J=0.5, kappa=0.6, h=0.0)),
                                    (("B", 2, 7, 14), ("B", 2, 3, 14),
                                     dict(J=0.4, kappa=0.6, kappa2=0.3))):
        out = []
        for args in (red_args, full_args):
            op = lt.TransferMatrix(*args, **kw)
            sym = op.is_symmetric()
            r = lt.solve(op, nev=1, tol=1e-13, want_left=not sym, want_vectors=True)
            r_vec = r.vecs[:, 0] if sym else r.vecs[:, 0].real
            l_vec = r_vec if sym else r.vecl
            dlam, dlnnf = op.grad_hf(l_vec, r_vec)
            out.append((r.evals[0].real, dlam, dlnnf))
            print(op.kind(), r.evals[0], dlam, flush=True)
        (l0r, gr, nr), (l0f, gf, nf) = out
        assert abs(l0r / l0f - 1) < 1e-12
        np.testing.assert_allclose(gr, gf, rtol=1e-9, atol=1e-12 * np.max(np.abs(gf)))
        np.testing.assert_allclose(nr, nf, rtol=1e-14)
