"""GPU parity at the benchmark configurations' own sizes.

  C2  reduced ANNNI L=28: the FP32 T.x (FP64 and FP32 vectors) against the
      oracle's FP64 T.x;
  C3  BNNNI/3NN diagonal TM (ZBT, TMat/tmat.cpp:480-662) at the scan widths
      14-22: T.x and T^T.x against the oracle (FP64 and FP32), the adjoint
      identity at L=22, lambda0(ZBT) = lambda0(ZBM reduced) at L=22, and the
      96-instance scan batch (mixed precision) against single FP64 solves;
  C4  DNNI (XDT, TMat/tmat.cpp:379-476) at L=22: T.x against the oracle and
      the dual-number dT/dtheta kernels against central differences of the
      oracle's T.x;
  C5  3D 5x5 layer: the reduced operator (M = 170,440, TMat/tmat3d.cpp:203-321)
      and the full 2^25 operator against the oracle, and lambda0 of the full
      operator (mixed, 20-vector basis) against the reduced one.
Tolerances: FP64 max|y - y_ref| <= 1e-13 max|y_ref|; FP32 <= 2e-5; eigenvalues
1e-12 relative."""
import numpy as np
import pytest

from tests.test_gpu_batch import scan_points

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 2e-5


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _parity(op, oop, seed, transpose=False, f32=True):
    x = np.random.default_rng(seed).uniform(-1, 1, op.cols())
    yr = oop.apply(x, transpose)
    e64 = _rel(op.apply(x, transpose), yr)
    print(f"  {op.kind()} n={op.cols()} tr={transpose} FP64 rel {e64:.2e}", flush=True)
    assert e64 <= TOL64
    if f32:
        e32 = _rel(op.apply_prec(x, 1, transpose), yr)
        print(f"  {op.kind()} n={op.cols()} tr={transpose} FP32 rel {e32:.2e}", flush=True)
        assert e32 <= TOL32


# ---- C3: ZBT at the scan widths ----------------------------------------
@pytest.mark.parametrize("L", [14, 16, 18, 20, 22])
@pytest.mark.parametrize("k1,k2", [(0.5, 0.0), (0.8, 0.5)])
def test_zbt_scan_widths(lt, oracle, L, k1, k2):
    op = lt.TransferMatrix("B", 2, 3, L, J=0.6, kappa=k1, kappa2=k2)
    oop = oracle.OracleOp(oracle.ZBT, L, 0, 0, 0.6, k1, k2, 0.0)
    _parity(op, oop, L, False)
    _parity(op, oop, L + 1, True)


def test_zbt_l22_adjoint_and_reduced(lt):
    op = lt.TransferMatrix("B", 2, 3, 22, J=0.5, kappa=0.5, kappa2=0.5)
    rng = np.random.default_rng(9)
    x, y = rng.uniform(-1, 1, op.cols()), rng.uniform(-1, 1, op.cols())
    a, b = float(y @ op.apply(x)), float(op.apply(y, True) @ x)
    print("adjoint", a, b, flush=True)
    assert abs(a - b) <= 1e-12 * abs(a)
    red = lt.TransferMatrix("B", 2, 7, 22, J=0.5, kappa=0.5, kappa2=0.5, neigs=1)
    rf = lt.solve(op, nev=2, tol=1e-12)
    rr = lt.solve(red, nev=1, tol=1e-12)
    print("lambda0 ZBT", rf.evals[0], "ZBM", rr.evals[0], flush=True)
    assert abs(rr.evals[0].real / rf.evals[0].real - 1) < 1e-12


def test_scan_batch_real_widths(lt):
    """The 96-instance scan (instance i -> grid point i mod 54, width
    (14, 16, 18, 20, 22)[i mod 5]) as one mixed-precision batch against
    single FP64 solves, lambda0 to 1e-12."""
    grid = scan_points()
    inst = [(grid[i % 54], (14, 16, 18, 20, 22)[i % 5]) for i in range(96)]
    ops = [lt.TransferMatrix("B", 2, 3, L, J=bJ, kappa=k1, kappa2=k2) for (bJ, k1, k2), L in inst]
    ref = [lt.solve(op, nev=2, tol=1e-12) for op in ops]
    bat = lt.solve_batch(ops, nev=2, ncv=12, tol=1e-12, precision=lt.TM_PREC_MIXED, nworkers=4)
    worst = 0.0
    for i, (a, b) in enumerate(zip(ref, bat)):
        e = abs(b.evals[0].real / a.evals[0].real - 1)
        worst = max(worst, e)
        assert e < 1e-12, (i, inst[i], a.evals[0], b.evals[0])
    print("96 instances, worst lambda0 rel diff", worst, flush=True)


# ---- C4: DNNI at L=22 ---------------------------------------------------------
def test_xdt_l22(lt, oracle):
    op = lt.TransferMatrix("D", 2, 1, 22, J=0.6, kappa=0.7, h=0.1)
    oop = oracle.OracleOp(oracle.XDT, 22, 0, 0, 0.6, 0.7, 0.0, 0.1)
    _parity(op, oop, 3)


def test_xdt_l22_dual_kernels_vs_oracle_fd(lt, oracle):
    """The dual-number kernels' <r, dT/dtheta r> / <r, r> against central
    differences of the oracle's T.x (parameters +- 1e-5), for betaJ,
    kappa1 and h (kappa2 does not enter the DNNI operator)."""
    th = [0.6, 0.7, 0.0, 0.1]
    op = lt.TransferMatrix("D", 2, 1, 22, J=th[0], kappa=th[1], h=th[3])
    r = np.random.default_rng(4).uniform(0.5, 1.0, op.cols())
    dlam, _ = op.grad_hf(r, r)
    rr = float(r @ r)
    for i in (0, 1, 3):
        tp, tm = list(th), list(th)
        tp[i] += 1e-5
        tm[i] -= 1e-5
        yp = oracle.OracleOp(oracle.XDT, 22, 0, 0, tp[0], tp[1], tp[2], tp[3]).apply(r)
        ym = oracle.OracleOp(oracle.XDT, 22, 0, 0, tm[0], tm[1], tm[2], tm[3]).apply(r)
        fd = float(r @ (yp - ym)) / 2e-5 / rr
        print(i, dlam[i], fd, flush=True)
        assert abs(dlam[i] - fd) <= 1e-7 * max(1.0, abs(fd))
    assert dlam[2] == 0.0


# ---- C2: the FP32 reduced T.x at L=28 ----------------------------------
def test_xam_l28_fp32(lt, oracle):
    import torch
    op = lt.TransferMatrix("A", 2, 5, 28, J=0.4, kappa=0.5, neigs=4)
    oop = oracle.OracleOp(oracle.XAM2D, 28, 0, 2, 0.4, 0.5, 0.0, 0.0)
    x = np.random.default_rng(11).uniform(0, 1, op.cols())
    yr = oop.apply(x)
    e = _rel(op.apply_prec(x, lt.TM_F32), yr)
    print("L=28 FP32 compute, FP64 vectors", e, flush=True)
    assert e <= TOL32
    # FP32 vectors in and out on the device (the mixed solve's basis columns)
    n = op.cols()
    xd = torch.from_numpy(x.astype(np.float32)).to("cuda")
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    op.apply_dev(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F32, lt.TM_F32)
    torch.cuda.synchronize()
    e2 = _rel(yd.cpu().numpy().astype(np.float64), oop.apply(x.astype(np.float32).astype(np.float64)))
    print("L=28 FP32 compute, FP32 vectors", e2, flush=True)
    assert e2 <= TOL32


# ---- C5: 3D 5x5 ---------------------------------------------------------------
def test_3d_5x5_reduced_parity(lt, oracle):
    op = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4)
    oop = oracle.OracleOp(oracle.XAM3D, 5, 5, 0, 0.25, 0.4, 0.0, 0.0)
    reps, dups = op.orbits()
    np.testing.assert_array_equal(reps, oop.reps())
    np.testing.assert_array_equal(dups, oop.dups())
    _parity(op, oop, 5)


def test_3d_5x5_full_parity(lt, oracle):
    op = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4, flags=lt.TM_FLAG_FULLSPACE)
    oop = oracle.OracleOp(oracle.XAT3D, 5, 5, 0, 0.25, 0.4, 0.0, 0.0)
    assert op.cols() == 1 << 25
    _parity(op, oop, 6)


def test_3d_5x5_full_equals_reduced(lt):
    full = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4, flags=lt.TM_FLAG_FULLSPACE)
    red = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4)
    rf = lt.solve(full, nev=2, ncv=20, precision=lt.TM_PREC_MIXED)
    rr = lt.solve(red, nev=2, tol=1e-12)
    print("lambda0 full", rf.evals[0], "reduced", rr.evals[0], flush=True)
    assert abs(rf.evals[0] / rr.evals[0] - 1) < 1e-12
