"""GPU parity of the register-resident Kronecker passes (the default plan of
the full-space XAT / XAT3D operators from L = 14 bits with FP64 and 15 with
FP32 arithmetic) against the CPU oracle.  The operators are made with
TM_FLAG_NO_AUTOTUNE so that the default plan runs, not a tuned alternative.
Sizes cover every plan shape: the first pass alone, a second pass of 1 to 11
bits (one, two and three register rounds, 128- and 64-byte runs) and a third
pass; boundary codes and the field cover every diagonal kind the kernel is
instantiated for.  Tolerances as in test_gpu_ops: FP64 1e-13, FP32 2e-5
(max-relative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 2e-5


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd
    return latttmat_amd


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _both(lt, op, oop, seed, fp32=True):
    x = np.random.default_rng(seed).uniform(-1.0, 1.0, op.cols())
    yr = oop.apply(x)
    e64 = _rel(op.apply(x), yr)
    print(f"  {op.kind()} n={op.cols()} fp64 rel={e64:.2e}", flush=True)
    assert e64 <= TOL64, e64
    if fp32:
        e32 = _rel(op.apply_prec(x, lt.TM_F32), yr)
        print(f"  {op.kind()} n={op.cols()} fp32 rel={e32:.2e}", flush=True)
        assert e32 <= TOL32, e32


@pytest.mark.parametrize("size", [14, 15, 16, 19, 21, 22, 24])
def test_xat_rpass(lt, oracle, size):
    op = lt.TransferMatrix("A", 2, 1, size, J=0.45, kappa=0.55, h=0.07,
                           flags=lt.TM_FLAG_NO_AUTOTUNE)
    oop = oracle.OracleOp(oracle.XAT, size, 0, 0, 0.45, 0.55, 0.0, 0.07)
    _both(lt, op, oop, size)


@pytest.mark.parametrize("size", [25, 26])
def test_xat_rpass_large(lt, oracle, size):
    op = lt.TransferMatrix("A", 2, 1, size, J=0.3, kappa=0.2, h=0.0,
                           flags=lt.TM_FLAG_NO_AUTOTUNE)
    oop = oracle.OracleOp(oracle.XAT, size, 0, 0, 0.3, 0.2, 0.0, 0.0)
    _both(lt, op, oop, 3)


@pytest.mark.parametrize("axialdir,bc", [(9, 1), (17, 5), (18, 9), (19, 13), (20, 17)])
def test_xat_rpass_boundaries(lt, oracle, axialdir, bc):
    op = lt.TransferMatrix("A", 2, axialdir, 17, J=0.45, kappa=0.3, h=0.2,
                           flags=lt.TM_FLAG_NO_AUTOTUNE)
    oop = oracle.OracleOp(oracle.XAT, 17, 0, bc, 0.45, 0.3, 0.0, 0.2)
    _both(lt, op, oop, bc)


def test_xat3d_rpass(lt, oracle):
    op = lt.TransferMatrix("A", 3, 1, (4, 5), J=0.25, kappa=0.4, h=0.05,
                           flags=lt.TM_FLAG_FULLSPACE | lt.TM_FLAG_NO_AUTOTUNE)
    oop = oracle.OracleOp(oracle.XAT3D, 4, 5, 0, 0.25, 0.4, 0.0, 0.05)
    _both(lt, op, oop, 5)


def test_xat_rpass_f32_io(lt, oracle):
    """FP32 vectors in and out (the mixed solve's FP32 phase)."""
    import torch
    n = 1 << 18
    op = lt.TransferMatrix("A", 2, 1, 18, J=0.5, kappa=0.6, h=0.0, flags=lt.TM_FLAG_NO_AUTOTUNE)
    oop = oracle.OracleOp(oracle.XAT, 18, 0, 0, 0.5, 0.6, 0.0, 0.0)
    x32 = np.random.default_rng(4).uniform(-1.0, 1.0, n).astype(np.float32)
    xd = torch.from_numpy(x32).to("cuda")
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    op.apply_dev(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F32, lt.TM_F32)
    torch.cuda.synchronize()
    err = _rel(yd.cpu().numpy().astype(np.float64), oop.apply(x32.astype(np.float64)))
    print(f"  XAT L=18 FP32 io rel={err:.2e}", flush=True)
    assert err <= TOL32, err


def test_xat_rpass_solve(lt):
    """lambda0 of the default (register-resident) plan equals the one of the
    tuned operator and the symmetric solve converges."""
    a = lt.TransferMatrix("A", 2, 1, 20, J=0.44, kappa=0.2, flags=lt.TM_FLAG_NO_AUTOTUNE)
    r = lt.solve(a, nev=2)
    b = lt.TransferMatrix("A", 2, 1, 20, J=0.44, kappa=0.2)
    r2 = lt.solve(b, nev=2)
    assert abs(r.evals[0] / r2.evals[0] - 1) < 1e-12
