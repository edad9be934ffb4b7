"""GPU parity of every operator path against the CPU oracle.

Each test builds the same operator through the C ABI (HIP kernels on the
MI355X) and in the oracle (a restatement of the reference algorithm), applies
both to the same seeded vectors, and compares.  Tolerances:
  FP64 compute:  max|y_gpu - y_ref| <= 1e-13 * max|y_ref|
  FP32 compute:  max|y_gpu - y_ref| <= 2e-5  * max|y_ref|   (single-precision
                 accumulation over L butterfly stages)
Orbit tables (representatives, orbit sizes) must be bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 2e-5


@pytest.fixture(scope="module")
def L():
    import latttmat_amd as lt
    return lt


def _vec(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _check_apply(lt, oracle, op, oop, transpose=False, seeds=(1, 2)):
    assert op.cols() == oop.n
    for sd in seeds:
        x = _vec(op.cols(), sd)
        y = op.apply(x, transpose)
        yr = oop.apply(x, transpose)
        err = _rel(y, yr)
        print(f"  {op.kind()} n={op.cols()} tr={transpose} seed={sd} rel={err:.2e}", flush=True)
        assert err <= TOL64, err


def _check_f32(lt, oracle, op, oop, transpose=False):
    n = op.cols()
    x = _vec(n, 7)
    yr = oop.apply(x, transpose)
    y = op.apply_prec(x, lt.TM_F32, transpose)
    err = _rel(y, yr)
    print(f"  {op.kind()} FP32 n={n} rel={err:.2e}", flush=True)
    assert err <= TOL32, err


@pytest.mark.parametrize("size", [4, 8, 12, 13, 14, 16, 18])
def test_xat_full(L, oracle, size):
    op = L.TransferMatrix("A", 2, 1, size, J=0.5, kappa=0.6, h=0.0)
    oop = oracle.OracleOp(oracle.XAT, size, 0, 0, 0.5, 0.6, 0.0, 0.0)
    assert abs(op.getnormfactor() / oop.normfactor - 1) < 1e-15
    _check_apply(L, oracle, op, oop)
    _check_f32(L, oracle, op, oop)


@pytest.mark.parametrize("axialdir,bc", [(9, 1), (17, 5), (18, 9), (19, 13), (20, 17)])
def test_xat_boundaries(L, oracle, axialdir, bc):
    op = L.TransferMatrix("A", 2, axialdir, 10, J=0.45, kappa=0.3, h=0.2)
    oop = oracle.OracleOp(oracle.XAT, 10, 0, bc, 0.45, 0.3, 0.0, 0.2)
    _check_apply(L, oracle, op, oop)


@pytest.mark.parametrize("size", [4, 6, 10, 14, 16])
@pytest.mark.parametrize("h", [0.0, 0.1])
def test_xdt(L, oracle, size, h):
    op = L.TransferMatrix("D", 2, 1, size, J=0.6, kappa=0.7, h=h)
    oop = oracle.OracleOp(oracle.XDT, size, 0, 0, 0.6, 0.7, 0.0, h)
    _check_apply(L, oracle, op, oop)
    if size >= 10:
        _check_f32(L, oracle, op, oop)


@pytest.mark.parametrize("axialdir,bc", [(17, 5), (18, 9)])
def test_xdt_confined(L, oracle, axialdir, bc):
    op = L.TransferMatrix("D", 2, axialdir, 9, J=0.4, kappa=0.3, h=0.05)
    oop = oracle.OracleOp(oracle.XDT, 9, 0, bc, 0.4, 0.3, 0.0, 0.05)
    _check_apply(L, oracle, op, oop)


@pytest.mark.parametrize("size", [4, 6, 10, 14, 16])
def test_zbt(L, oracle, size):
    op = L.TransferMatrix("B", 2, 3, size, J=0.5, kappa=0.5, kappa2=0.3, h=0.05)
    oop = oracle.OracleOp(oracle.ZBT, size, 0, 0, 0.5, 0.5, 0.3, 0.05)
    assert not op.is_symmetric()
    _check_apply(L, oracle, op, oop, transpose=False)
    _check_apply(L, oracle, op, oop, transpose=True)
    if size >= 10:
        _check_f32(L, oracle, op, oop)


@pytest.mark.parametrize("size", [8, 10, 12, 14, 16, 18, 20])
@pytest.mark.parametrize("neigs", [1, 4])
def test_xam2d_periodic(L, oracle, size, neigs):
    ext = 2 if neigs > 1 else 0
    for flags in (0, L.TM_FLAG_GENERIC_REDUCED):
        op = L.TransferMatrix("A", 2, 5, size, J=0.4, kappa=0.5, h=0.0, neigs=neigs, flags=flags)
        oop = oracle.OracleOp(oracle.XAM2D, size, 0, ext, 0.4, 0.5, 0.0, 0.0)
        reps, dups = op.orbits()
        np.testing.assert_array_equal(reps, oop.reps())
        np.testing.assert_array_equal(dups, oop.dups())
        _check_apply(L, oracle, op, oop)
        _check_f32(L, oracle, op, oop)


@pytest.mark.parametrize("size", [9, 12, 15])
def test_xam2d_field_and_open(L, oracle, size):
    # h != 0 disables the flip symmetry; axialdir 13 is the open boundary
    for axialdir, ext in ((5, 0), (13, 1)):
        op = L.TransferMatrix("A", 2, axialdir, size, J=0.35, kappa=0.2, h=0.15)
        oop = oracle.OracleOp(oracle.XAM2D, size, 0, ext, 0.35, 0.2, 0.0, 0.15)
        reps, dups = op.orbits()
        np.testing.assert_array_equal(reps, oop.reps())
        np.testing.assert_array_equal(dups, oop.dups())
        _check_apply(L, oracle, op, oop)


@pytest.mark.parametrize("s1,s2,kappa,h", [(2, 3, 0.4, 0.0), (3, 3, 0.0, 0.0), (3, 4, 0.4, 0.0),
                                           (4, 4, 0.0, 0.0), (3, 4, 0.4, 0.1), (4, 4, 0.3, 0.0)])
def test_xam3d(L, oracle, s1, s2, kappa, h):
    op = L.TransferMatrix("A", 3, 1, (s1, s2), J=0.25, kappa=kappa, h=h)
    oop = oracle.OracleOp(oracle.XAM3D, s1, s2, 0, 0.25, kappa, 0.0, h)
    reps, dups = op.orbits()
    np.testing.assert_array_equal(reps, oop.reps())
    np.testing.assert_array_equal(dups, oop.dups())
    _check_apply(L, oracle, op, oop)


@pytest.mark.parametrize("s1,s2", [(2, 3), (3, 3), (3, 4), (4, 4), (4, 5)])
def test_xat3d_full(L, oracle, s1, s2):
    op = L.TransferMatrix("A", 3, 1, (s1, s2), J=0.25, kappa=0.4, h=0.05,
                          flags=L.TM_FLAG_FULLSPACE)
    oop = oracle.OracleOp(oracle.XAT3D, s1, s2, 0, 0.25, 0.4, 0.0, 0.05)
    _check_apply(L, oracle, op, oop)
    _check_f32(L, oracle, op, oop)


def _check_f32_io(lt, op, oop, transpose=False):
    """FP32 vectors in and out of the device apply (what the FP32 phase of a
    mixed solve hands the operator), against the oracle on the same rounded
    input."""
    import torch
    n = op.cols()
    x32 = _vec(n, 11).astype(np.float32)
    xd = torch.from_numpy(x32).to("cuda")
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    op.apply_dev(xd.data_ptr(), yd.data_ptr(), n, 1, lt.TM_F32, lt.TM_F32, transpose)
    torch.cuda.synchronize()
    y = yd.cpu().numpy().astype(np.float64)
    yr = oop.apply(x32.astype(np.float64), transpose)
    err = _rel(y, yr)
    print(f"  {op.kind()} FP32 io n={n} tr={transpose} rel={err:.2e}", flush=True)
    assert err <= TOL32, err


@pytest.mark.parametrize("case", ["xat14", "xat13_bc", "xdt12", "zbt12", "xam16", "xam16_generic",
                                  "xam3d_34", "xat3d_34"])
def test_apply_dev_f32_io(L, oracle, case):
    O = oracle
    if case == "xat14":
        op = L.TransferMatrix("A", 2, 1, 14, J=0.5, kappa=0.6)
        oop = O.OracleOp(O.XAT, 14, 0, 0, 0.5, 0.6, 0.0, 0.0)
    elif case == "xat13_bc":
        op = L.TransferMatrix("A", 2, 9, 13, J=0.45, kappa=0.3, h=0.2)
        oop = O.OracleOp(O.XAT, 13, 0, 1, 0.45, 0.3, 0.0, 0.2)
    elif case == "xdt12":
        op = L.TransferMatrix("D", 2, 1, 12, J=0.4, kappa=0.3, h=0.1)
        oop = O.OracleOp(O.XDT, 12, 0, 0, 0.4, 0.3, 0.0, 0.1)
    elif case == "zbt12":
        op = L.TransferMatrix("B", 2, 3, 12, J=0.3, kappa=0.5, kappa2=0.2)
        oop = O.OracleOp(O.ZBT, 12, 0, 0, 0.3, 0.5, 0.2, 0.0)
    elif case in ("xam16", "xam16_generic"):
        flags = L.TM_FLAG_GENERIC_REDUCED if case.endswith("generic") else 0
        op = L.TransferMatrix("A", 2, 5, 16, J=0.4, kappa=0.5, neigs=4, flags=flags)
        oop = O.OracleOp(O.XAM2D, 16, 0, 2, 0.4, 0.5, 0.0, 0.0)
    elif case == "xam3d_34":
        op = L.TransferMatrix("A", 3, 1, (3, 4), J=0.25, kappa=0.4)
        oop = O.OracleOp(O.XAM3D, 3, 4, 0, 0.25, 0.4, 0.0, 0.0)
    else:
        op = L.TransferMatrix("A", 3, 1, (3, 4), J=0.25, kappa=0.4, h=0.05,
                              flags=L.TM_FLAG_FULLSPACE)
        oop = O.OracleOp(O.XAT3D, 3, 4, 0, 0.25, 0.4, 0.0, 0.05)
    _check_f32_io(L, op, oop)
    if case == "zbt12":
        _check_f32_io(L, op, oop, transpose=True)
