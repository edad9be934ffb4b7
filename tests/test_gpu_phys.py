"""GPU parity of tm_op_moments (the X_phys.dat figures reduced on the device)
against the oracle's sequential restatement of the reference's moments() /
phys_extend() for every operator kind, on random positive vector pairs; and
at the benchmark width (reduced ANNNI L=28, 4,806,078 orbits) against the
binomial closed forms of a uniform weight (q_o = sqrt(d_o / 2^L)).
Tolerance 1e-12 relative (the device adds in a different order); the
figure count and the NaN entries of the 3D operators exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as m
    return m


# (label, TransferMatrix args, kwargs, oracle kind, oracle size1, size2, bc)
CASES = [
    ("XAT", ("A", 2, 1, 12), {}, "XAT", 12, 0, 0),
    ("XAT-open", ("A", 2, 9, 12), {}, "XAT", 12, 0, 1),
    ("XAT-c5", ("A", 2, 17, 11), {}, "XAT", 11, 0, 5),
    ("XAM2D", ("A", 2, 5, 16), {}, "XAM2D", 16, 0, 0),
    ("XAM2D-open", ("A", 2, 13, 14), {}, "XAM2D", 14, 0, 1),
    ("XAM2D-field", ("A", 2, 5, 14), {"h": 0.1}, "XAM2D", 14, 0, 0),
    ("XDT", ("D", 2, 1, 12), {}, "XDT", 12, 0, 0),
    ("XDT-c5", ("D", 2, 17, 12), {}, "XDT", 12, 0, 5),
    ("XDM", ("D", 2, 5, 14), {}, "XDM", 14, 0, 0),
    ("ZBT", ("B", 2, 3, 12), {"kappa2": 0.2}, "ZBT", 12, 0, 0),
    ("ZBM", ("B", 2, 7, 14), {"kappa2": 0.2}, "ZBM", 14, 0, 0),
    ("YAT", ("A", 2, 2, 7), {}, "YAT", 7, 0, 0),
    ("Y2AT", ("A", 2, 0, 5), {}, "Y2AT", 5, 0, 0),
    ("YAM", ("A", 2, 6, 7), {}, "YAM", 7, 0, 0),
    ("YBM", ("B", 2, 6, 7), {"kappa2": 0.2, "h": 0.1}, "YBM", 7, 0, 0),
    ("XAM3D", ("A", 3, 1, (3, 4)), {}, "XAM3D", 3, 4, 0),
    ("XAT3D", ("A", 3, 1, (3, 4)), {"flags": 1}, "XAT3D", 3, 4, 0),
]


@pytest.mark.parametrize("label,args,kw,okind,s1,s2,bc", CASES, ids=[c[0] for c in CASES])
def test_moments_vs_oracle(lt, oracle, label, args, kw, okind, s1, s2, bc):
    kw = dict(kw)
    J, k = 0.4, 0.3
    h = kw.pop("h", 0.0)
    k2 = kw.pop("kappa2", 0.0)
    op = lt.TransferMatrix(*args, J=J, kappa=k, kappa2=k2, h=h, **kw)
    assert op.kind() == okind
    oop = oracle.OracleOp(getattr(oracle, okind), s1, s2, bc, J, k, k2, h)
    n = op.cols()
    assert n == oop.n
    rng = np.random.default_rng(len(label))
    qr = rng.uniform(0.1, 1.0, n)
    ql = rng.uniform(0.1, 1.0, n)
    for pair in ((qr, None), (qr, ql)):
        g = op.moments(*pair)
        o = oop.moments(*pair)
        print(f"  {label} n={n} gpu={g} oracle={o}", flush=True)
        assert len(g) == len(o)
        nan = np.isnan(o)
        np.testing.assert_array_equal(np.isnan(g), nan)
        np.testing.assert_allclose(g[~nan], o[~nan], rtol=1e-12, atol=1e-14)


def test_moments_uniform_l28(lt):
    L = 28
    op = lt.TransferMatrix("A", 2, 5, L, J=0.4, kappa=0.5)
    _, dups = op.orbits()
    d = dups.astype(np.float64)
    assert d.sum() == 2.0**L
    f = op.moments(np.sqrt(d / d.sum()))
    print("  XAM L=28 uniform", f, flush=True)
    # flip-folded orbits: the even moments and node counts are the invariant ones
    np.testing.assert_allclose(f[[1, 3]], [1.0, 3.0 * L - 2.0], rtol=1e-12)
    np.testing.assert_allclose(f[4:], [L / 2, (L * L + L) / 4], rtol=1e-12)


def test_moments_reject_null(lt):
    import ctypes as C
    op = lt.TransferMatrix("A", 2, 1, 8)
    out = np.empty(8)
    n = C.c_int(0)
    st = lt._abi.lib().tm_op_moments(op._h, None, None, out.ctypes.data_as(C.POINTER(C.c_double)),
                                     C.byref(n))
    assert st != 0
