"""Hellmann-Feynman derivatives (tm_grad_hf, dual-number operator kernels)
against central finite differences of the leading eigenvalue: dense LAPACK
eigenvalues of the oracle's matrices at small sizes, and GPU eigensolves for
the DNNI L=22 workload (f = -ln(lambda0)/L, step 1e-5, FP64).  The reference
has no derivatives, so these are checked against the mathematics only."""
import numpy as np
import pytest
import torch

from tests.hf_autograd import free_energy

pytestmark = pytest.mark.gpu

PARAMS = ("betaJ", "kappa", "kappa2", "h")


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def dense_lambda0(oracle, kind, s1, s2, bc, th):
    T = oracle.OracleOp(kind, s1, s2, bc, *th).dense()
    if kind in (oracle.ZBT, oracle.ZBM):
        w = np.linalg.eigvals(T)
    else:
        w = np.linalg.eigvalsh(0.5 * (T + T.T))
    return float(np.real(w[np.argmax(np.abs(w))]))


def fd_dense(oracle, kind, s1, s2, bc, th, i, step=1e-5):
    tp, tm = list(th), list(th)
    tp[i] += step
    tm[i] -= step
    return (dense_lambda0(oracle, kind, s1, s2, bc, tp)
            - dense_lambda0(oracle, kind, s1, s2, bc, tm)) / (2 * step)


CASES = [
    # (model, dim, axialdir, size, oracle kind, s1, s2, oracle bc, (bJ, k, k2, h))
    ("A", 2, 1, 10, "XAT", 10, 0, 0, (0.5, 0.6, 0.0, 0.1)),
    ("A", 2, 9, 9, "XAT", 9, 0, 1, (0.45, 0.3, 0.0, 0.2)),
    ("A", 2, 17, 9, "XAT", 9, 0, 5, (0.45, 0.3, 0.0, 0.0)),
    ("D", 2, 1, 10, "XDT", 10, 0, 0, (0.6, 0.7, 0.0, 0.1)),
    ("D", 2, 18, 9, "XDT", 9, 0, 9, (0.6, 0.3, 0.0, 0.1)),
    ("B", 2, 3, 8, "ZBT", 8, 0, 0, (0.5, 0.55, 0.3, 0.1)),
    ("B", 2, 3, 8, "ZBT", 8, 0, 0, (0.4, 0.2, 0.6, 0.0)),
    ("A", 2, 5, 12, "XAM2D", 12, 0, 0, (0.4, 0.45, 0.0, 0.0)),
    ("A", 2, 13, 12, "XAM2D", 12, 0, 1, (0.4, 0.3, 0.0, 0.15)),
    ("A", 3, 1, (3, 4), "XAM3D", 3, 4, 0, (0.25, 0.4, 0.0, 0.0)),
    ("D", 2, 5, 10, "XDM", 10, 0, 0, (0.6, 0.7, 0.0, 0.1)),
    ("D", 2, 6, 10, "XDM", 10, 0, 0, (0.5, 0.3, 0.0, 0.0)),
    ("B", 2, 7, 8, "ZBM", 8, 0, 0, (0.5, 0.55, 0.3, 0.1)),
    ("B", 2, 7, 10, "ZBM", 10, 0, 0, (0.4, 0.2, 0.6, 0.0)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[4]}-{c[2]}-{c[3]}")
def test_hf_vs_dense_fd(lt, oracle, case):
    model, dim, ax, size, okind, s1, s2, obc, th = case
    print("case", case, flush=True)
    op = lt.TransferMatrix(model, dim, ax, size, J=th[0], kappa=th[1], kappa2=th[2], h=th[3])
    sym = op.is_symmetric()
    r = lt.solve(op, nev=1, tol=1e-12, want_vectors=True, want_left=not sym)
    lam = float(np.real(r.evals[0]))
    right = np.real(r.vecs[:, 0])
    left = right if sym else r.vecl
    dlam, dlnnf = op.grad_hf(left, right)
    # finite differences stay inside one normfactor branch (kappa != 0.5, ...)
    kind = getattr(oracle, okind)
    for i in range(4):
        fd = fd_dense(oracle, kind, s1, s2, obc, th, i)
        print(PARAMS[i], dlam[i], fd, flush=True)
        assert abs(dlam[i] - fd) <= 1e-7 * abs(lam) + 1e-7 * abs(fd), (PARAMS[i], dlam[i], fd)
    # d ln(nf) against the normfactor's own finite difference
    for i in range(4):
        tp, tm = list(th), list(th)
        tp[i] += 1e-6
        tm[i] -= 1e-6
        nfp = oracle.OracleOp(kind, s1, s2, obc, *tp).normfactor
        nfm = oracle.OracleOp(kind, s1, s2, obc, *tm).normfactor
        assert abs(dlnnf[i] - (np.log(nfp) - np.log(nfm)) / 2e-6) < 1e-6


def test_dnni_l22_free_energy_gradient(lt):
    """DNNI, axialdir 1, L=22 (4,194,304 states), betaJ=0.6, kappa1=0.7,
    h=0.1: grad of f = -ln(lambda0)/L by autograd (Hellmann-Feynman backward)
    against central differences with step 1e-5."""
    spec = dict(model="D", axialdir=1, size=22, layer=22)
    theta = torch.tensor([0.6, 0.7, 0.0, 0.1], dtype=torch.float64, requires_grad=True)
    f = free_energy(theta, spec)
    f.backward()
    g = theta.grad.numpy()
    print("f", float(f.detach()), "grad", g, flush=True)
    for i in (0, 1, 3):
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[i] += 1e-5
        tm[i] -= 1e-5
        fd = (float(free_energy(tp, spec)) - float(free_energy(tm, spec))) / 2e-5
        print(PARAMS[i], g[i], fd, flush=True)
        assert abs(g[i] - fd) <= 1e-7 * max(1.0, abs(fd))
    assert g[2] == 0.0  # kappa2 does not enter the DNNI operator
