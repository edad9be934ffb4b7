"""tm_solve_batch: many operators of different kinds and sizes solved
concurrently (one stream per operator, several host threads) give exactly
the results of solving them one by one, and the BNNNI parameter-scan rule
(instance i -> grid point i mod 54, width (14,16,18,20,22)[i mod 5]) is
reproduced at small widths against dense eigenvalues."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


def scan_points():
    bJ = [round(0.2 + 0.1 * i, 10) for i in range(9)]
    return list(itertools.product(bJ, (0.2, 0.5, 0.8), (0.0, 0.5)))


def test_batch_equals_sequential(lt):
    ops = []
    for i, L in enumerate((8, 10, 12, 14, 10, 8)):
        ops.append(lt.TransferMatrix("B", 2, 3, L, J=0.3 + 0.05 * i, kappa=0.4, kappa2=0.3))
    ops.append(lt.TransferMatrix("A", 2, 1, 12, J=0.5, kappa=0.6))
    ops.append(lt.TransferMatrix("D", 2, 1, 12, J=0.6, kappa=0.7, h=0.1))
    seq = [lt.solve(op, nev=2, want_vectors=True) for op in ops]
    for workers in (1, 3, 8):
        bat = lt.solve_batch(ops, nev=2, want_vectors=True, nworkers=workers)
        for a, b in zip(seq, bat):
            np.testing.assert_array_equal(a.evals, b.evals)
            np.testing.assert_array_equal(a.vecs, b.vecs)


def test_scan_rule_small_widths(lt, oracle):
    grid = scan_points()
    assert len(grid) == 54
    widths = (6, 8, 10)
    ops, meta = [], []
    for i in range(24):
        bJ, k1, k2 = grid[i % 54]
        L = widths[i % len(widths)]
        ops.append(lt.TransferMatrix("B", 2, 3, L, J=bJ, kappa=k1, kappa2=k2))
        meta.append((L, bJ, k1, k2))
    res = lt.solve_batch(ops, nev=2, precision=lt.TM_PREC_MIXED, nworkers=4)
    for (L, bJ, k1, k2), r in zip(meta, res):
        D = oracle.OracleOp(oracle.ZBT, L, 0, 0, bJ, k1, k2, 0.0).dense()
        w = np.linalg.eigvals(D)
        lam0 = w[np.argmax(np.abs(w))]
        print(L, bJ, k1, k2, r.evals[0], lam0, flush=True)
        assert abs(r.evals[0] / lam0 - 1) < 1e-12
