"""Restart cycles replayed from hipGraphs against the same solve launched
eagerly.  Every replay feeds the next restart (eager host work, restart
kernels and a synchronisation sit between replays), so equal eigenvalues
and vectors, bit for bit, check every replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


CASES = [
    ("XAT L=14", dict(model="A", axialdir=1, size=14, J=0.44, kappa=0.2), dict(nev=2)),
    ("XAM L=20 mixed", dict(model="A", axialdir=5, size=20, J=0.4, kappa=0.5, neigs=4),
     dict(nev=4, precision=1)),
    ("ZBT L=12 left", dict(model="B", axialdir=3, size=12, J=0.5, kappa=0.5, kappa2=0.3),
     dict(nev=2, want_left=True)),
    ("XDT L=16", dict(model="D", axialdir=1, size=16, J=0.6, kappa=0.7, h=0.1), dict(nev=1)),
]


@pytest.mark.parametrize("name,op_kw,solve_kw", CASES, ids=[c[0] for c in CASES])
def test_graph_replay_equals_eager(lt, name, op_kw, solve_kw):
    kw = dict(op_kw)
    model, ax, size = kw.pop("model"), kw.pop("axialdir"), kw.pop("size")
    op = lt.TransferMatrix(model, 2, ax, size, **kw)
    ref = lt.solve(op, want_vectors=True, use_graph=False, **solve_kw)
    assert ref.extra["graph_cycles"] == 0
    for rep in range(3):
        g = lt.solve(op, want_vectors=True, use_graph=True, **solve_kw)
        print(name, rep, "graph cycles", g.extra["graph_cycles"], "restarts", g.restarts,
              flush=True)
        assert g.extra["graph_cycles"] > 0
        np.testing.assert_array_equal(g.evals, ref.evals)
        np.testing.assert_array_equal(g.vecs, ref.vecs)
        if solve_kw.get("want_left"):
            assert g.eval_left == ref.eval_left
        e = lt.solve(op, want_vectors=True, use_graph=False, **solve_kw)
        np.testing.assert_array_equal(e.evals, ref.evals)
