"""The tile-size autotuner: an operator whose first apply timed candidate
tile sizes gives the same T.x (to rounding: the butterfly stages are grouped
differently) as one fixed to the default tiles (TM_FLAG_NO_AUTOTUNE), and the
tuned choice stays in force for later applies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("XAT L=20", ("A", 2, 1, 20), dict(J=0.4, kappa=0.5, h=0.1)),
    ("XAT3D 4x5", ("A", 3, 1, (4, 5)), dict(J=0.25, kappa=0.4, flags=1)),
    ("ZBT L=16", ("B", 2, 3, 16), dict(J=0.5, kappa=0.5, kappa2=0.3)),
    ("XDT L=18", ("D", 2, 1, 18), dict(J=0.6, kappa=0.7, h=0.1)),
]


@pytest.mark.parametrize("name,args,kw", CASES, ids=[c[0] for c in CASES])
def test_autotuned_equals_default_tiles(name, args, kw):
    import latttmat_amd as lt
    flags = kw.pop("flags", 0)
    tuned = lt.TransferMatrix(*args, flags=flags, **kw)
    fixed = lt.TransferMatrix(*args, flags=flags | lt.TM_FLAG_NO_AUTOTUNE, **kw)
    rng = np.random.default_rng(11)
    for rep in range(3):
        x = rng.uniform(-1, 1, tuned.cols())
        for tr in (False, True):
            a = tuned.apply(x, transpose=tr)
            b = fixed.apply(x, transpose=tr)
            e = np.max(np.abs(a - b)) / np.max(np.abs(b))
            print(name, rep, tr, e, flush=True)
            assert e < 1e-13
    r1 = lt.solve(tuned, nev=1)
    r2 = lt.solve(fixed, nev=1)
    assert abs(r1.evals[0] / r2.evals[0] - 1) < 1e-12
