"""Device orbit tables at full widths against the symmetry group of each
reduced operator (tests/groups.py): representatives are the minima of their
orbits, sizes are the number of distinct images, the sizes sum to 2^n and
the count is Burnside's over the group's cycle index.  The tables also match
the oracle's sequential enumerations bit for bit.

  ZBM  L=24, Z group with and without the flip   (TMat/tmatreduce.cpp:242-334)
  XDM  L=26, dihedral group with the flip         (TMat/tmatreduce.cpp:19-111)
  YAM  L=12 (24-bit two-layer states), with flip (TMat/tmatreduce.cpp:114-239)
  YBM  L=12, without flip
  XAM3D 5x5 at kappa=0: the quarter turn joins the group (TMat/tmat3d.cpp:17-98)
"""
import numpy as np
import pytest

from tests import groups

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as m
    return m


CASES = [
    ("ZBM", ("B", 2, 7, 24), dict(J=0.4, kappa=0.5, kappa2=0.3), lambda: groups.gens_z(24), 1),
    ("ZBM", ("B", 2, 7, 24), dict(J=0.4, kappa=0.5, kappa2=0.3, h=0.1), lambda: groups.gens_z(24), 0),
    ("XDM", ("D", 2, 5, 26), dict(J=0.5, kappa=0.6), lambda: groups.gens_x(26), 1),
    ("YAM", ("A", 2, 6, 12), dict(J=0.45, kappa=0.35), lambda: groups.gens_y(12), 1),
    ("YBM", ("B", 2, 6, 12), dict(J=0.45, kappa=0.3, kappa2=0.2, h=0.1), lambda: groups.gens_y(12), 0),
    ("XAM3D", ("A", 3, 1, (5, 5)), dict(J=0.25, kappa=0.0), lambda: groups.gens_3d(5, 5, True), 1),
]


@pytest.mark.parametrize("kind,args,kw,gens,flip", CASES,
                         ids=[f"{c[0]}-flip{c[4]}" for c in CASES])
def test_orbit_tables_full_width(lt, oracle, kind, args, kw, gens, flip):
    op = lt.TransferMatrix(*args, **kw)
    assert op.kind() == kind
    reps, dups = op.orbits()
    print(f"  {kind} flip={flip} orbits={len(reps)}", flush=True)
    groups.check_orbit_tables(reps, dups, gens(), flip)
    if kind in ("ZBM",):
        oreps, odups = oracle.orbits_z2d(args[3], flip)
    elif kind == "XDM":
        oreps, odups = oracle.orbits_x2d(args[3], 0, flip)
    elif kind in ("YAM", "YBM"):
        oreps, odups, _ = oracle.orbits_y2d(args[3], flip)
    else:
        oreps, odups = oracle.orbits_x3d(5, 5, flip, 1)
    np.testing.assert_array_equal(reps, oreps)
    np.testing.assert_array_equal(dups, odups)
