"""Full-size parity at the benchmark configuration (L=28 reduced ANNNI,
4,806,078 orbits) and the 5x5 3D layer, through size-independent properties
and the oracle where it finishes in seconds:
  - orbit count against Burnside's lemma, sizes summing to 2^L;
  - representatives and sizes bit-exact against the oracle's sequential
    enumeration (TMat/tmatreduce.cpp:19-111 restated);
  - the fused two-pass T.x against the generic expand/butterfly/gather T.x
    and against the oracle's T.x, FP64, to 1e-13;
  - symmetry <y, T x> = <T y, x> of the reduced operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as lt
    return lt


@pytest.fixture(scope="module")
def op28(lt):
    return lt.TransferMatrix("A", 2, 5, 28, J=0.4, kappa=0.5, neigs=4)


def test_orbits_l28(lt, op28, oracle):
    print("enumerating on the CPU oracle (sequential, ~20 s)", flush=True)
    reps, dups = op28.orbits()
    assert len(reps) == 4806078
    assert int(dups.astype(np.int64).sum()) == 1 << 28
    oreps, odups = oracle.orbits_x2d(28, 0, 0)
    np.testing.assert_array_equal(reps, oreps)
    np.testing.assert_array_equal(dups, odups)


def test_fused_vs_generic_and_oracle_l28(lt, op28, oracle):
    gen = lt.TransferMatrix("A", 2, 5, 28, J=0.4, kappa=0.5, neigs=4,
                            flags=lt.TM_FLAG_GENERIC_REDUCED)
    x = np.random.default_rng(5).uniform(-1, 1, op28.cols())
    y1 = op28.apply(x)
    y2 = gen.apply(x)
    e = np.max(np.abs(y1 - y2)) / np.max(np.abs(y2))
    print("fused vs generic", e, flush=True)
    assert e < 1e-13
    oop = oracle.OracleOp(oracle.XAM2D, 28, 0, 2, 0.4, 0.5, 0.0, 0.0)
    yo = oop.apply(x)
    e2 = np.max(np.abs(y1 - yo)) / np.max(np.abs(yo))
    print("fused vs oracle", e2, flush=True)
    assert e2 < 1e-13
    z = np.random.default_rng(6).uniform(-1, 1, op28.cols())
    a = float(z @ y1)
    b = float(op28.apply(z) @ x)
    assert abs(a - b) <= 1e-13 * (abs(a) + abs(b))


def test_3d_5x5_orbits(lt):
    op = lt.TransferMatrix("A", 3, 1, (5, 5), J=0.25, kappa=0.4)
    reps, dups = op.orbits()
    assert len(reps) == 170440
    assert int(dups.astype(np.int64).sum()) == 1 << 25
