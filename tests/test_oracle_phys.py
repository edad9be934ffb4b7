"""Pins of the oracle's observables (or_op_moments, the X_phys.dat figures)
by closed forms that do not restate it:
  - uniform weight over the states: m = 2 popcount - L is a sum of L
    independent +-1 spins, so <m> = <m^3> = 0, <m^2> = L, <m^4> = 3L^2 - 2L;
    the periodic node count is a sum of L Bernoulli(1/2) bond indicators
    with pairwise-independent terms, <n> = L/2, <n^2> = (L^2 + L)/4;
  - reduced operators: q_o = sqrt(d_o) gives each state of the orbit the
    same weight, so the figures equal the full-space ones -- all of them
    without spin-flip folding, the flip-invariant ones (even moments, node
    counts, |m|, |m_st|, squares) with it, since a folded orbit's
    representative carries one sign of m for both;
  - weight on one configuration: the figures are that state's values (which
    layer carries m for the two-layer operators, the staggered magnetisation
    of the DNNI observables);
  - the non-symmetric normalisation by sum ql qr: scaling ql leaves the
    figures unchanged.
"""
from math import comb

import numpy as np
import pytest


def _binomial(L):
    return [0.0, 1.0, 0.0, 3.0 * L - 2.0]


def _uniform_reduced(op):
    d = op.dups().astype(np.float64)
    return np.sqrt(d / d.sum())


@pytest.mark.parametrize("L", [8, 11])
def test_xat_uniform(oracle, L):
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, 0.4, 0.3, 0.0, 0.0)
    q = np.full(op.n, 1.0 / np.sqrt(op.n))
    f = op.moments(q)
    assert len(f) == 6
    np.testing.assert_allclose(f[:4], _binomial(L), atol=1e-12)
    np.testing.assert_allclose(f[4:], [L / 2, (L * L + L) / 4], rtol=1e-13)


def test_xat_single_state(oracle):
    L = 10
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, 0.4, 0.3, 0.0, 0.0)
    s = 0b0101010111
    b = format(s, f"0{L}b")
    up = b.count("1")
    walls = sum(b[i] != b[(i + 1) % L] for i in range(L))  # periodic domain walls
    assert (up, walls) == (6, 8)
    q = np.zeros(op.n)
    q[s] = 1.0
    m = 2 * up - L
    f = op.moments(q)
    np.testing.assert_allclose(f, [m / L, m**2 / L, m**3 / L, m**4 / L, walls, walls**2],
                               rtol=1e-14)


def _check_moments(f, L, folded):
    want = _binomial(L)
    idx = [1, 3] if folded else [0, 1, 2, 3]
    np.testing.assert_allclose(np.asarray(f)[idx], np.asarray(want)[idx], atol=1e-12)


@pytest.mark.parametrize("ext", [0, 2])
@pytest.mark.parametrize("kind,L", [("XAM2D", 12), ("XDM", 10)])
def test_reduced_uniform_equals_full(oracle, kind, L, ext):
    op = oracle.OracleOp(getattr(oracle, kind), L, 0, ext, 0.4, 0.3, 0.0, 0.0)
    f = op.moments(_uniform_reduced(op))
    _check_moments(f[:4], L, ext == 0)
    if kind == "XAM2D":
        np.testing.assert_allclose(f[4:], [L / 2, (L * L + L) / 4], rtol=1e-13)
    else:
        # flipping the odd sites is a bijection: m_st has m's distribution
        e_abs = sum(comb(L, k) * abs(2 * k - L) for k in range(L + 1)) / 2**L
        np.testing.assert_allclose(f[4:], [e_abs / L, e_abs / L, 1 / L, 1 / L], rtol=1e-13)


def test_xdm_neel_orbit(oracle):
    L = 10
    op = oracle.OracleOp(oracle.XDM, L, 0, 0, 0.4, 0.3, 0.0, 0.0)
    reps = op.reps()
    q = np.zeros(op.n)
    q[np.searchsorted(reps, 0b0101010101)] = 1.0
    f = op.moments(q)
    # m = 0; the staggered magnetisation is +-L
    np.testing.assert_allclose(f, [0, 0, 0, 0, 1, 0, 1, 0], atol=1e-15)


@pytest.mark.parametrize("s1,s2", [(3, 4), (2, 5)])
def test_3d_uniform(oracle, s1, s2):
    L = s1 * s2
    for kind in ("XAT3D", "XAM3D"):
        op = oracle.OracleOp(getattr(oracle, kind), s1, s2, 0, 0.25, 0.4, 0.0, 0.0)
        q = np.full(op.n, 1.0 / np.sqrt(op.n)) if kind == "XAT3D" else _uniform_reduced(op)
        f = op.moments(q)
        _check_moments(f[:4], L, kind == "XAM3D")
        assert len(f) == 6 and np.all(np.isnan(f[4:]))


@pytest.mark.parametrize("kind,L,ext", [("ZBT", 10, 0), ("ZBM", 12, 0), ("ZBM", 12, 2)])
def test_gen_uniform_and_scale(oracle, kind, L, ext):
    op = oracle.OracleOp(getattr(oracle, kind), L, 0, ext, 0.4, 0.3, 0.2, 0.0)
    q = np.full(op.n, 1.0 / np.sqrt(op.n)) if kind == "ZBT" else _uniform_reduced(op)
    f = op.moments(q, 3.7 * q)
    assert len(f) == 4
    _check_moments(f, L, kind == "ZBM" and ext == 0)


@pytest.mark.parametrize("kind", ["YAT", "YAM", "YBM"])
def test_two_layer_magnetisation_layer(oracle, kind):
    """YAT weighs the low layer of the 4^L index (GenMatPhys), YAM / YBM the
    representative's high layer (tmatreduce.cpp:882-909): put all weight on
    states whose carrying layer is all up and the other layer uniform (no
    flip folding, ext = 2, so that such representatives exist)."""
    L = 5
    op = oracle.OracleOp(getattr(oracle, kind), L, 0, 2, 0.4, 0.3, 0.2, 0.0)
    lm = (1 << L) - 1
    states = np.arange(op.n, dtype=np.uint64) if kind == "YAT" else op.reps()
    layer = (states & np.uint64(lm)) if kind == "YAT" else (states >> np.uint64(L))
    q = (layer == lm).astype(np.float64)
    assert q.sum() > 1
    f = op.moments(q, q)
    np.testing.assert_allclose(f, [1.0, L, L**2, L**3], rtol=1e-14)
    # uniform over everything: the binomial figures of one layer
    qu = np.full(op.n, 1.0) if kind == "YAT" else np.sqrt(op.dups().astype(np.float64))
    np.testing.assert_allclose(op.moments(qu, qu), _binomial(L), atol=1e-12)


def test_y2at_zeros(oracle):
    op = oracle.OracleOp(oracle.Y2AT, 4, 0, 0, 0.4, 0.3, 0.0, 0.0)
    q = np.random.default_rng(1).uniform(0, 1, op.n)
    f = op.moments(q, q)
    np.testing.assert_array_equal(f, np.zeros(4))


def test_planted_layer_mistake_is_caught(oracle):
    """A YAT weighting by the high layer instead (the planted mistake) gives
    the uniform figures, not the all-up ones the test above demands."""
    L = 5
    op = oracle.OracleOp(oracle.YAT, L, 0, 0, 0.4, 0.3, 0.0, 0.0)
    lm = (1 << L) - 1
    st = np.arange(op.n, dtype=np.uint64)
    q = ((st & np.uint64(lm)) == lm).astype(np.float64)
    m_high = 2.0 * np.array([bin(int(s) >> L).count("1") for s in st]) - L
    wrong = [float(np.sum(m_high**k * q * q) / (L * np.sum(q * q))) for k in range(1, 5)]
    assert not np.allclose(wrong, [1.0, L, L**2, L**3])
