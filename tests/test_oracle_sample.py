"""Pins of the oracle's configuration sampler (or_sample_configs, the
reference's dumpconfig chains) by the distributions its draws must follow.

With the uniforms of many chains placed on a regular grid, the number of
chains that land on an outcome k equals M p_k to within the grid's
resolution (|count - M p_k| <= 1 for one uniform per draw, <= M1 + M2 for a
grid of M1 x M2 uniform pairs).  The probabilities are written here from the
lattice couplings, not from the oracle's code:
  start               p(s) = vec_s^2 / sum vec^2           (SymDumpConfig::initq)
  XAT move a -> c     p(c) ~ exp(betaJ sum_i s_i(a) s_i(c)) vec_c
                      (the factor exp(self(a)) is common to all c)
  XAM / XDM move      p(o) ~ (T e_a)_o vec_o over orbits, then within orbit o
                      p(s) ~ exp(cross(a, s)), cross the in-plane coupling of
                      the two layers (DNNI: plus the diagonal -kappa bonds)
"""
import numpy as np
import pytest

from tests import groups


def _spins(s, L):
    return np.array([1 - 2 * ((int(s) >> i) & 1) for i in range(L)], dtype=np.float64)


def _cross(a, s, L, bJ, k, dnni):
    sa, sb = _spins(a, L), _spins(s, L)
    x = bJ * float(sa @ sb)
    if dnni:
        x -= k * bJ * float(sa @ (np.roll(sb, -1) + np.roll(sb, 1)))
    return x


def _vec(n, seed):
    return np.random.default_rng(seed).uniform(0.2, 1.0, n) * np.where(
        np.random.default_rng(seed + 1).uniform(size=n) < 0.3, -1.0, 1.0)


def test_start_distribution(oracle):
    L, M = 6, 20000
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, 0.4, 0.3, 0.0, 0.0)
    v = _vec(op.n, 3)
    u = (np.arange(M) + 0.5) / M
    st = op.sample_configs(v, u, M, 1)[:, 0]
    cnt = np.bincount(st.astype(np.int64), minlength=op.n)
    p = v * v / np.sum(v * v)
    assert np.max(np.abs(cnt - M * p)) <= 1.0


def test_xat_move_distribution(oracle):
    L, M, bJ = 6, 20000, 0.4
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, bJ, 0.3, 0.0, 0.2)
    v = _vec(op.n, 5)
    u = np.empty((M, 2))
    u[:, 0] = 0.37
    u[:, 1] = (np.arange(M) + 0.5) / M
    st = op.sample_configs(v, u.ravel(), M, 2)
    a = int(st[0, 0])
    assert np.all(st[:, 0] == a)
    w = np.array([np.exp(_cross(a, c, L, bJ, 0, False)) * abs(v[c]) for c in range(op.n)])
    cnt = np.bincount(st[:, 1].astype(np.int64), minlength=op.n)
    assert np.max(np.abs(cnt - M * w / w.sum())) <= 1.0


@pytest.mark.parametrize("kind,ext", [("XAM2D", 0), ("XAM2D", 2), ("XAM2D", 1), ("XDM", 0)])
def test_reduced_move_distribution(oracle, kind, ext):
    L, bJ, k = 6, 0.45, 0.35
    op = oracle.OracleOp(getattr(oracle, kind), L, 0, ext, bJ, k, 0.0, 0.0)
    n = op.n
    reps, dups = op.reps(), op.dups()
    v = _vec(n, 7)
    M1 = M2 = 150
    g1 = (np.arange(M1) + 0.5) / M1
    g2 = (np.arange(M2) + 0.5) / M2
    u = np.empty((M1 * M2, 3))
    u[:, 0] = 0.61
    u[:, 1] = np.repeat(g1, M2)
    u[:, 2] = np.tile(g2, M1)
    st = op.sample_configs(v, u.ravel(), M1 * M2, 2)
    a = int(st[0, 0])
    assert np.all(st[:, 0] == a)
    qa = int(np.searchsorted(reps, a))
    assert reps[qa] == a  # a chain starts at a representative
    e = np.zeros(n)
    e[qa] = 1.0
    po = op.apply(e) * np.abs(v)
    po /= po.sum()
    # the orbit members from the group itself
    gens = groups.gens_x(L) if (ext & 1) == 0 else [tuple(L - 1 - i for i in range(L))]
    G = groups.closure(gens)
    flip = (ext & 2) == 0
    mask = (1 << L) - 1
    want = {}
    for o, r in enumerate(reps):
        mem = set()
        for g in G:
            s = sum(((int(r) >> i) & 1) << g[i] for i in range(L))
            mem.add(s)
            if flip:
                mem.add(~s & mask)
        assert len(mem) == dups[o]
        w = {s: np.exp(_cross(a, s, L, bJ, k, kind == "XDM")) for s in mem}
        tot = sum(w.values())
        for s, ws in w.items():
            want[s] = po[o] * ws / tot
    cnt = np.bincount(st[:, 1].astype(np.int64), minlength=1 << L)
    exp_cnt = np.zeros(1 << L)
    for s, p in want.items():
        exp_cnt[s] = M1 * M2 * p
    assert np.max(np.abs(cnt - exp_cnt)) <= M1 + M2
    assert np.sum(cnt[exp_cnt == 0]) == 0


def test_planted_self_term_mistake_is_caught(oracle):
    """Weighting a move by exp(self(c)) as well (a plausible slip) moves the
    distribution by far more than the grid tolerance."""
    L, M, bJ, k, h = 6, 20000, 0.4, 0.3, 0.2
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, bJ, k, 0.0, h)
    v = _vec(op.n, 5)
    u = np.empty((M, 2))
    u[:, 0] = 0.37
    u[:, 1] = (np.arange(M) + 0.5) / M
    st = op.sample_configs(v, u.ravel(), M, 2)
    a = int(st[0, 0])

    def self_e(c):
        s = _spins(c, L)
        return h * bJ * s.sum() + bJ * float(s @ np.roll(s, -1)) - k * bJ * float(s @ np.roll(s, -2))
    w = np.array([np.exp(self_e(c) + _cross(a, c, L, bJ, 0, False)) * abs(v[c])
                  for c in range(op.n)])
    cnt = np.bincount(st[:, 1].astype(np.int64), minlength=op.n)
    assert np.max(np.abs(cnt - M * w / w.sum())) > 50
