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
  non-symmetric       start p(s) ~ |vec_s vecl_s|; YAT moves (b, a) -> (c, b)
                      with p(c) ~ exp(-kappa betaJ sum s_i(a) s_i(c)) vec_(c,b);
                      ZBM members ~ exp(cross) of the diagonal transfer matrix;
                      YAM / YBM members restricted to low layer = current high
                      layer, ~ exp(-kappa betaJ sum s_i(a_low) s_i(s_high))
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


def _cross_zbm(a, s, L, bJ, k, k2):
    """ZBM2DDumpConfig::crossintr from the couplings: -kappa2 on the same
    site, -kappa at distance two, betaJ from each even site of a to the odd
    neighbours of s (DumpConfig/dumpreduce.cpp:389-407)."""
    sa, sb = _spins(a, L), _spins(s, L)
    x = -k2 * bJ * float(sa @ sb) - k * bJ * float(sa @ (np.roll(sb, -2) + np.roll(sb, 2)))
    ev = np.arange(0, L, 2)
    x += bJ * float(sa[ev] @ (np.roll(sb, 1)[ev] + np.roll(sb, -1)[ev]))
    return x


def test_yat_move_distribution(oracle):
    """YAT: start ~ |vec vecl| over 4^L; the state (b, a) moves to (c, b) with
    p(c) ~ exp(-kappa betaJ sum_i s_i(a) s_i(c)) |vec_(c,b)|."""
    L, M, bJ, k = 4, 20000, 0.45, 0.35
    op = oracle.OracleOp(oracle.YAT, L, 0, 0, bJ, k, 0.0, 0.1)
    v, vl = _vec(op.n, 11), _vec(op.n, 13)
    u = (np.arange(M) + 0.5) / M
    st = op.sample_configs(v, u, M, 1, vl)[:, 0]
    cnt = np.bincount(st.astype(np.int64), minlength=op.n)
    p = np.abs(v * vl) / np.sum(np.abs(v * vl))
    assert np.max(np.abs(cnt - M * p)) <= 1.0
    uu = np.empty((M, 2))
    uu[:, 0] = 0.43
    uu[:, 1] = u
    st = op.sample_configs(v, uu.ravel(), M, 2, vl)
    s0 = int(st[0, 0])
    lm = (1 << L) - 1
    a, b = s0 & lm, s0 >> L
    assert np.all(st[:, 1] & np.uint64(lm) == b)
    w = np.array([np.exp(-k * bJ * float(_spins(a, L) @ _spins(c, L))) * abs(v[(c << L) | b])
                  for c in range(1 << L)])
    cnt = np.bincount((st[:, 1] >> np.uint64(L)).astype(np.int64), minlength=1 << L)
    assert np.max(np.abs(cnt - M * w / w.sum())) <= 1.0


@pytest.mark.parametrize("kind,L,ext,h", [("ZBM", 6, 0, 0.0), ("ZBM", 8, 2, 0.1), ("YAM", 4, 0, 0.0),
                                          ("YBM", 4, 2, 0.0)])
def test_nonsymmetric_reduced_move_distribution(oracle, kind, L, ext, h):
    bJ, k, k2 = 0.45, 0.35, 0.25
    op = oracle.OracleOp(getattr(oracle, kind), L, 0, ext, bJ, k, k2, h)
    n = op.n
    reps, dups = op.reps(), op.dups()
    v, vl = _vec(n, 17), _vec(n, 19)
    M1 = M2 = 150
    u = np.empty((M1 * M2, 3))
    u[:, 0] = 0.52
    u[:, 1] = np.repeat((np.arange(M1) + 0.5) / M1, M2)
    u[:, 2] = np.tile((np.arange(M2) + 0.5) / M2, M1)
    st = op.sample_configs(v, u.ravel(), M1 * M2, 2, vl)
    a = int(st[0, 0])
    assert np.all(st[:, 0] == a)
    qa = int(np.searchsorted(reps, a))
    assert reps[qa] == a
    e = np.zeros(n)
    e[qa] = 1.0
    po = op.apply(e) * np.abs(v)
    po /= po.sum()
    two = kind != "ZBM"
    nb = 2 * L if two else L
    G = groups.closure(groups.gens_y(L) if two else groups.gens_z(L))
    flip = h == 0 and (ext & 2) == 0
    full = (1 << nb) - 1
    lm = (1 << L) - 1
    want = {}
    for o, r in enumerate(reps):
        mem = set()
        for g in G:
            s = sum(((int(r) >> i) & 1) << g[i] for i in range(nb))
            mem.add(s)
            if flip:
                mem.add(~s & full)
        assert len(mem) == dups[o]
        if two:
            w = {s: np.exp(-k * bJ * float(_spins(a & lm, L) @ _spins(s >> L, L)))
                 for s in mem if (s & lm) == (a >> L)}
        else:
            w = {s: np.exp(_cross_zbm(a, s, L, bJ, k, k2)) for s in mem}
        tot = sum(w.values())
        for s, ws in w.items():
            want[s] = want.get(s, 0.0) + po[o] * ws / tot
    cnt = np.bincount(st[:, 1].astype(np.int64), minlength=1 << nb)
    exp_cnt = np.zeros(1 << nb)
    for s, p in want.items():
        exp_cnt[s] = M1 * M2 * p
    assert np.max(np.abs(cnt - exp_cnt)) <= M1 + M2
    assert np.sum(cnt[exp_cnt == 0]) == 0
