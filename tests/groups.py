"""Symmetry groups of the reduced operators as permutations of bit positions,
and orbit counts by Burnside's lemma over the group's cycle index.

These are independent of the orbit enumerations under test (which walk
states): the count is (1/|G|) sum_g fix(g), where a position permutation
with c cycles fixes 2^c states, and the permutation followed by the global
spin flip fixes 2^c states when every cycle has even length (the bits
alternate round each cycle) and none otherwise.  A permutation p sends bit
i of a state to bit p[i].

Groups (bit i = site i; 3D bit r*s2 + c = row r, column c):
  X   rotations and the mirror i -> L-1-i  (genX2DDuplist,
      TMat/tmatreduce.cpp:19-111)
  Z   rotations by two sites and the mirror i -> -i mod L  (genZ2DDuplist,
      TMat/tmatreduce.cpp:242-334)
  Y   both layers of a two-layer state rotated / mirrored together
      (genY2DDuplist, TMat/tmatreduce.cpp:114-239)
  3D  row and column translations, row and column mirrors, and the
      quarter turn of a square layer  (genX3DDuplist, TMat/tmat3d.cpp:17-98)
"""
import numpy as np


def gens_x(L):
    return [tuple((i + 1) % L for i in range(L)), tuple(L - 1 - i for i in range(L))]


def gens_z(L):
    return [tuple((i + 2) % L for i in range(L)), tuple((-i) % L for i in range(L))]


def gens_y(L):
    """Two-layer index (high << L) | low; both layers move together."""
    return [tuple((i % L + 1) % L + (i // L) * L for i in range(2 * L)),
            tuple(L - 1 - (i % L) + (i // L) * L for i in range(2 * L))]


def gens_3d(s1, s2, rotate):
    maps = [lambda r, c: ((r + 1) % s1, c), lambda r, c: (r, (c + 1) % s2),
            lambda r, c: (s1 - 1 - r, c), lambda r, c: (r, s2 - 1 - c)]
    if rotate:
        assert s1 == s2
        maps.append(lambda r, c: (c, s1 - 1 - r))
    out = []
    for f in maps:
        p = [0] * (s1 * s2)
        for r in range(s1):
            for c in range(s2):
                rr, cc = f(r, c)
                p[r * s2 + c] = rr * s2 + cc
        out.append(tuple(p))
    return out


def closure(gens):
    """Every element of the group the permutations generate."""
    e = tuple(range(len(gens[0])))
    seen, todo = {e}, [e]
    while todo:
        p = todo.pop()
        for g in gens:
            q = tuple(g[i] for i in p)
            if q not in seen:
                seen.add(q)
                todo.append(q)
    return sorted(seen)


def cycle_lengths(p):
    seen = [False] * len(p)
    out = []
    for i in range(len(p)):
        n, j = 0, i
        while not seen[j]:
            seen[j] = True
            j = p[j]
            n += 1
        if n:
            out.append(n)
    return out


def burnside_count(gens, flip):
    G = closure(gens)
    tot = 0
    for g in G:
        c = cycle_lengths(g)
        tot += 2 ** len(c)
        if flip and all(x % 2 == 0 for x in c):
            tot += 2 ** len(c)
    d = len(G) * (2 if flip else 1)
    assert tot % d == 0
    return tot // d


def _apply(states, p):
    """Image of every state under the position permutation p, moving all
    bits that share one shift distance together."""
    by_shift = {}
    for i, j in enumerate(p):
        by_shift[j - i] = by_shift.get(j - i, 0) | (1 << i)
    t = np.zeros_like(states)
    for d, m in by_shift.items():
        part = states & np.uint64(m)
        t |= (part << np.uint64(d)) if d >= 0 else (part >> np.uint64(-d))
    return t


def check_orbit_tables(reps, dups, gens, flip, chunk=1 << 16):
    """Orbit tables (representatives ascending, sizes) against the group:
    every representative is the minimum of its orbit, its size is the
    number of distinct images, the sizes sum to 2^n and the count equals
    Burnside's.  Distinct minima plus sizes summing to 2^n make the table
    a partition of the state space."""
    n = len(gens[0])
    G = closure(gens)
    reps = np.asarray(reps, np.uint64)
    dups = np.asarray(dups)
    assert len(reps) == burnside_count(gens, flip)
    assert int(dups.astype(np.int64).sum()) == 1 << n
    assert np.all(reps[1:] > reps[:-1])
    mask = np.uint64((1 << n) - 1)
    for a in range(0, len(reps), chunk):
        s = reps[a:a + chunk]
        imgs = np.empty((len(G) * (2 if flip else 1), len(s)), np.uint64)
        for k, g in enumerate(G):
            imgs[k] = _apply(s, g)
        if flip:
            imgs[len(G):] = ~imgs[:len(G)] & mask
        np.testing.assert_array_equal(imgs.min(axis=0), s)
        srt = np.sort(imgs, axis=0)
        size = 1 + np.count_nonzero(np.diff(srt, axis=0), axis=0)
        np.testing.assert_array_equal(size, dups[a:a + chunk].astype(np.int64))
