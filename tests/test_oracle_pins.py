"""Pins of the CPU oracle (oracle/tm_oracle.c) against checks that do not
restate its algorithm: textbook bit operations, dense transfer matrices built
from the lattice Hamiltonians, the Kaufman closed form, brute-force orbit
closure, Burnside counts and operator identities.

The reference ships no test vectors or fixtures for this path (its tests only
print, TMat/tests.cpp:17-196), so these checks are what pins the oracle."""
import numpy as np
import pytest

from tests import groups, physics


def test_bit_primitives_exhaustive(oracle):
    L = oracle.lib()
    for n in range(2, 13):
        for s in range(1 << n):
            b = format(s, f"0{n}b")
            for d in (1, 2):
                if d >= n:
                    continue
                # rotate right: bit i+d -> bit i
                assert L.or_offset_state(s, n, d) == int(b[-d:] + b[:-d], 2)
                assert L.or_lffset_state(s, n, d) == int(b[d:] + b[:d], 2)
            assert L.or_bitreversal(s, n) == int(b[::-1], 2)
            assert L.or_spinbzexp(s, 0, n) == 2 * b.count("1") - n


def test_powdi(oracle):
    import ctypes as C
    L = oracle.lib()
    for K in (0.1, 0.5, 1.3):
        a = (C.c_double * 2)(np.exp(K), np.exp(-K))
        for b in range(-12, 13):
            assert abs(L.or_powdi(a, b) / np.exp(-K * b) - 1) < 1e-14


def test_3d_ops(oracle):
    L = oracle.lib()
    s1, s2 = 3, 4
    rng = np.random.default_rng(0)
    for s in rng.integers(0, 1 << 12, 200):
        s = int(s)
        grid = [[(s >> (r * s2 + c)) & 1 for c in range(s2)] for r in range(s1)]

        def enc(g):
            return sum(g[r][c] << (r * s2 + c) for r in range(len(g)) for c in range(len(g[0])))
        # rffset: each row rotated right by 1 (column c+1 -> c)
        assert L.or_rffset3d(s, s1, s2, 1) == enc([[row[(c + 1) % s2] for c in range(s2)] for row in grid])
        # dffset: rows rotated (row r+1 -> r)
        assert L.or_dffset3d(s, s1, s2, 1) == enc([grid[(r + 1) % s1] for r in range(s1)])
        assert L.or_reflect3d(s, s1, s2, 0) == enc([row[::-1] for row in grid])
        assert L.or_reflect3d(s, s1, s2, 1) == enc(grid[::-1])
    for s in rng.integers(0, 1 << 9, 100):
        s = int(s)
        g = [[(s >> (r * 3 + c)) & 1 for c in range(3)] for r in range(3)]
        # the quarter turn: site (r, c) moves to (c, size-1-r)
        rot = [[g[2 - c][r] for c in range(3)] for r in range(3)]
        got = L.or_rotate3d(s, 3)
        gg = [[(got >> (r * 3 + c)) & 1 for c in range(3)] for r in range(3)]
        assert gg == rot
        # of order 4, its square the point reflection
        g2 = L.or_rotate3d(got, 3)
        assert g2 == int(format(s, "09b")[::-1], 2)
        assert L.or_rotate3d(L.or_rotate3d(g2, 3), 3) == s


@pytest.mark.parametrize("L,bJ,k,h", [(6, 0.5, 0.6, 0.0), (8, 0.3, 0.2, 0.1), (7, 0.8, 0.0, -0.2)])
def test_xat_dense_hamiltonian(oracle, L, bJ, k, h):
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, bJ, k, 0.0, h)
    D = physics.xat_dense(L, bJ, k, h) * op.normfactor
    T = op.dense()
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14


@pytest.mark.parametrize("L,bJ,k,h", [(6, 0.6, 0.7, 0.1), (8, 0.4, 0.3, 0.0), (5, 0.5, 0.9, 0.05)])
def test_xdt_dense_hamiltonian(oracle, L, bJ, k, h):
    op = oracle.OracleOp(oracle.XDT, L, 0, 0, bJ, k, 0.0, h)
    D = physics.xdt_dense(L, bJ, k, h) * op.normfactor
    T = op.dense()
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14
    assert np.max(np.abs(T - T.T)) / np.max(np.abs(T)) < 1e-14


@pytest.mark.parametrize("L,bJ,k,k2,h", [(6, 0.5, 0.5, 0.3, 0.1), (8, 0.3, 0.8, 0.0, 0.0),
                                          (8, 0.7, 0.2, 0.5, -0.1)])
def test_zbt_dense_hamiltonian_and_adjoint(oracle, L, bJ, k, k2, h):
    op = oracle.OracleOp(oracle.ZBT, L, 0, 0, bJ, k, k2, h)
    D = physics.zbt_dense(L, bJ, k, k2, h) * op.normfactor
    T = op.dense()
    TT = op.dense(transpose=True)
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14
    # the transpose operator (ZBTMatVecProd2D_T) is the exact adjoint
    assert np.max(np.abs(TT - T.T)) / np.max(np.abs(T)) < 1e-14
    # non-symmetric, positive (Perron): leading eigenvalue real and simple
    w = np.linalg.eigvals(T)
    w = w[np.argsort(-np.abs(w))]
    assert abs(w[0].imag) < 1e-12 * abs(w[0]) and abs(w[0]) > abs(w[1])
    assert np.max(np.abs(T - T.T)) > 1e-3 * np.max(np.abs(T))


@pytest.mark.parametrize("L,bJ,k,h", [(6, 0.5, 0.6, 0.0), (7, 0.3, 0.2, 0.1)])
def test_xat_open_dense_hamiltonian(oracle, L, bJ, k, h):
    # bc 1: the open chain's L-1 and L-2 bonds (TMat/tmat.hpp:233-247)
    op = oracle.OracleOp(oracle.XAT, L, 0, 1, bJ, k, 0.0, h)
    D = physics.xat_dense_open(L, bJ, k, h) * op.normfactor
    T = op.dense()
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14


@pytest.mark.parametrize("s1,s2,k,h", [(3, 3, 0.4, 0.0), (3, 3, 0.0, 0.0), (3, 3, 0.4, 0.1),
                                        (3, 4, 0.4, 0.0), (3, 4, 0.0, 0.05)])
def test_xam3d_projection(oracle, s1, s2, k, h):
    """3D reduced operator (TMat/tmat3d.cpp:102-201) against the projection
    of the dense 3D Hamiltonian matrix onto orbit sums of its group: row and
    column translations and mirrors, the flip when h = 0, and the quarter
    turn when kappa = 0 on a square layer."""
    bJ = 0.25
    red = oracle.OracleOp(oracle.XAM3D, s1, s2, 0, bJ, k, 0.0, h)
    L = s1 * s2
    N = 1 << L
    e = np.array([physics.layer3d_energy(t, s1, s2, bJ, k, h) for t in range(N)])
    pcs = np.array([bin(t).count("1") for t in range(N)])
    idx = np.arange(N)
    cross = bJ * (L - 2 * pcs[idx[:, None] ^ idx[None, :]])
    T = np.exp(0.5 * e[:, None] + 0.5 * e[None, :] + cross) * red.normfactor
    censym = h == 0
    rotate = k == 0 and s1 == s2
    G = groups.closure(groups.gens_3d(s1, s2, rotate))
    states = np.arange(N, dtype=np.uint64)
    imgs = [groups._apply(states, g) for g in G]
    if censym:
        imgs += [~m & np.uint64(N - 1) for m in imgs]
    orbit_min = np.min(np.stack(imgs), axis=0)
    reps, dups = red.reps(), red.dups()
    np.testing.assert_array_equal(reps, np.unique(orbit_min))
    orb_of = np.searchsorted(reps, orbit_min)
    np.testing.assert_array_equal(dups, np.bincount(orb_of))
    P = np.zeros((N, len(reps)))
    P[idx, orb_of] = 1.0
    d = dups.astype(np.float64)
    Tr = (P.T @ T @ P) / np.sqrt(np.outer(d, d))
    Ro = red.dense()
    assert np.max(np.abs(Ro - Tr)) / np.max(np.abs(Tr)) < 1e-13


@pytest.mark.parametrize("s1,s2", [(2, 3), (3, 3), (2, 4)])
def test_xat3d_dense_hamiltonian(oracle, s1, s2):
    op = oracle.OracleOp(oracle.XAT3D, s1, s2, 0, 0.3, 0.4, 0.0, 0.1)
    D = physics.xat3d_dense(s1, s2, 0.3, 0.4, 0.1) * op.normfactor
    T = op.dense()
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14


@pytest.mark.parametrize("L,K", [(8, 0.5), (10, 0.3), (10, 0.5), (9, 0.44)])
def test_kaufman_closed_form(oracle, L, K):
    op = oracle.OracleOp(oracle.XAT, L, 0, 0, K, 0.0, 0.0, 0.0)
    w = np.sort(np.linalg.eigvalsh(op.dense()))[::-1] / op.normfactor
    l0, l1 = physics.kaufman(L, K)
    assert abs(w[0] / l0 - 1) < 1e-13
    assert abs(w[1] / l1 - 1) < 1e-12


@pytest.mark.parametrize("L", [4, 5, 8, 10, 12])
@pytest.mark.parametrize("bc,censym", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_orbits_2d_bruteforce(oracle, L, bc, censym):
    reps, dups = oracle.orbits_x2d(L, bc, censym)
    br, bd = physics.orbit_closure_2d(L, bc, censym)
    np.testing.assert_array_equal(reps, br)
    np.testing.assert_array_equal(dups, bd)
    assert int(dups.sum()) == 1 << L
    assert np.all(np.diff(reps.astype(np.int64)) > 0)


BURNSIDE_NOFLIP = {12: 224, 14: 687, 16: 2250, 18: 7685, 20: 27012}
BURNSIDE_FLIP = {12: 122, 14: 362, 16: 1162, 18: 3914, 20: 13648}


def _burnside(L, flip):
    """Orbit count of D_L (x Z2 flip) on L-bit necklaces by Burnside's lemma."""
    from math import gcd
    tot = 0
    # rotations: a rotation by r has gcd(r,L) cycles; with flip it fixes a
    # string only if the cycle length L/gcd is even (then 2^(gcd)) ...
    for r in range(L):
        g = gcd(r, L)
        tot += 2 ** g
        if flip:
            cyc = L // g
            tot += 2 ** g if cyc % 2 == 0 else 0
    # reflections
    for _ in range(L):
        pass
    if L % 2 == 0:
        refl = (L // 2) * 2 ** (L // 2 + 1) + (L // 2) * 2 ** (L // 2)
        refl_f = (L // 2) * 0 + (L // 2) * 2 ** (L // 2)
    else:
        refl = L * 2 ** ((L + 1) // 2)
        refl_f = 0
    tot += refl + (refl_f if flip else 0)
    return tot // (2 * L * (2 if flip else 1))


@pytest.mark.parametrize("L", sorted(BURNSIDE_NOFLIP))
def test_orbit_counts_burnside(oracle, L):
    assert len(oracle.orbits_x2d(L, 0, 0)[0]) == BURNSIDE_NOFLIP[L] == _burnside(L, False)
    assert len(oracle.orbits_x2d(L, 0, 1)[0]) == BURNSIDE_FLIP[L] == _burnside(L, True)


@pytest.mark.parametrize("L", range(4, 31))
def test_x_closed_form_equals_cycle_index(L):
    """The dihedral closed form above against the cycle-index count."""
    gx = groups.gens_x(L)
    assert _burnside(L, False) == groups.burnside_count(gx, False)
    assert _burnside(L, True) == groups.burnside_count(gx, True)


@pytest.mark.parametrize("censym", [0, 1])
def test_orbit_tables_against_groups(oracle, censym):
    """Orbit tables of every group (X, Z, two-layer Y, 3D with and without
    the quarter turn) against the group itself: representatives are orbit
    minima, sizes are image counts, counts are Burnside's."""
    for L in range(4, 21):
        groups.check_orbit_tables(*oracle.orbits_x2d(L, 0, censym), groups.gens_x(L), censym)
        groups.check_orbit_tables(*oracle.orbits_z2d(L, censym), groups.gens_z(L), censym)
    for L in range(3, 10):
        reps, dups, _ = oracle.orbits_y2d(L, censym)
        groups.check_orbit_tables(reps, dups, groups.gens_y(L), censym)
    for s1, s2, rot in ((3, 3, 1), (3, 4, 0), (4, 4, 1), (4, 4, 0), (4, 5, 0)):
        groups.check_orbit_tables(*oracle.orbits_x3d(s1, s2, censym, rot),
                                  groups.gens_3d(s1, s2, rot), censym)


def test_orbits_3d_closure(oracle):
    """3D orbits (translations x row/column mirrors x flip [x rotation]) by
    brute-force closure on 3x3 and 3x4 layers."""
    for s1, s2, censym, rotate in ((3, 3, 1, 1), (3, 4, 1, 0), (3, 4, 0, 0), (2, 4, 1, 0)):
        reps, dups = oracle.orbits_x3d(s1, s2, censym, rotate)
        L = oracle.lib()
        N = 1 << (s1 * s2)
        mask = N - 1
        seen = np.zeros(N, bool)
        br, bd = [], []
        for s in range(N):
            if seen[s]:
                continue
            orb, fr = set(), [s]
            while fr:
                t = fr.pop()
                if t in orb:
                    continue
                orb.add(t)
                fr += [L.or_rffset3d(t, s1, s2, 1), L.or_dffset3d(t, s1, s2, 1),
                       L.or_reflect3d(t, s1, s2, 0), L.or_reflect3d(t, s1, s2, 1)]
                if censym:
                    fr.append(~t & mask)
                if rotate:
                    fr.append(L.or_rotate3d(t, s1))
            for t in orb:
                seen[t] = True
            br.append(min(orb))
            bd.append(len(orb))
        np.testing.assert_array_equal(reps, np.array(br, np.uint64))
        np.testing.assert_array_equal(dups, np.array(bd, np.uint16))


def test_orbit_count_3d_5x5_order(oracle):
    # group order 200 on a 5x5 layer without the quarter turn (400 with it):
    # the literal counts the full-size GPU tests assert, from the cycle index
    assert len(groups.closure(groups.gens_3d(5, 5, False))) * 2 == 200
    assert groups.burnside_count(groups.gens_3d(5, 5, False), True) == 170440
    assert groups.burnside_count(groups.gens_3d(5, 5, True), True) == 86056
    assert groups.burnside_count(groups.gens_x(28), False) == 4806078
    reps, dups = oracle.orbits_x3d(3, 4, 1, 0)
    assert int(dups.sum()) == 1 << 12


@pytest.mark.parametrize("L,ext", [(8, 0), (8, 2), (10, 0), (10, 2), (9, 1), (10, 3)])
def test_xam_projection(oracle, L, ext):
    """T_red(o,o') = (d_o d_o')^-1/2 sum_{s in o, s' in o'} T(s,s') built from
    the dense Hamiltonian, against the oracle's reduced operator."""
    h = 0.0
    bJ, k = 0.45, 0.55
    red = oracle.OracleOp(oracle.XAM2D, L, 0, ext, bJ, k, 0.0, h)
    bc = ext & 1
    if bc:
        # open boundary: the reduced operator's energies subtract L for every
        # bond count (TMat/tmatreduce.cpp:388-399) where the physical chain
        # has L-1 nearest- and L-2 next-nearest-neighbour bonds, so each half
        # diagonal carries exp(bJ/2 - bJ kappa) more than the Hamiltonian's:
        # the whole matrix is nf exp(bJ (1 - 2 kappa)) times the physical one
        T = physics.xat_dense_open(L, bJ, k, h) * red.normfactor * np.exp(bJ * (1 - 2 * k))
    else:
        T = physics.xat_dense(L, bJ, k, h) * red.normfactor
    reps, dups = red.reps(), red.dups()
    censym = (h == 0 and (ext & 2) == 0)
    cr, cd = physics.orbit_closure_2d(L, bc, censym)
    np.testing.assert_array_equal(reps, cr)
    np.testing.assert_array_equal(dups, cd)
    N = 1 << L
    mask = N - 1
    orb_of = np.empty(N, int)

    def rev(s):
        return int(format(s, f"0{L}b")[::-1], 2)
    for o, r in enumerate(reps):
        fr, orb = [int(r)], set()
        while fr:
            t = fr.pop()
            if t in orb:
                continue
            orb.add(t)
            fr += [rev(t)] + ([] if bc else [physics.rotr(t, L, 1)]) + \
                ([~t & mask] if censym else [])
        for t in orb:
            orb_of[t] = o
    M = len(reps)
    P = np.zeros((N, M))
    P[np.arange(N), orb_of] = 1.0
    d = dups.astype(np.float64)
    Tr = (P.T @ T @ P) / np.sqrt(np.outer(d, d))
    Ro = red.dense()
    assert np.max(np.abs(Ro - Tr)) / np.max(np.abs(Tr)) < 1e-13


def test_reduced_lambda0_equals_full(oracle):
    for L in (10, 12):
        full = oracle.OracleOp(oracle.XAT, L, 0, 0, 0.4, 0.5, 0.0, 0.0)
        red = oracle.OracleOp(oracle.XAM2D, L, 0, 2, 0.4, 0.5, 0.0, 0.0)
        wf = np.linalg.eigvalsh(full.dense())[-1]
        wr = max(abs(np.linalg.eigvals(red.dense())))
        assert abs(wr / wf - 1) < 1e-13
    # 3D reduced vs full
    full = oracle.OracleOp(oracle.XAT3D, 3, 3, 0, 0.25, 0.4, 0.0, 0.0)
    red = oracle.OracleOp(oracle.XAM3D, 3, 3, 0, 0.25, 0.4, 0.0, 0.0)
    wf = np.linalg.eigvalsh(full.dense())[-1]
    wr = max(abs(np.linalg.eigvals(red.dense())))
    assert abs(wr / wf - 1) < 1e-13


def _closure(L, gens):
    """Orbits of the group generated by gens on L-bit states: (reps ascending,
    sizes, orbit index of every state), by breadth-first closure."""
    N = 1 << L
    orb = np.full(N, -1, dtype=np.int64)
    reps, sizes = [], []
    for s in range(N):
        if orb[s] >= 0:
            continue
        o = len(reps)
        orb[s] = o
        todo, n = [s], 1
        while todo:
            t = todo.pop()
            for g in gens:
                u = g(t)
                if orb[u] < 0:
                    orb[u] = o
                    n += 1
                    todo.append(u)
        reps.append(s)
        sizes.append(n)
    return np.array(reps, np.uint64), np.array(sizes, np.uint16), orb


def _x_gens(L, censym):
    mask = (1 << L) - 1
    g = [lambda s: physics.rotr(s, L, 1), lambda s: int(format(s, f"0{L}b")[::-1], 2)]
    return g + ([lambda s: ~s & mask] if censym else [])


def _z_gens(L, censym):
    # translations by two sites and the mirror i -> -i
    mask = (1 << L) - 1
    g = [lambda s: physics.rotr(s, L, 2),
         lambda s: int(format(physics.rotr(s, L, 1), f"0{L}b")[::-1], 2)]
    return g + ([lambda s: ~s & mask] if censym else [])


def _projection(T, orb, sizes):
    P = np.zeros((T.shape[0], len(sizes)))
    P[np.arange(T.shape[0]), orb] = 1.0 / np.sqrt(sizes.astype(np.float64))[orb]
    return P.T @ T @ P


@pytest.mark.parametrize("L", [4, 6, 8, 10, 12])
@pytest.mark.parametrize("censym", [0, 1])
def test_orbits_z2d_closure(oracle, L, censym):
    reps, dups = oracle.orbits_z2d(L, censym)
    br, bd, _ = _closure(L, _z_gens(L, censym))
    np.testing.assert_array_equal(reps, br)
    np.testing.assert_array_equal(dups, bd)
    assert int(dups.sum()) == 1 << L


@pytest.mark.parametrize("L,bJ,k,h", [(8, 0.5, 0.3, 0.0), (9, 0.4, 0.7, 0.1), (10, 0.6, 0.4, 0.0)])
def test_xdt_commutes_with_dihedral_group(L, bJ, k, h):
    """The DNNI layer transfer matrix is invariant under the group XDM
    reduces by (so P^T T P carries its symmetric-sector spectrum)."""
    T = physics.xdt_dense(L, bJ, k, h)
    for g in _x_gens(L, h == 0):
        perm = np.array([g(s) for s in range(1 << L)])
        assert np.max(np.abs(T[np.ix_(perm, perm)] - T)) <= 1e-14 * np.max(np.abs(T))


@pytest.mark.parametrize("L,bJ,k,k2,h", [(8, 0.5, 0.5, 0.3, 0.0), (10, 0.3, 0.8, 0.1, 0.1)])
def test_zbt_commutes_with_z_group(L, bJ, k, k2, h):
    T = physics.zbt_dense(L, bJ, k, k2, h)
    for g in _z_gens(L, h == 0):
        perm = np.array([g(s) for s in range(1 << L)])
        assert np.max(np.abs(T[np.ix_(perm, perm)] - T)) <= 1e-14 * np.max(np.abs(T))
    # the plain mirror i -> L-1-i is not a symmetry of the diagonal TM
    rev = np.array([int(format(s, f"0{L}b")[::-1], 2) for s in range(1 << L)])
    assert np.max(np.abs(T[np.ix_(rev, rev)] - T)) > 1e-3 * np.max(np.abs(T))


@pytest.mark.parametrize("L,ext,h", [(8, 0, 0.0), (8, 2, 0.0), (10, 0, 0.1), (10, 0, 0.0),
                                     (9, 2, 0.0)])
def test_xdm_projection(oracle, L, ext, h):
    """The restated XDMMatVecProd2D::perform_op equals (d_o d_o')^-1/2 times
    the orbit sums of the dense DNNI transfer matrix."""
    bJ, k = 0.45, 0.35
    red = oracle.OracleOp(oracle.XDM, L, 0, ext, bJ, k, 0.0, h)
    censym = h == 0 and (ext & 2) == 0
    cr, cd, orb = _closure(L, _x_gens(L, censym))
    np.testing.assert_array_equal(red.reps(), cr)
    np.testing.assert_array_equal(red.dups(), cd)
    T = physics.xdt_dense(L, bJ, k, h) * red.normfactor
    Tr = _projection(T, orb, cd)
    R = red.dense()
    assert np.max(np.abs(R - Tr)) / np.max(np.abs(Tr)) < 1e-13
    assert np.max(np.abs(R - R.T)) / np.max(np.abs(R)) < 1e-13


@pytest.mark.parametrize("L,ext,k,k2,h", [(8, 0, 0.5, 0.3, 0.0), (8, 2, 0.2, 0.1, 0.0),
                                          (10, 0, 0.8, 0.4, 0.1), (10, 0, 0.3, 0.7, 0.0),
                                          (10, 2, 0.6, 0.0, 0.0)])
def test_zbm_projection(oracle, L, ext, k, k2, h):
    """The restated ZBMMatVecProd2D::perform_op_internal (both directions)
    equals the orbit-sum projection of the dense BNNNI/3NN diagonal TM."""
    bJ = 0.4
    red = oracle.OracleOp(oracle.ZBM, L, 0, ext, bJ, k, k2, h)
    censym = h == 0 and (ext & 2) == 0
    cr, cd, orb = _closure(L, _z_gens(L, censym))
    np.testing.assert_array_equal(red.reps(), cr)
    np.testing.assert_array_equal(red.dups(), cd)
    T = physics.zbt_dense(L, bJ, k, k2, h) * red.normfactor
    Tr = _projection(T, orb, cd)
    R, RT = red.dense(), red.dense(transpose=True)
    assert np.max(np.abs(R - Tr)) / np.max(np.abs(Tr)) < 1e-13
    assert np.max(np.abs(RT - Tr.T)) / np.max(np.abs(Tr)) < 1e-13


def test_xdm_zbm_lambda0_equal_full(oracle):
    """The Perron vector lies in the symmetric sector: lambda0 of the reduced
    operators equals that of the full ones (normfactors agree)."""
    xdt = oracle.OracleOp(oracle.XDT, 10, 0, 0, 0.5, 0.6, 0.0, 0.0)
    xdm = oracle.OracleOp(oracle.XDM, 10, 0, 0, 0.5, 0.6, 0.0, 0.0)
    assert xdm.normfactor == xdt.normfactor
    wf = np.linalg.eigvalsh(xdt.dense())[-1]
    wr = np.linalg.eigvalsh(xdm.dense())[-1]
    assert abs(wr / wf - 1) < 1e-13
    zbt = oracle.OracleOp(oracle.ZBT, 10, 0, 0, 0.4, 0.6, 0.3, 0.0)
    zbm = oracle.OracleOp(oracle.ZBM, 10, 0, 0, 0.4, 0.6, 0.3, 0.0)
    assert abs(zbm.normfactor / zbt.normfactor - 1) < 1e-15
    wf = max(abs(np.linalg.eigvals(zbt.dense())))
    wr = max(abs(np.linalg.eigvals(zbm.dense())))
    assert abs(wr / wf - 1) < 1e-13


@pytest.mark.parametrize("L,bJ,k,h", [(3, 0.5, 0.3, 0.0), (4, 0.4, 0.6, 0.1), (5, 0.3, 0.2, 0.0)])
def test_yat_y2at_dense_hamiltonian(oracle, L, bJ, k, h):
    """The restated YAT (and its transpose class) and Y2AT against the
    two-layer transfer matrices built from the axial ANNNI Hamiltonian; odd L
    included (the restatement keeps the final layer swap out of place)."""
    yat = oracle.OracleOp(oracle.YAT, L, 0, 0, bJ, k, 0.0, h)
    D = physics.yat_dense(L, bJ, k, h) * yat.normfactor
    T = yat.dense()
    assert np.max(np.abs(T - D)) / np.max(np.abs(D)) < 1e-14
    assert np.max(np.abs(yat.dense(transpose=True) - T.T)) / np.max(np.abs(D)) < 1e-14
    y2 = oracle.OracleOp(oracle.Y2AT, L, 0, 0, bJ, k, 0.0, h)
    D2 = physics.y2at_dense(L, bJ, k, h) * y2.normfactor
    T2 = y2.dense()
    assert np.max(np.abs(T2 - D2)) / np.max(np.abs(D2)) < 1e-14
    # the reference's transposed Y2AT instance is the operator itself
    assert np.array_equal(y2.dense(transpose=True), T2)
    # two axial layers per step: lambda0(Y2AT) / nf = (lambda0(YAT) / nf)^2
    l1 = max(abs(np.linalg.eigvals(T))) / yat.normfactor
    l2 = max(abs(np.linalg.eigvals(T2))) / y2.normfactor
    assert abs(l1 * l1 / l2 - 1) < 1e-13


def _closure_layers(L, censym):
    """Orbits of two-layer states (index (c << L) | p) under rotation and
    reflection of both layers together (and the flip of both)."""
    lm = (1 << L) - 1

    def rot(s):
        return (physics.rotr(s >> L, L, 1) << L) | physics.rotr(s & lm, L, 1)

    def ref(s):
        return ((int(format(s >> L, f"0{L}b")[::-1], 2) << L)
                | int(format(s & lm, f"0{L}b")[::-1], 2))
    gens = [rot, ref] + ([lambda s: ~s & ((1 << (2 * L)) - 1)] if censym else [])
    return _closure(2 * L, gens)


@pytest.mark.parametrize("kind,L,k,k2,h,ext", [("YAM", 4, 0.3, 0.0, 0.0, 0), ("YAM", 4, 0.3, 0.0, 0.0, 2),
                                               ("YAM", 5, 0.6, 0.0, 0.1, 0), ("YBM", 4, 0.3, 0.2, 0.0, 0),
                                               ("YBM", 5, 0.6, 0.4, 0.0, 2)])
def test_yam_ybm_projection(oracle, kind, L, k, k2, h, ext):
    """The restated YAMMatVecProd2D::perform_op (and _T, and YBM's layer)
    equals the orbit-sum projection of the symmetrised two-layer transfer
    matrix built from the Hamiltonian; the orbit tables and the layer-swap
    pairing match a brute-force closure."""
    bJ = 0.45
    red = oracle.OracleOp(getattr(oracle, kind), L, 0, ext, bJ, k, k2, h)
    censym = h == 0 and (ext & 2) == 0
    cr, cd, orb = _closure_layers(L, censym)
    np.testing.assert_array_equal(red.reps(), cr)
    np.testing.assert_array_equal(red.dups(), cd)
    reps, dups, swp = oracle.orbits_y2d(L, int(censym))
    lm = (1 << L) - 1
    for o, r in enumerate(reps):
        r = int(r)
        assert orb[((r & lm) << L) | (r >> L)] == swp[o]
    T = physics.yam_full_dense(L, bJ, k, h, k2 if kind == "YBM" else None) * red.normfactor
    Tr = _projection(T, orb, cd)
    assert np.max(np.abs(red.dense() - Tr)) / np.max(np.abs(Tr)) < 1e-13
    assert np.max(np.abs(red.dense(transpose=True) - Tr.T)) / np.max(np.abs(Tr)) < 1e-13


def test_yam_lambda0_equals_yat(oracle):
    yam = oracle.OracleOp(oracle.YAM, 5, 0, 0, 0.45, 0.35, 0.0, 0.0)
    yat = oracle.OracleOp(oracle.YAT, 5, 0, 0, 0.45, 0.35, 0.0, 0.0)
    assert yam.normfactor == yat.normfactor
    l1 = max(abs(np.linalg.eigvals(yam.dense())))
    l2 = max(abs(np.linalg.eigvals(yat.dense())))
    assert abs(l1 / l2 - 1) < 1e-13
