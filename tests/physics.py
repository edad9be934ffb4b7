"""Dense transfer matrices written directly from the lattice Hamiltonians.

These are independent of the operator algorithms (butterflies, sweeps,
orbit scatter/gather): each entry T(a, b) is the Boltzmann weight of the two
layers a, b summed from the bond definitions.  Spin convention of the
reference: bit i of a state is site i, a set bit is a down spin
(readme.md:100), so sum_i s_i s'_i = L - 2 popcount(a ^ b).

Hamiltonian sources in the reference (the same couplings as the operators):
  ANNNI perp-TM  DumpConfig/dumpreduce.cpp:26-45
  DNNI           DumpConfig/dumpreduce.cpp:495-514, TMat/tmat.cpp:713-734
  BNNNI/3NN      DumpConfig/dumpreduce.cpp:374-407 (diagonal TM)
  3D ANNNI       TMat/tmat3d.cpp:188-199 (layer), Th = prod K over sites
"""
import numpy as np


def pc(x):
    return bin(int(x)).count("1")


def rotr(s, n, d):
    return (s >> d) | ((s & ((1 << d) - 1)) << (n - d))


def rotl(s, n, d):
    return ((s << d) & ((1 << n) - 1)) | (s >> (n - d))


def cross_nn(a, b, L):
    """sum_i s_i s'_i"""
    return L - 2 * pc(a ^ b)


def pairsum(x, L):
    return L - 2 * pc(x)


def xat_dense(L, bJ, k, h):
    """Symmetrised ANNNI perpendicular TM (without the normfactor)."""
    N = 1 << L
    e = np.array([bJ * pairsum(s ^ rotr(s, L, 1), L) - k * bJ * pairsum(s ^ rotr(s, L, 2), L)
                  + h * bJ * pairsum(s, L) for s in range(N)])
    T = np.empty((N, N))
    for a in range(N):
        for b in range(N):
            T[a, b] = np.exp(0.5 * e[a] + 0.5 * e[b] + bJ * cross_nn(a, b, L))
    return T


def xat_dense_open(L, bJ, k, h):
    """The open-chain ANNNI perpendicular TM (without the normfactor): L-1
    nearest-neighbour and L-2 next-nearest-neighbour bonds in the layer."""
    N = 1 << L
    m1, m2 = (N - 1) >> 1, (N - 1) >> 2
    e = np.array([bJ * pairsum((s & m1) ^ (s >> 1), L - 1)
                  - k * bJ * pairsum((s & m2) ^ (s >> 2), L - 2)
                  + h * bJ * pairsum(s, L) for s in range(N)])
    pcs = np.array([pc(t) for t in range(N)])
    idx = np.arange(N)
    cross = bJ * (L - 2 * pcs[idx[:, None] ^ idx[None, :]])
    return np.exp(0.5 * e[:, None] + 0.5 * e[None, :] + cross)


def xdt_dense(L, bJ, k, h):
    """Symmetrised DNNI TM: NN in-layer + field on the layer, NN and the two
    diagonal NNN (-kappa) across layers."""
    N = 1 << L
    e = np.array([bJ * pairsum(s ^ rotr(s, L, 1), L) + h * bJ * pairsum(s, L) for s in range(N)])
    T = np.empty((N, N))
    for a in range(N):
        for b in range(N):
            c = (bJ * cross_nn(a, b, L) - k * bJ * pairsum(a ^ rotr(b, L, 1), L)
                 - k * bJ * pairsum(a ^ rotl(b, L, 1), L))
            T[a, b] = np.exp(0.5 * e[a] + 0.5 * e[b] + c)
    return T


def yat_dense(L, bJ, k, h):
    """ANNNI transfer along the axial (NNN) direction over two-layer states
    (previous layer p, current layer c), index (p << L) | c; one step adds
    layer n coupled to c (NN) and p (NNN, -kappa), with its own in-layer NN
    bonds and field (without the normfactor)."""
    n1 = 1 << L
    lay = np.array([bJ * pairsum(s ^ rotr(s, L, 1), L) + h * bJ * pairsum(s, L)
                    for s in range(n1)])
    cr = np.array([[cross_nn(a, b, L) for b in range(n1)] for a in range(n1)], float)
    T = np.zeros((n1 * n1, n1 * n1))
    for p in range(n1):          # new state (c_old, n): previous = old current
        for n in range(n1):
            for q in range(n1):  # old state (q, p)
                T[(p << L) | n, (q << L) | p] = np.exp(lay[n] + bJ * cr[n, p] - k * bJ * cr[n, q])
    return T


def yam_full_dense(L, bJ, k, h, k2=None):
    """The symmetrised axial transfer over two-layer states (c, p) = (current,
    previous), index (c << L) | p: T[(n, c), (c, p)] = W(n, c) K2(n, p) W(c, p)
    with W(c, p) = exp(half the in-layer energy of c and of the c-p bonds),
    K2 the axial NNN (-kappa).  k2 given: the BNNNI layer (NNN -kappa inside
    the layer, diagonal -k2 between the layers).  Without the normfactor."""
    n1 = 1 << L
    W = np.empty((n1, n1))
    for c in range(n1):
        for p in range(n1):
            e = (bJ * pairsum(c ^ rotr(c, L, 1), L) + bJ * cross_nn(c, p, L)
                 + h * bJ * pairsum(c, L))
            if k2 is not None:
                e += (-k * bJ * pairsum(c ^ rotr(c, L, 2), L)
                      - k2 * bJ * (cross_nn(c, rotr(p, L, 1), L) + cross_nn(c, rotl(p, L, 1), L)))
            W[c, p] = np.exp(0.5 * e)
    cr = np.array([[cross_nn(a, b, L) for b in range(n1)] for a in range(n1)], float)
    T = np.zeros((n1 * n1, n1 * n1))
    for n in range(n1):
        for c in range(n1):
            for p in range(n1):
                T[(n << L) | c, (c << L) | p] = W[n, c] * np.exp(-k * bJ * cr[n, p]) * W[c, p]
    return T


def y2at_dense(L, bJ, k, h):
    """ANNNI transfer by two axial layers over layer pairs, index
    (r << L) | l with r the older and l the newer layer of the pair: NN
    between the new pair's older layer and the old pair's newer one, NNN
    (-kappa) between equal positions of the two pairs; each pair's internal
    energy split half to each side (without the normfactor)."""
    n1 = 1 << L
    N = n1 * n1
    e = np.empty(N)
    for s in range(N):
        lf, rt = s & (n1 - 1), s >> L
        e[s] = (bJ * (pairsum(lf ^ rotr(lf, L, 1), L) + pairsum(rt ^ rotr(rt, L, 1), L)
                      + pairsum(lf ^ rt, L)) + h * bJ * pairsum(s, 2 * L))
    cr = np.array([[cross_nn(a, b, L) for b in range(n1)] for a in range(n1)], float)
    T = np.empty((N, N))
    for new in range(N):
        nl, nr = new & (n1 - 1), new >> L
        for old in range(N):
            ol, orr = old & (n1 - 1), old >> L
            T[new, old] = np.exp(0.5 * e[new] + 0.5 * e[old] + bJ * cr[nr, ol]
                                 - k * bJ * (cr[nl, ol] + cr[nr, orr]))
    return T


def zbt_dense(L, bJ, k, k2, h):
    """BNNNI + 3NN on the diagonal TM (row index = layer a, column = b)."""
    N = 1 << L
    e = np.array([h * bJ * pairsum(s, L) + bJ * pairsum(s ^ rotr(s, L, 1), L)
                  - k2 * bJ * pairsum(s ^ rotr(s, L, 2), L) for s in range(N)])

    def cross(ra, rb):
        t = (-k2 * bJ * pairsum(ra ^ rb, L) - k * bJ * pairsum(ra ^ rotr(rb, L, 2), L)
             - k * bJ * pairsum(ra ^ rotl(rb, L, 2), L))
        # zig-zag nearest neighbours between the two diagonal layers; the
        # rotated copy keeps bit L (= the top bit again) so that the last pair
        # wraps around the ring
        rb2 = (rb << 1) | (rb >> (L - 1))
        nanti = 0
        for _ in range(L // 2):
            nanti += (ra ^ rb2) & 1
            rb2 >>= 2
            nanti += (ra ^ rb2) & 1
            ra >>= 2
        return t + bJ * (L - 2 * nanti)

    T = np.empty((N, N))
    for a in range(N):
        for b in range(N):
            T[a, b] = np.exp(0.5 * e[a] + 0.5 * e[b] + cross(a, b))
    return T


def layer3d_energy(s, s1, s2, bJ, k, h):
    L = s1 * s2
    m = (1 << s2) - 1

    def rff(x, d):
        r = 0
        for i in range(s1):
            r |= rotr((x >> (i * s2)) & m, s2, d) << (i * s2)
        return r
    nn = pairsum(s ^ rff(s, 1), L) + pairsum(s ^ rotr(s, L, s2), L)
    return bJ * nn - k * bJ * pairsum(s ^ rff(s, 2), L) + h * bJ * pairsum(s, L)


def xat3d_dense(s1, s2, bJ, k, h):
    L = s1 * s2
    N = 1 << L
    e = np.array([layer3d_energy(s, s1, s2, bJ, k, h) for s in range(N)])
    T = np.empty((N, N))
    for a in range(N):
        for b in range(N):
            T[a, b] = np.exp(0.5 * e[a] + 0.5 * e[b] + bJ * cross_nn(a, b, L))
    return T


def kaufman(L, K):
    """lambda0, lambda1 of the 2D Ising perpendicular TM (kappa=0, h=0),
    periodic width L, coupling K = betaJ (Kaufman 1949)."""
    Ks = np.arctanh(np.exp(-2 * K))

    def gam(k):
        c = (np.cosh(2 * Ks) * np.cosh(2 * K) - np.sinh(2 * Ks) * np.sinh(2 * K)
             * np.cos(np.pi * k / L))
        return np.arccosh(c)
    pre = (2 * np.sinh(2 * K)) ** (L / 2)
    odd = sum(gam(k) for k in range(1, 2 * L, 2))
    even = sum(gam(k) for k in range(2, 2 * L, 2))
    l0 = pre * np.exp(0.5 * odd)
    l1 = pre * np.exp(0.5 * (2 * (K - Ks) + even))
    return l0, l1


def orbit_closure_2d(L, bc, censym):
    """Brute-force orbits: reps (minima, ascending) and sizes."""
    N = 1 << L
    mask = N - 1
    seen = np.zeros(N, bool)
    reps, dups = [], []

    def rev(s):
        return int(format(s, f"0{L}b")[::-1], 2)
    for s in range(N):
        if seen[s]:
            continue
        orb = set()
        frontier = [s]
        while frontier:
            t = frontier.pop()
            if t in orb:
                continue
            orb.add(t)
            nxt = [rev(t)]
            if bc == 0:
                nxt.append(rotr(t, L, 1))
            if censym:
                nxt.append(~t & mask)
            frontier.extend(nxt)
        for t in orb:
            seen[t] = True
        reps.append(min(orb))
        dups.append(len(orb))
    return np.array(reps, np.uint64), np.array(dups, np.uint16)
