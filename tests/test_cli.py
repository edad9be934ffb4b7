"""The `tmat` program: the positional input parser (TMat/read_input.cpp:14-92)
on the CPU, and end-to-end runs on the GPU whose output files
(X_val.dat, X_vec.dat, X_vecl.dat, X_phys.dat; TMat/pfunc.hpp:116-199) are
checked against the oracle's dense matrices and closed forms."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TMAT = os.path.join(ROOT, "latttmat_amd", "tmat")

INPUT = """modeltype = {model} # model type. A: ANNNI; B: BNNNI and 3NN; D: DNNI
dim = {dim}       # dimensionality. 2 or 3
axialdir = {axialdir}  # indices of the type of boundary condition (see source)
neigs = {neigs}     # number of eigenvalues been calculated
savevec = {savevec} # options (1 or 0) to save vectors or computing layer magnetication (or not)
outfile = {out}  # prefix for output files
size = {size}        # system width L
J = {J}          # inverse temperature beta * J
kappa = {kappa}    # kappa1 and kappa2 values
h = {h}            # external field
"""


def write_input(tmp, **kw):
    d = dict(model="A", dim=2, axialdir=1, neigs=1, savevec="0 0", out="T", size=8, J=0.5,
             kappa="0.6 0", h=0)
    d.update(kw)
    path = os.path.join(tmp, "input.dat")
    with open(path, "w") as f:
        f.write(INPUT.format(**d))
    return path


def run(args, cwd, timeout=120):
    if not os.path.exists(TMAT):
        pytest.fail(f"{TMAT} is not built; run __graft_entry__.build()")
    return subprocess.run([TMAT, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


# ---------------------------------------------------------------- CPU ----

def test_parse_ok_and_echo(tmp_path):
    p = write_input(str(tmp_path), model="B", axialdir=3, kappa="0.5 0.3", size=10, h=0.25)
    r = run(["--check-input", p], str(tmp_path))
    assert r.returncode == 0, r.stdout
    lines = r.stdout.splitlines()
    assert lines[0] == f"Reading input from file {p}"
    assert lines[1] == "     MODEL : 2D_B3"
    assert lines[2] == "Save & flg : T 0 0"
    assert lines[3] == "      size : 10"
    assert lines[4] == "     betaJ : 0.5"
    assert lines[5] == "     kappa : 0.5"
    assert lines[6] == "    kappa2 : 0.3"
    assert lines[7] == "         h : 0.25"


def test_parse_3d_sizes(tmp_path):
    p = write_input(str(tmp_path), dim=3, size="4 5")
    r = run(["--check-input", p], str(tmp_path))
    assert r.returncode == 0
    assert "      size : 4 5" in r.stdout.splitlines()


def test_parse_kappa2_defaults_to_zero(tmp_path):
    p = write_input(str(tmp_path), model="B", axialdir=3, kappa="0.7")
    r = run(["--check-input", p], str(tmp_path))
    assert r.returncode == 0
    assert "    kappa2 : 0" in r.stdout.splitlines()


@pytest.mark.parametrize("kw,code,msg", [
    (dict(dim=4), 4, "Invalid dim 4"),
    (dict(model="C"), 4, "Invalid model C"),
    (dict(size=3), 4, "Invalid size (<4) 3"),
    (dict(J="abc"), 3, "Error reading input file"),
])
def test_parse_errors(tmp_path, kw, code, msg):
    p = write_input(str(tmp_path), **kw)
    r = run(["--check-input", p], str(tmp_path))
    assert r.returncode == code
    assert msg in r.stdout


def test_missing_file(tmp_path):
    r = run(["--check-input", str(tmp_path / "nope.dat")], str(tmp_path))
    assert r.returncode == 2
    assert "Can't open" in r.stdout
    # without --check-input the program exits 0 like the reference's main()
    r = run([str(tmp_path / "nope.dat")], str(tmp_path))
    assert r.returncode == 0


def test_default_input_name(tmp_path):
    write_input(str(tmp_path))
    r = run(["--check-input"], str(tmp_path))
    assert r.returncode == 0
    assert "Reading input from file input.dat" in r.stdout


# ---------------------------------------------------------------- GPU ----

def moments_full(q2, L, nodes_bc=None):
    s = np.arange(len(q2), dtype=np.uint64)
    pc = np.array([bin(int(v)).count("1") for v in s], dtype=np.float64)
    m = 2 * pc - L
    mom = [float(np.sum(m ** k * q2)) / L for k in range(1, 5)]
    return mom


@pytest.mark.gpu
def test_cli_xat_outputs(tmp_path, oracle):
    L = 10
    p = write_input(str(tmp_path), size=L, neigs=2, savevec="1 1")
    r = run([p], str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    D = oracle.OracleOp(oracle.XAT, L, 0, 0, 0.5, 0.6, 0.0, 0.0)
    T = D.dense()
    w, V = np.linalg.eigh(0.5 * (T + T.T))
    nf = D.normfactor
    vals = np.loadtxt(tmp_path / "T_val.dat")
    assert vals.shape == (2,)
    np.testing.assert_allclose(vals, w[::-1][:2] / nf, rtol=1e-12)
    raw = open(tmp_path / "T_val.dat").read()
    assert not raw.endswith("\n")
    vec = np.fromfile(tmp_path / "T_vec.dat", dtype=np.float64)
    assert vec.size == 2 * (1 << L)
    v0 = vec[: 1 << L]
    assert abs(abs(v0 @ V[:, -1]) - 1) < 1e-10
    phys = open(tmp_path / "T_phys.dat").read().split("\n")
    mom = [float(x) for x in phys[0].split("\t")]
    ref = moments_full(V[:, -1] ** 2, L)
    np.testing.assert_allclose(mom, ref, rtol=1e-8, atol=1e-10)
    assert "Number of iteration:" in r.stdout


def read_phys(path):
    return [[float(x) for x in ln.split("\t")] for ln in open(path).read().split("\n")]


@pytest.mark.gpu
def test_cli_reduced_phys(tmp_path, oracle):
    """The reduced operator's X_phys.dat weights each orbit by the
    magnetisation of its representative (TMat/tmatreduce.cpp:565-613), checked
    against the oracle's orbit basis and dense Perron vector.  Even moments and
    node counts are invariant under the group (spin flip included), so they
    also equal the full operator's."""
    L = 14
    d1, d2 = tmp_path / "full", tmp_path / "red"
    d1.mkdir()
    d2.mkdir()
    p1 = write_input(str(d1), size=L, axialdir=1, savevec="0 1", out="F", J=0.3, kappa="0.4 0")
    p2 = write_input(str(d2), size=L, axialdir=5, savevec="1 1", out="R", J=0.3, kappa="0.4 0")
    r1 = run([p1], str(d1))
    r2 = run([p2], str(d2))
    assert r1.returncode == 0 and r2.returncode == 0
    f = read_phys(d1 / "F_phys.dat")
    g = read_phys(d2 / "R_phys.dat")
    print(f, g, flush=True)
    # oracle: reduced basis with spin-flip folding (neigs=1 -> censym on)
    oop = oracle.OracleOp(oracle.XAM2D, L, 0, 0, 0.3, 0.4, 0.0, 0.0)
    T = oop.dense()
    w, V = np.linalg.eigh(0.5 * (T + T.T))
    q = V[:, -1]
    reps = oop.reps().astype(np.uint64)
    m = 2.0 * np.array([bin(int(s)).count("1") for s in reps]) - L
    mom = [float(np.sum(m ** k * q * q)) / L for k in range(1, 5)]
    sh = (reps >> np.uint64(1)) | ((reps & np.uint64(1)) << np.uint64(L - 1))
    nd = np.array([bin(int(a ^ b)).count("1") for a, b in zip(reps, sh)], float)
    node = [float(np.sum(nd * q * q)), float(np.sum(nd * nd * q * q))]
    np.testing.assert_allclose(g[0], mom, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(g[1], node, rtol=1e-8)
    np.testing.assert_allclose([f[0][1], f[0][3]], [g[0][1], g[0][3]], rtol=1e-8)
    np.testing.assert_allclose(f[1], g[1], rtol=1e-8)
    assert abs(f[0][0]) < 1e-9 and abs(f[0][2]) < 1e-8  # full vector is flip-even
    vr = np.fromfile(d2 / "R_vec.dat", dtype=np.float64)
    assert vr.size == len(reps) and abs(abs(vr @ q) - 1) < 1e-10
    vf = np.loadtxt(d1 / "F_val.dat")
    vv = np.loadtxt(d2 / "R_val.dat")
    assert abs(float(vv) / float(vf) - 1) < 1e-12


@pytest.mark.gpu
def test_cli_zbt_left_and_vecs(tmp_path, oracle):
    L = 8
    p = write_input(str(tmp_path), model="B", axialdir=3, size=L, kappa="0.5 0.3", savevec="1 1",
                    out="Z")
    r = run([p], str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    D = oracle.OracleOp(oracle.ZBT, L, 0, 0, 0.5, 0.5, 0.3, 0.0)
    T = D.dense()
    w, VR = np.linalg.eig(T)
    i0 = np.argmax(np.abs(w))
    lam = w[i0].real
    raw = open(tmp_path / "Z_val.dat").read()
    assert raw.startswith("(") and raw.endswith(")")
    re_, im_ = raw.strip("()").split(",")
    assert abs(float(re_) / (lam / D.normfactor) - 1) < 1e-12 and float(im_) == 0.0
    vr = np.fromfile(tmp_path / "Z_vec.dat", dtype=np.float64)
    vl = np.fromfile(tmp_path / "Z_vecl.dat", dtype=np.float64)
    assert vr.size == vl.size == 1 << L
    assert np.linalg.norm(T @ vr - lam * vr) / abs(lam) < 1e-8
    assert np.linalg.norm(T.T @ vl - lam * vl) / abs(lam) < 1e-8
    chk = [ln for ln in r.stdout.splitlines() if ln.startswith("Eigensolver check: ")]
    assert len(chk) == 1 and abs(float(chk[0].split()[-1]) - 1) < 1e-9
    # non-symmetric moments: weights ql*qr normalised by their sum, no second line
    phys = open(tmp_path / "Z_phys.dat").read()
    assert "\n" not in phys
    mom = [float(x) for x in phys.split("\t")]
    wgt = vl * vr
    s = np.arange(1 << L)
    m = 2 * np.array([bin(int(v)).count("1") for v in s], float) - L
    ref = [float(np.sum(m ** k * wgt) / (L * wgt.sum())) for k in range(1, 5)]
    np.testing.assert_allclose(mom, ref, rtol=1e-8, atol=1e-10)


@pytest.mark.gpu
def test_cli_xdm_phys(tmp_path, oracle):
    """Reduced DNNI (D, axialdir 5): X_phys.dat holds the four moments over
    the orbit representatives and the four staggered / uniform magnetisation
    observables of XDMMatVecProd2D::phys_extend (TMat/tmatreduce.cpp:1289-1317)."""
    L = 12
    p = write_input(str(tmp_path), model="D", axialdir=5, size=L, savevec="1 1", out="X",
                    J=0.5, kappa="0.6 0")
    r = run([p], str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    oop = oracle.OracleOp(oracle.XDM, L, 0, 0, 0.5, 0.6, 0.0, 0.0)
    T = oop.dense()
    w, V = np.linalg.eigh(0.5 * (T + T.T))
    q = V[:, -1]
    assert abs(float(np.loadtxt(tmp_path / "X_val.dat")) / (w[-1] / oop.normfactor) - 1) < 1e-12
    reps = oop.reps()
    pc = np.array([bin(int(s)).count("1") for s in reps], float)
    m = 2 * pc - L
    stag = 2 * np.array([bin(int(s) ^ 0xAAA).count("1") for s in reps], float) - L
    q2 = q * q
    mom = [float(np.sum(m ** k * q2)) / L for k in range(1, 5)]
    ext = [float(np.sum(np.abs(stag) * q2)) / L, float(np.sum(np.abs(m) * q2)) / L,
           float(np.sum(stag ** 2 * q2)) / L ** 2, float(np.sum(m ** 2 * q2)) / L ** 2]
    g = read_phys(tmp_path / "X_phys.dat")
    print(g, mom, ext, flush=True)
    np.testing.assert_allclose(g[0], mom, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(g[1], ext, rtol=1e-8, atol=1e-12)


@pytest.mark.gpu
def test_cli_zbm_left_and_phys(tmp_path, oracle):
    """Reduced diagonal BNNNI (B, axialdir 7): complex X_val.dat, right and
    left Perron vectors over the orbits, moments weighted by ql*qr (no second
    line: ZBMMatVecProd2D::phys_extend reports nothing)."""
    L = 10
    p = write_input(str(tmp_path), model="B", axialdir=7, size=L, kappa="0.5 0.3",
                    savevec="1 1", out="Z")
    r = run([p], str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    D = oracle.OracleOp(oracle.ZBM, L, 0, 0, 0.5, 0.5, 0.3, 0.0)
    T = D.dense()
    w = np.linalg.eigvals(T)
    lam = w[np.argmax(np.abs(w))].real
    raw = open(tmp_path / "Z_val.dat").read()
    re_, im_ = raw.strip("()").split(",")
    # the solve stops at the reference's residual tolerance 1e-10
    # (pfunc.hpp:52): for a non-symmetric matrix the eigenvalue error is first order
    assert abs(float(re_) / (lam / D.normfactor) - 1) < 1e-10 and float(im_) == 0.0
    vr = np.fromfile(tmp_path / "Z_vec.dat", dtype=np.float64)
    vl = np.fromfile(tmp_path / "Z_vecl.dat", dtype=np.float64)
    assert vr.size == vl.size == D.n
    assert np.linalg.norm(T @ vr - lam * vr) / abs(lam) < 1e-8
    assert np.linalg.norm(T.T @ vl - lam * vl) / abs(lam) < 1e-8
    phys = open(tmp_path / "Z_phys.dat").read()
    assert "\n" not in phys
    mom = [float(x) for x in phys.split("\t")]
    wgt = vl * vr
    m = 2 * np.array([bin(int(v)).count("1") for v in D.reps()], float) - L
    ref = [float(np.sum(m ** k * wgt) / (L * wgt.sum())) for k in range(1, 5)]
    np.testing.assert_allclose(mom, ref, rtol=1e-8, atol=1e-10)


@pytest.mark.gpu
def test_cli_axial_outputs(tmp_path, oracle):
    """Axial ANNNI over two-layer states: YAT (axialdir 2) writes its left
    vector from the transpose class and the generic moments over 4^L states
    with the low layer's magnetisation; Y2AT (axialdir 0) takes its own
    operator as the transpose (TMat/tmat.hpp:382-384) and writes zero
    moments (TMat/tmat.cpp:312-317)."""
    L = 5
    d1, d2 = tmp_path / "yat", tmp_path / "y2at"
    d1.mkdir()
    d2.mkdir()
    p1 = write_input(str(d1), axialdir=2, size=L, savevec="1 1", out="Y", J=0.45, kappa="0.3 0")
    p2 = write_input(str(d2), axialdir=0, size=L, savevec="1 1", out="W", J=0.45, kappa="0.3 0")
    r1 = run([p1], str(d1))
    r2 = run([p2], str(d2))
    print(r1.stdout, r1.stderr, r2.stdout, r2.stderr, flush=True)
    assert r1.returncode == 0 and r2.returncode == 0
    D = oracle.OracleOp(oracle.YAT, L, 0, 0, 0.45, 0.3, 0.0, 0.0)
    T = D.dense()
    w = np.linalg.eigvals(T)
    lam = w[np.argmax(np.abs(w))].real
    re_, im_ = open(d1 / "Y_val.dat").read().strip("()").split(",")
    assert abs(float(re_) / (lam / D.normfactor) - 1) < 1e-10 and float(im_) == 0.0
    vr = np.fromfile(d1 / "Y_vec.dat", dtype=np.float64)
    vl = np.fromfile(d1 / "Y_vecl.dat", dtype=np.float64)
    assert vr.size == vl.size == 1 << (2 * L)
    assert np.linalg.norm(T @ vr - lam * vr) / abs(lam) < 1e-8
    assert np.linalg.norm(T.T @ vl - lam * vl) / abs(lam) < 1e-8
    mom = [float(x) for x in open(d1 / "Y_phys.dat").read().split("\t")]
    wgt = vl * vr
    s = np.arange(1 << (2 * L))
    m = 2 * np.array([bin(int(v) & ((1 << L) - 1)).count("1") for v in s], float) - L
    ref = [float(np.sum(m ** k * wgt) / (L * wgt.sum())) for k in range(1, 5)]
    np.testing.assert_allclose(mom, ref, rtol=1e-8, atol=1e-10)
    # Y2AT
    D2 = oracle.OracleOp(oracle.Y2AT, L, 0, 0, 0.45, 0.3, 0.0, 0.0)
    w2 = np.linalg.eigvals(D2.dense())
    lam2 = w2[np.argmax(np.abs(w2))].real
    re_, im_ = open(d2 / "W_val.dat").read().strip("()").split(",")
    assert abs(float(re_) / (lam2 / D2.normfactor) - 1) < 1e-10
    chk = [ln for ln in r2.stdout.splitlines() if ln.startswith("Eigensolver check: ")]
    assert len(chk) == 1 and abs(float(chk[0].split()[-1]) - 1) < 1e-9
    assert open(d2 / "W_phys.dat").read() == "0\t0\t0\t0"


@pytest.mark.gpu
def test_cli_yam_phys(tmp_path, oracle):
    """Reduced axial ANNNI (A, axialdir 6): moments over the representatives
    with the current (high) layer's magnetisation, weighted by ql*qr
    (TMat/tmatreduce.cpp:882-909)."""
    L = 6
    p = write_input(str(tmp_path), axialdir=6, size=L, savevec="1 1", out="M", J=0.45,
                    kappa="0.3 0")
    r = run([p], str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    D = oracle.OracleOp(oracle.YAM, L, 0, 0, 0.45, 0.3, 0.0, 0.0)
    T = D.dense()
    w = np.linalg.eigvals(T)
    lam = w[np.argmax(np.abs(w))].real
    re_, im_ = open(tmp_path / "M_val.dat").read().strip("()").split(",")
    assert abs(float(re_) / (lam / D.normfactor) - 1) < 1e-10
    vr = np.fromfile(tmp_path / "M_vec.dat", dtype=np.float64)
    vl = np.fromfile(tmp_path / "M_vecl.dat", dtype=np.float64)
    assert vr.size == vl.size == D.n
    assert np.linalg.norm(T @ vr - lam * vr) / abs(lam) < 1e-8
    assert np.linalg.norm(T.T @ vl - lam * vl) / abs(lam) < 1e-8
    mom = [float(x) for x in open(tmp_path / "M_phys.dat").read().split("\t")]
    wgt = vl * vr
    m = 2 * np.array([bin(int(v) >> L).count("1") for v in D.reps()], float) - L
    ref = [float(np.sum(m ** k * wgt) / (L * wgt.sum())) for k in range(1, 5)]
    np.testing.assert_allclose(mom, ref, rtol=1e-8, atol=1e-10)
