"""GPU parity of tm_sample_configs (dumpconfig's configuration chains as
device Markov chains) against the oracle's sequential restatement of the
dump() of every DumpConfig model (XAT, YAT: DumpConfig/dumpconfig.cpp:88-222;
XAM, YAM, YBM, ZBM, XDM: DumpConfig/dumpreduce.cpp:47-590): given the same
uniforms, every state of every chain is the same
(integer output, compared exactly; the weights are summed in a different
order, so a draw could differ only when a uniform falls within rounding of a
cumulative boundary).  Cases: eigenvectors from the device solver and random
signed vectors, periodic / open / unfolded XAM, XDM, and the benchmark
widths XAT L=24 and XAM L=28 for a few steps; plus the dumpconfig program's
files."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import latttmat_amd as m
    return m


def _rvec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.05, 1.0, n) * np.where(rng.uniform(size=n) < 0.3, -1.0, 1.0)


# (label, TransferMatrix args, kwargs, oracle kind, oracle bc, vector source, nchain, length)
CASES = [
    ("XAT-L10-eig", ("A", 2, 1, 10), {}, "XAT", 0, "eig", 3, 25),
    ("XAT-L18-rand", ("A", 2, 1, 18), {"h": 0.1}, "XAT", 0, "rand", 2, 12),
    ("XAM-L14-eig", ("A", 2, 5, 14), {}, "XAM2D", 0, "eig", 3, 20),
    ("XAM-L20-rand", ("A", 2, 5, 20), {}, "XAM2D", 0, "rand", 2, 10),
    ("XAM-L12-open", ("A", 2, 13, 12), {}, "XAM2D", 1, "rand", 2, 15),
    ("XAM-L12-unfolded", ("A", 2, 5, 12), {"neigs": 4}, "XAM2D", 2, "rand", 2, 15),
    ("XDM-L12-eig", ("D", 2, 5, 12), {}, "XDM", 0, "eig", 3, 20),
    ("XDM-L16-rand", ("D", 2, 5, 16), {"h": 0.05}, "XDM", 0, "rand", 2, 10),
    ("YAT-L6-eig", ("A", 2, 2, 6), {}, "YAT", 0, "eig", 3, 20),
    ("YAT-L9-rand", ("A", 2, 2, 9), {"h": 0.1}, "YAT", 0, "rand", 2, 12),
    ("ZBM-L12-eig", ("B", 2, 7, 12), {"kappa2": 0.25}, "ZBM", 0, "eig", 3, 15),
    ("ZBM-L16-rand", ("B", 2, 7, 16), {"kappa2": 0.25, "h": 0.1}, "ZBM", 0, "rand", 2, 10),
    ("YAM-L6-eig", ("A", 2, 6, 6), {}, "YAM", 0, "eig", 3, 15),
    ("YAM-L8-rand", ("A", 2, 6, 8), {}, "YAM", 0, "rand", 2, 10),
    ("YBM-L7-rand", ("B", 2, 6, 7), {"kappa2": 0.25, "h": 0.1}, "YBM", 0, "rand", 2, 10),
]


@pytest.mark.parametrize("label,args,kw,okind,obc,src,nchain,length", CASES,
                         ids=[c[0] for c in CASES])
def test_chains_vs_oracle(lt, oracle, label, args, kw, okind, obc, src, nchain, length):
    kw = dict(kw)
    bJ, k, h = 0.45, 0.35, kw.pop("h", 0.0)
    k2 = kw.pop("kappa2", 0.0)
    op = lt.TransferMatrix(*args, J=bJ, kappa=k, kappa2=k2, h=h, **kw)
    assert op.kind() == okind
    oop = oracle.OracleOp(getattr(oracle, okind), args[3], 0, obc, bJ, k, k2, h)
    n = op.cols()
    assert n == oop.n
    sym = op.is_symmetric()
    vl = None
    if src == "eig":
        r = lt.solve(op, nev=1, want_vectors=True, want_left=not sym)
        v = np.real(r.vecs[:, 0])
        if not sym:
            vl = np.asarray(r.vecl)
    else:
        v = _rvec(n, len(label))
        if not sym:
            vl = _rvec(n, 2 * len(label))
    u = np.random.default_rng(100 + len(label)).uniform(0, 1, op.sample_uniforms_needed(nchain, length))
    g = op.sample_configs(v, u, nchain, length, vl)
    o = oop.sample_configs(v, u, nchain, length, vl)
    print(f"  {label}: n={n} first chain {g[0][:6]}", flush=True)
    np.testing.assert_array_equal(g, o)
    nb = 2 * args[3] if okind in ("YAT", "YAM", "YBM") else args[3]
    assert np.all(g < (1 << nb))


@pytest.mark.parametrize("label,args,okind,length", [("XAT-L24", ("A", 2, 1, 24), "XAT", 6),
                                                     ("XAM-L28", ("A", 2, 5, 28), "XAM2D", 3)])
def test_chains_full_width(lt, oracle, label, args, okind, length):
    bJ, k = 0.4, 0.5
    op = lt.TransferMatrix(*args, J=bJ, kappa=k)
    oop = oracle.OracleOp(getattr(oracle, okind), args[3], 0, 0, bJ, k, 0.0, 0.0)
    v = _rvec(op.cols(), 9)
    u = np.random.default_rng(11).uniform(0, 1, op.sample_uniforms_needed(1, length))
    print(f"  {label}: sampling on the device", flush=True)
    g = op.sample_configs(v, u, 1, length)
    print(f"  {label}: device {g[0]}; oracle (sequential, seconds)", flush=True)
    o = oop.sample_configs(v, u, 1, length)
    np.testing.assert_array_equal(g, o)


def test_sampler_rejects_unsupported(lt):
    op = lt.TransferMatrix("B", 2, 3, 10, J=0.4, kappa=0.3, kappa2=0.2)
    with pytest.raises(lt.TMError):
        op.sample_configs(np.ones(op.cols()), np.full(100, 0.5), 1, 5, np.ones(op.cols()))
    # the non-symmetric samplers need the left vector
    op = lt.TransferMatrix("B", 2, 7, 10, J=0.4, kappa=0.3, kappa2=0.2)
    with pytest.raises(lt.TMError):
        op.sample_configs(np.ones(op.cols()), np.full(100, 0.5), 1, 5)


def _write_input(d, axialdir, size, out, model="A"):
    p = os.path.join(d, "input.dat")
    kap = "0.35 0.25" if model == "B" else "0.35"
    with open(p, "w") as f:
        f.write(f"model = {model}\ndim = 2\naxialdir = {axialdir}\nneigs = 1\nsavevec = 1 1\n"
                f"outprefix = {out}\nsize = {size}\nJ = 0.45\nkappa = {kap}\nh = 0\n")
    return p


@pytest.mark.parametrize("model,axialdir,size", [("A", 1, 10), ("A", 5, 14), ("D", 5, 12),
                                                 ("A", 2, 6), ("A", 6, 6), ("B", 6, 6),
                                                 ("B", 7, 12)])
def test_dumpconfig_program(lt, tmp_path, model, axialdir, size):
    """tmat writes X_vec.dat, dumpconfig plants N chains of the given length
    (DumpConfig/main.cpp:30-144): N files of `length` lines of `size` bits;
    the same seed gives the same files, and every line is a layer state."""
    here = os.path.dirname(lt.LIB_PATH)
    inp = _write_input(str(tmp_path), axialdir, size, "P", model)
    r = subprocess.run([os.path.join(here, "tmat"), inp], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    outs = []
    for rep in range(2):
        r = subprocess.run([os.path.join(here, "dumpconfig"), inp, "3", "40", "--seed", "7"],
                           cwd=tmp_path, capture_output=True, text=True, timeout=120)
        print(r.stdout[-300:], r.stderr[-300:], flush=True)
        assert r.returncode == 0, r.stderr
        files = [open(tmp_path / f"P_config_{i}.dat").read() for i in range(3)]
        outs.append(files)
    assert outs[0] == outs[1]
    nlines = 41 if (model, axialdir) == ("A", 2) else 40  # YAT adds the first low layer
    for txt in outs[0]:
        lines = txt.split("\n")
        assert lines[-1] == "" and len(lines) == nlines + 1
        for ln in lines[:-1]:
            assert len(ln) == size and set(ln) <= {"0", "1"}
    # the same chains through the C ABI with the program's draws (std::mt19937
    # seeded 7, restated in tests/std_random.py) and its output lines
    from tests import std_random
    kap2 = 0.25 if model == "B" else 0.0
    op = lt.TransferMatrix(model, 2, axialdir, size, J=0.45, kappa=0.35, kappa2=kap2)
    n = op.cols()
    vec = np.fromfile(tmp_path / "P_vec.dat", dtype=np.float64)[:n]
    vecl = None if op.is_symmetric() else np.fromfile(tmp_path / "P_vecl.dat", dtype=np.float64)[:n]
    u = np.array(std_random.uniforms(7, op.sample_uniforms_needed(3, 40)))
    st = op.sample_configs(vec, u, 3, 40, vecl)
    lm = (1 << size) - 1
    kind = op.kind()
    for c in range(3):
        if kind == "YAT":
            want = [int(st[c, 0]) & lm] + [int(x) >> size for x in st[c]]
        elif kind in ("YAM", "YBM"):
            want = [int(x) & lm for x in st[c]]
        else:
            want = [int(x) for x in st[c]]
        text = "".join("".join(str((s >> b) & 1) for b in range(size)) + "\n" for s in want)
        assert outs[0][c] == text
