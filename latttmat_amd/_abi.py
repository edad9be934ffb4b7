"""ctypes mirror of include/latttmat.h."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblatttmat.so")

TM_OK, TM_EINVAL, TM_ENOMEM, TM_EHIP, TM_ENOCONV, TM_EUNSUPPORTED = range(6)
TM_F64, TM_F32 = 0, 1
TM_PREC_F64, TM_PREC_MIXED = 0, 1
TM_FLAG_FULLSPACE = 1
TM_FLAG_GENERIC_REDUCED = 2
TM_FLAG_NO_AUTOTUNE = 4
KIND_NAMES = {1: "XAT", 2: "XDT", 3: "ZBT", 4: "XAM2D", 5: "XAM3D", 6: "XAT3D", 7: "XDM",
              8: "ZBM", 9: "YAT", 10: "Y2AT", 11: "YAM", 12: "YBM"}

# every function the header declares (checked by the CPU tests)
EXPORTS = [
    "tm_op_create", "tm_op_destroy", "tm_op_rows", "tm_op_cols", "tm_op_normfactor",
    "tm_op_is_symmetric", "tm_op_kind", "tm_op_suggested_ncv", "tm_op_layer_size",
    "tm_op_matvec_count", "tm_op_apply", "tm_op_apply_dev", "tm_op_apply_dev_noalloc",
    "tm_op_apply_prec",
    "tm_op_time_apply", "tm_op_orbits", "tm_op_moments", "tm_sample_configs",
    "tm_solve_opts_default", "tm_solve", "tm_solve_batch", "tm_grad_hf", "tm_last_error",
    "tm_version",
]


class TMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("model", C.c_char), ("dim", C.c_int), ("axialdir", C.c_int), ("neigs", C.c_int),
        ("size1", C.c_int), ("size2", C.c_int), ("betaJ", C.c_double), ("kappa", C.c_double),
        ("kappa2", C.c_double), ("h", C.c_double), ("flags", C.c_int),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("nev", C.c_int), ("ncv", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
        ("maxtrial", C.c_int), ("precision", C.c_int), ("seed", C.c_uint64),
        ("want_left", C.c_int), ("want_vectors", C.c_int), ("mixed_tol", C.c_double),
        ("verbose", C.c_int), ("refine", C.c_int), ("use_graph", C.c_int),
        ("basis_f32", C.c_int), ("refine_extra", C.c_int), ("verify", C.c_int),
        ("block", C.c_int),
    ]


class SolveOut(C.Structure):
    _fields_ = [
        ("nconv", C.c_int), ("evals_re", C.POINTER(C.c_double)),
        ("evals_im", C.POINTER(C.c_double)), ("resid", C.POINTER(C.c_double)),
        ("vecs", C.POINTER(C.c_double)), ("vecl", C.POINTER(C.c_double)),
        ("eval_left", C.c_double), ("matvecs_f64", C.c_int64), ("matvecs_f32", C.c_int64),
        ("restarts", C.c_int64), ("trials", C.c_int), ("seconds", C.c_double),
        ("graph_cycles", C.c_int64), ("resid_check", C.c_double),
    ]


_lib = None


def lib():
    """Load liblatttmat.so (raises if it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch wheels bundle their own HIP runtime under the same soname
        # (libamdhip64.so.7); if torch is used in this process it must be
        # loaded first so that only one runtime owns the device.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
        pd = C.POINTER(C.c_double)
        L.tm_op_create.restype = i32
        L.tm_op_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.tm_op_destroy.restype = None
        L.tm_op_destroy.argtypes = [vp]
        for nm in ("tm_op_rows", "tm_op_cols", "tm_op_matvec_count"):
            getattr(L, nm).restype = i64
            getattr(L, nm).argtypes = [vp]
        L.tm_op_normfactor.restype = dbl
        L.tm_op_normfactor.argtypes = [vp]
        for nm in ("tm_op_is_symmetric", "tm_op_kind", "tm_op_layer_size"):
            getattr(L, nm).restype = i32
            getattr(L, nm).argtypes = [vp]
        L.tm_op_suggested_ncv.restype = i32
        L.tm_op_suggested_ncv.argtypes = [vp, i32]
        L.tm_op_apply.restype = i32
        L.tm_op_apply.argtypes = [vp, pd, pd, i32]
        L.tm_op_apply_dev.restype = i32
        L.tm_op_apply_dev.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, vp]
        L.tm_op_apply_dev_noalloc.restype = i32
        L.tm_op_apply_dev_noalloc.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, vp]
        L.tm_op_apply_prec.restype = i32
        L.tm_op_apply_prec.argtypes = [vp, pd, pd, i32, i32]
        L.tm_op_time_apply.restype = i32
        L.tm_op_time_apply.argtypes = [vp, i32, i32, i32, i32, i32, pd]
        L.tm_op_orbits.restype = i32
        L.tm_op_orbits.argtypes = [vp, vp, vp]
        L.tm_op_moments.restype = i32
        L.tm_op_moments.argtypes = [vp, pd, pd, pd, C.POINTER(i32)]
        L.tm_sample_configs.restype = i32
        L.tm_sample_configs.argtypes = [vp, pd, pd, pd, i32, i32, vp]
        L.tm_solve_opts_default.restype = None
        L.tm_solve_opts_default.argtypes = [C.POINTER(SolveOpts)]
        L.tm_solve.restype = i32
        L.tm_solve.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(SolveOut)]
        L.tm_solve_batch.restype = i32
        L.tm_solve_batch.argtypes = [vp, i32, C.POINTER(SolveOpts), C.POINTER(SolveOut), i32]
        L.tm_grad_hf.restype = i32
        L.tm_grad_hf.argtypes = [vp, pd, pd, pd, pd]
        L.tm_last_error.restype = C.c_char_p
        L.tm_last_error.argtypes = []
        L.tm_version.restype = C.c_char_p
        L.tm_version.argtypes = []
        _lib = L
    return _lib


def exported_symbols():
    """Names from EXPORTS that the loaded library resolves."""
    L = lib()
    out = []
    for nm in EXPORTS:
        try:
            getattr(L, nm)
            out.append(nm)
        except AttributeError:
            pass
    return out


def _check(st):
    if st != TM_OK:
        raise TMError(st, lib().tm_last_error().decode())


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class TransferMatrix:
    """Spectra-style operator over the C ABI: rows(), cols(), perform_op(x, y).

    Arguments follow the tmat input keys (TMat/read_input.cpp:29-60):
    model, dim, axialdir, neigs, size (int, or (size1, size2) in 3D),
    J (= betaJ), kappa, kappa2, h.
    """

    def __init__(self, model="A", dim=2, axialdir=1, size=12, J=0.5, kappa=0.0,
                 kappa2=0.0, h=0.0, neigs=1, flags=0):
        if isinstance(size, (tuple, list)):
            s1, s2 = int(size[0]), int(size[1])
        else:
            s1, s2 = int(size), 0
        self.params = Params(model.encode() if isinstance(model, str) else model, dim, axialdir,
                             neigs, s1, s2, J, kappa, kappa2, h, flags)
        h_ = C.c_void_p()
        _check(lib().tm_op_create(C.byref(self.params), C.byref(h_)))
        self._h = h_

    def close(self):
        if getattr(self, "_h", None):
            lib().tm_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- the reference operator surface ---
    def rows(self):
        return int(lib().tm_op_rows(self._h))

    def cols(self):
        return int(lib().tm_op_cols(self._h))

    def getnormfactor(self):
        return float(lib().tm_op_normfactor(self._h))

    def is_symmetric(self):
        return bool(lib().tm_op_is_symmetric(self._h))

    def kind(self):
        return KIND_NAMES.get(int(lib().tm_op_kind(self._h)), "?")

    def getLayerSize(self):
        return int(lib().tm_op_layer_size(self._h))

    def matvec_count(self):
        return int(lib().tm_op_matvec_count(self._h))

    def suggested_nev(self, neigs):
        return int(lib().tm_op_suggested_ncv(self._h, neigs))

    def perform_op(self, x, y, transpose=False):
        assert x.dtype == np.float64 and y.dtype == np.float64
        assert x.flags.c_contiguous and y.flags.c_contiguous
        assert x.shape == (self.cols(),) and y.shape == (self.rows(),)
        _check(lib().tm_op_apply(self._h, _pd(x), _pd(y), 1 if transpose else 0))

    def apply(self, x, transpose=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.rows())
        self.perform_op(x, y, transpose)
        return y

    def apply_prec(self, x, compute=TM_F64, transpose=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.rows())
        _check(lib().tm_op_apply_prec(self._h, _pd(x), _pd(y), 1 if transpose else 0, compute))
        return y

    def time_apply(self, io=TM_F64, compute=TM_F64, transpose=False, warmup=3, reps=10):
        ms = np.zeros(1)
        _check(lib().tm_op_time_apply(self._h, io, compute, 1 if transpose else 0, warmup, reps,
                                      _pd(ms)))
        return float(ms[0])

    def apply_dev(self, x_ptr, y_ptr, ld, nvec=1, dtype=TM_F64, compute=TM_F64,
                  transpose=False, stream=None):
        _check(lib().tm_op_apply_dev(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), ld, nvec,
                                     dtype, compute, 1 if transpose else 0,
                                     C.c_void_p(stream) if stream else None))

    def apply_dev_noalloc(self, x_ptr, y_ptr, ld, nvec=1, dtype=TM_F64, compute=TM_F64,
                          transpose=False, stream=None):
        _check(lib().tm_op_apply_dev_noalloc(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), ld,
                                             nvec, dtype, compute, 1 if transpose else 0,
                                             C.c_void_p(stream) if stream else None))

    def orbits(self):
        n = self.cols()
        reps = np.empty(n, np.uint64)
        dups = np.empty(n, np.uint16)
        _check(lib().tm_op_orbits(self._h, reps.ctypes.data, dups.ctypes.data))
        return reps, dups

    def moments(self, qr, ql=None):
        """X_phys.dat figures of an eigenvector pair, reduced on the GPU: the
        four magnetisation moments, then the kind's extra observables."""
        qr = np.ascontiguousarray(qr, dtype=np.float64)
        if ql is not None:
            ql = np.ascontiguousarray(ql, dtype=np.float64)
        out = np.empty(8)
        nout = C.c_int(0)
        _check(lib().tm_op_moments(self._h, None if ql is None else _pd(ql), _pd(qr), _pd(out),
                                   C.byref(nout)))
        return out[:nout.value].copy()

    def sample_uniforms_needed(self, nchain, length):
        """Uniform draws one tm_sample_configs call consumes."""
        per = length if self.kind() in ("XAT", "YAT") else 2 * length - 1
        return nchain * per

    def sample_configs(self, vec, u, nchain, length, vecl=None):
        """Configuration chains planted from the eigenvector vec (and the left
        eigenvector vecl for the non-symmetric kinds) with the uniforms u
        (dumpconfig's draws, in order): (nchain, length) states."""
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert vec.shape == (self.cols(),)
        assert u.size >= self.sample_uniforms_needed(nchain, length)
        if vecl is not None:
            vecl = np.ascontiguousarray(vecl, dtype=np.float64)
            assert vecl.shape == (self.cols(),)
        out = np.empty((nchain, length), np.uint64)
        _check(lib().tm_sample_configs(self._h, _pd(vec), None if vecl is None else _pd(vecl),
                                       _pd(u), nchain, length, out.ctypes.data))
        return out

    def grad_hf(self, l, r):
        l = np.ascontiguousarray(l, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        dlam = np.empty(4)
        dlnnf = np.empty(4)
        _check(lib().tm_grad_hf(self._h, _pd(l), _pd(r), _pd(dlam), _pd(dlnnf)))
        return dlam, dlnnf


@dataclass
class SolveResult:
    status: int
    nconv: int
    evals: np.ndarray
    resid: np.ndarray
    vecs: np.ndarray | None
    vecl: np.ndarray | None
    eval_left: float
    matvecs_f64: int
    matvecs_f32: int
    restarts: int
    trials: int
    seconds: float
    extra: dict = field(default_factory=dict)


def _make_opts(nev, ncv, tol, maxit, precision, seed, want_left, want_vectors, mixed_tol,
               verbose, maxtrial, refine, use_graph, basis_f32=True, refine_extra=1, verify=True,
               block=0):
    o = SolveOpts()
    lib().tm_solve_opts_default(C.byref(o))
    o.nev, o.ncv, o.tol, o.maxit, o.maxtrial = nev, ncv, tol, maxit, maxtrial
    o.precision, o.seed = precision, seed
    o.want_left, o.want_vectors = int(want_left), int(want_vectors)
    o.mixed_tol, o.verbose, o.refine = mixed_tol, int(verbose), int(refine)
    o.use_graph = int(use_graph)
    o.basis_f32 = int(basis_f32)
    o.refine_extra = int(refine_extra)
    o.verify = int(verify)
    o.block = int(block)
    return o


class _OutBufs:
    """Host buffers behind one SolveOut."""

    def __init__(self, op, nev, want_left, want_vectors):
        n = op.cols()
        self.n, self.nev = n, nev
        self.sym = op.is_symmetric()
        self.er, self.ei, self.rs = np.zeros(nev), np.zeros(nev), np.zeros(nev)
        self.vecs = None
        if want_vectors:
            self.vecs = np.zeros(nev * n if self.sym else 2 * nev * n)
        self.vecl = np.zeros(n) if (want_left and not self.sym and want_vectors) else None
        out = SolveOut()
        out.evals_re, out.evals_im, out.resid = _pd(self.er), _pd(self.ei), _pd(self.rs)
        out.vecs = _pd(self.vecs) if self.vecs is not None else None
        out.vecl = _pd(self.vecl) if self.vecl is not None else None
        self.out = out

    def result(self, st):
        out, n, nev = self.out, self.n, self.nev
        evals = self.er + 1j * self.ei if not self.sym else self.er
        vecs = self.vecs
        if vecs is not None:
            if self.sym:
                vecs = vecs.reshape(nev, n).T.copy()
            else:
                v = vecs.reshape(nev, n, 2)
                vecs = (v[..., 0] + 1j * v[..., 1]).T.copy()
        return SolveResult(st, out.nconv, evals, self.rs, vecs, self.vecl, out.eval_left,
                           out.matvecs_f64, out.matvecs_f32, out.restarts, out.trials,
                           out.seconds, {"graph_cycles": int(out.graph_cycles),
                                         "resid_check": float(out.resid_check)})


def solve(op: TransferMatrix, nev=1, ncv=0, tol=1e-10, maxit=1000, precision=TM_PREC_F64,
          seed=1, want_left=False, want_vectors=False, mixed_tol=2e-7, verbose=False,
          maxtrial=3, raise_on_noconv=True, refine=0, use_graph=True,
          basis_f32=True, refine_extra=1, verify=True, block=0) -> SolveResult:
    o = _make_opts(nev, ncv, tol, maxit, precision, seed, want_left, want_vectors, mixed_tol,
                   verbose, maxtrial, refine, use_graph, basis_f32, refine_extra, verify, block)
    b = _OutBufs(op, nev, want_left, want_vectors)
    st = lib().tm_solve(op._h, C.byref(o), C.byref(b.out))
    if st not in (TM_OK, TM_ENOCONV) or (st == TM_ENOCONV and raise_on_noconv):
        _check(st)
    return b.result(st)


def solve_batch(ops, nev=1, ncv=0, tol=1e-10, maxit=1000, precision=TM_PREC_F64, seed=1,
                want_left=False, want_vectors=False, mixed_tol=2e-7, maxtrial=3, refine=0,
                use_graph=True, nworkers=4, raise_on_error=True, basis_f32=True, verify=True):
    """tm_solve_batch over a list of operators with one set of options."""
    n = len(ops)
    o = _make_opts(nev, ncv, tol, maxit, precision, seed, want_left, want_vectors, mixed_tol,
                   False, maxtrial, refine, use_graph, basis_f32, 1, verify)
    opts = (SolveOpts * n)(*([o] * n))
    bufs = [_OutBufs(op, nev, want_left, want_vectors) for op in ops]
    outs = (SolveOut * n)(*[b.out for b in bufs])
    hs = (C.c_void_p * n)(*[op._h for op in ops])
    st = lib().tm_solve_batch(hs, n, opts, outs, nworkers)
    if st != TM_OK and raise_on_error:
        _check(st)
    res = []
    for i, b in enumerate(bufs):
        b.out = outs[i]
        res.append(b.result(TM_OK if outs[i].nconv >= nev else TM_ENOCONV))
    return res