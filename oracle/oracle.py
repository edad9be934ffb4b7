"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product package ``latttmat_amd`` never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

XAT, XDT, ZBT, XAM2D, XAM3D, XAT3D, XDM, ZBM, YAT, Y2AT, YAM, YBM = range(1, 13)

_lib = None


def build() -> None:
    """Compile liboracle.so with the committed Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run oracle.build()")
        L = C.CDLL(_LIB_PATH)
        u64, i64, dbl, i32 = C.c_uint64, C.c_int64, C.c_double, C.c_int
        pdbl = C.POINTER(C.c_double)
        L.or_offset_state.restype = u64
        L.or_offset_state.argtypes = [u64, i32, i32]
        L.or_lffset_state.restype = u64
        L.or_lffset_state.argtypes = [u64, i32, i32]
        L.or_bitreversal.restype = u64
        L.or_bitreversal.argtypes = [u64, i32]
        L.or_spinbzexp.restype = i32
        L.or_spinbzexp.argtypes = [u64, u64, i32]
        L.or_powdi.restype = dbl
        L.or_powdi.argtypes = [pdbl, i32]
        L.or_rffset3d.restype = u64
        L.or_rffset3d.argtypes = [u64, i32, i32, i32]
        L.or_dffset3d.restype = u64
        L.or_dffset3d.argtypes = [u64, i32, i32, i32]
        L.or_rotate3d.restype = u64
        L.or_rotate3d.argtypes = [u64, i32]
        L.or_reflect3d.restype = u64
        L.or_reflect3d.argtypes = [u64, i32, i32, i32]
        L.or_op_create.restype = C.c_void_p
        L.or_op_create.argtypes = [i32, i32, i32, i32, dbl, dbl, dbl, dbl]
        L.or_op_destroy.restype = None
        L.or_op_destroy.argtypes = [C.c_void_p]
        L.or_op_apply.restype = None
        L.or_op_apply.argtypes = [C.c_void_p, pdbl, pdbl, i32]
        L.or_op_ncol.restype = i64
        L.or_op_ncol.argtypes = [C.c_void_p]
        L.or_op_normfactor.restype = dbl
        L.or_op_normfactor.argtypes = [C.c_void_p]
        L.or_op_tvec.restype = pdbl
        L.or_op_tvec.argtypes = [C.c_void_p]
        L.or_op_reps.restype = C.POINTER(C.c_uint64)
        L.or_op_reps.argtypes = [C.c_void_p]
        L.or_op_dups.restype = C.POINTER(C.c_uint16)
        L.or_op_dups.argtypes = [C.c_void_p]
        L.or_orbits_x2d.restype = i64
        L.or_orbits_x2d.argtypes = [i32, i32, i32, C.c_void_p, C.c_void_p]
        L.or_orbits_z2d.restype = i64
        L.or_orbits_z2d.argtypes = [i32, i32, C.c_void_p, C.c_void_p]
        L.or_orbits_y2d.restype = i64
        L.or_orbits_y2d.argtypes = [i32, i32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_orbits_x3d.restype = i64
        L.or_orbits_x3d.argtypes = [i32, i32, i32, i32, C.c_void_p, C.c_void_p]
        L.or_tvec_xat.restype = None
        L.or_tvec_xat.argtypes = [i32, i32, dbl, dbl, dbl, pdbl]
        L.or_tvec_xdt.restype = None
        L.or_tvec_xdt.argtypes = [i32, i32, dbl, dbl, pdbl]
        L.or_tvec_zbt.restype = None
        L.or_tvec_zbt.argtypes = [i32, dbl, dbl, dbl, pdbl]
        L.or_op_moments.restype = None
        L.or_op_moments.argtypes = [C.c_void_p, C.c_void_p, pdbl, pdbl, C.POINTER(i32)]
        L.or_sample_configs.restype = i32
        L.or_sample_configs.argtypes = [C.c_void_p, pdbl, C.c_void_p, pdbl, i32, i32, C.c_void_p]
        L.or_num_threads.restype = i32
        L.or_num_threads.argtypes = []
        L.or_set_num_threads.restype = None
        L.or_set_num_threads.argtypes = [i32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleOp:
    """A reference-algorithm operator on the CPU (FP64)."""

    def __init__(self, kind, size1, size2=0, bc=0, betaJ=0.5, kappa=0.0,
                 kappa2=0.0, h=0.0):
        self.kind = kind
        self.size1, self.size2 = size1, size2
        self._h = lib().or_op_create(kind, size1, size2, bc, betaJ, kappa, kappa2, h)
        if not self._h:
            raise ValueError("oracle rejected operator parameters")
        self.n = int(lib().or_op_ncol(self._h))
        self.normfactor = float(lib().or_op_normfactor(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_op_destroy(h)
            self._h = None

    def apply(self, x, transpose=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.n,)
        y = np.empty_like(x)
        lib().or_op_apply(self._h, _p(x), _p(y), 1 if transpose else 0)
        return y

    def dense(self, transpose=False):
        """raw_Tmat-style extraction (TMat/tmat.hpp:90-111) without /normfactor."""
        n = self.n
        M = np.empty((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[:] = 0.0
            e[j] = 1.0
            M[:, j] = self.apply(e, transpose)
        return M

    def tvec(self):
        return np.ctypeslib.as_array(lib().or_op_tvec(self._h), shape=(self.n,)).copy()

    def sample_configs(self, vec, u, nchain, length, vecl=None):
        """dumpconfig's chains with the given draws: (nchain, length) states."""
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert vec.shape == (self.n,)
        if vecl is not None:
            vecl = np.ascontiguousarray(vecl, dtype=np.float64)
            assert vecl.shape == (self.n,)
        out = np.empty((nchain, length), np.uint64)
        if lib().or_sample_configs(self._h, _p(vec), None if vecl is None else vecl.ctypes.data,
                                   _p(u), nchain, length, out.ctypes.data):
            raise ValueError("no sampler for this operator kind")
        return out

    def moments(self, qr, ql=None):
        """X_phys.dat figures: four moments, then the extra observables."""
        qr = np.ascontiguousarray(qr, dtype=np.float64)
        assert qr.shape == (self.n,)
        if ql is not None:
            ql = np.ascontiguousarray(ql, dtype=np.float64)
            assert ql.shape == (self.n,)
        out = np.empty(8)
        nout = C.c_int(0)
        lib().or_op_moments(self._h, None if ql is None else ql.ctypes.data, _p(qr), _p(out),
                            C.byref(nout))
        return out[:nout.value].copy()

    def reps(self):
        p = lib().or_op_reps(self._h)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy() if p else None

    def dups(self):
        p = lib().or_op_dups(self._h)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy() if p else None


def orbits_x2d(L, bc=0, censym=0):
    m = lib().or_orbits_x2d(L, bc, censym, None, None)
    reps = np.empty(m, np.uint64)
    dups = np.empty(m, np.uint16)
    lib().or_orbits_x2d(L, bc, censym, reps.ctypes.data, dups.ctypes.data)
    return reps, dups


def orbits_z2d(L, censym=0):
    m = lib().or_orbits_z2d(L, censym, None, None)
    reps = np.empty(m, np.uint64)
    dups = np.empty(m, np.uint16)
    lib().or_orbits_z2d(L, censym, reps.ctypes.data, dups.ctypes.data)
    return reps, dups


def orbits_y2d(L, censym=0):
    m = lib().or_orbits_y2d(L, censym, None, None, None)
    reps = np.empty(m, np.uint64)
    dups = np.empty(m, np.uint16)
    swp = np.empty(m, np.int64)
    lib().or_orbits_y2d(L, censym, reps.ctypes.data, dups.ctypes.data, swp.ctypes.data)
    return reps, dups, swp


def orbits_x3d(s1, s2, censym=1, rotate=0):
    m = lib().or_orbits_x3d(s1, s2, censym, rotate, None, None)
    reps = np.empty(m, np.uint64)
    dups = np.empty(m, np.uint16)
    lib().or_orbits_x3d(s1, s2, censym, rotate, reps.ctypes.data, dups.ctypes.data)
    return reps, dups


def tvec_xat(L, bc, betaJ, kappa, h):
    V = np.empty(1 << L)
    lib().or_tvec_xat(L, bc, betaJ, kappa, h, _p(V))
    return V


def num_threads():
    return int(lib().or_num_threads())


def set_num_threads(n):
    lib().or_set_num_threads(int(n))
