"""latttmat_amd -- MI355X-native transfer-matrix eigensolver (Python binding).

The compute path is the native library ``liblatttmat.so`` (hand-written HIP
kernels for gfx950 behind the C ABI declared in ``include/latttmat.h``).
This module is a thin ctypes binding used by the tests and the benchmark;
it has no fallback: if the library is not built, importing the binding
raises, and creating an operator on a machine without a GPU returns the
library's TM_EHIP error.

The operator object mirrors the reference's Spectra operator concept
(``rows()``, ``cols()``, ``perform_op(x, y)``, TMat/tmat.hpp:86-89) and the
solve mirrors ``pfunc_inner`` (TMat/pfunc.hpp:79-202).
"""
from ._abi import (  # noqa: F401
    LIB_PATH,
    TM_EHIP,
    TM_EINVAL,
    TM_ENOCONV,
    TM_ENOMEM,
    TM_EUNSUPPORTED,
    TM_OK,
    TM_F32,
    TM_F64,
    TM_FLAG_FULLSPACE,
    TM_FLAG_GENERIC_REDUCED,
    TM_FLAG_NO_AUTOTUNE,
    TM_PREC_F64,
    TM_PREC_MIXED,
    TMError,
    TransferMatrix,
    SolveResult,
    exported_symbols,
    lib,
    solve,
    solve_batch,
)

__all__ = [
    "LIB_PATH", "TM_OK", "TM_EINVAL", "TM_ENOMEM", "TM_EHIP", "TM_ENOCONV", "TM_EUNSUPPORTED", "TM_F32", "TM_F64", "TM_FLAG_FULLSPACE", "TM_FLAG_GENERIC_REDUCED", "TM_FLAG_NO_AUTOTUNE",
    "TM_PREC_F64", "TM_PREC_MIXED", "TMError", "TransferMatrix", "SolveResult",
    "exported_symbols", "lib", "solve", "solve_batch",
]
