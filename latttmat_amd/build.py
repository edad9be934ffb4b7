"""Build the native library ``liblatttmat.so`` (HIP, gfx950) and the ``tmat``
and ``dumpconfig`` command-line programs in-tree with hipcc.  Called by ``__graft_entry__.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblatttmat.so")
CLI = os.path.join(HERE, "tmat")
DUMPCONFIG = os.path.join(HERE, "dumpconfig")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(HERE, "..", "include"),
]

LIB_SOURCES = [
    "tm_api.cpp", "tm_runtime.cpp", "tm_ops.cpp", "tm_solver.cpp", "tm_dense.cpp", "tm_grad.hip",
    "tm_kron.hip", "tm_reduced.hip", "tm_sweep.hip", "tm_krylov.hip", "tm_axial.hip",
    "tm_phys.hip", "tm_sample.hip", "tm_noalloc.hip",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(LIB_SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for src, exe in (("tmat_main.cpp", CLI), ("dumpconfig_main.cpp", DUMPCONFIG)):
        tmpc = exe + ".tmp"
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(CSRC, src), "-o", tmpc, "-L", HERE, "-llatttmat",
               "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{os.path.basename(exe)} build failed\n{r.stdout}\n{r.stderr}")
        os.replace(tmpc, exe)
    if verbose:
        print(f"built {LIB}")


if __name__ == "__main__":
    build()
    sys.exit(0)
