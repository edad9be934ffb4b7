"""Host code under sanitizers, on the CPU.

tests/native builds the library's host code (the eigen driver tm_solver.cpp,
the buffer/copy/stream layer tm_runtime.cpp, the C ABI tm_api.cpp) against a
deferred-execution stand-in of the HIP runtime: kernels and asynchronous
copies run only when the host waits, other streams progress in between,
device-to-host destinations read as NaN until the wait, and graph capture
records closures that replays run with their captured pointers.  Its
scenarios run plain, with AddressSanitizer + UndefinedBehaviorSanitizer and
with ThreadSanitizer.  Each scenario is a regression test of a defect found
this way or by review:

  graph_reuse        repeated solves of one operator (cached graphs, a larger
                     basis, mixed precision) give the dense spectrum
  nonsym_left        non-symmetric solve and left vector
  apply_dev_streams  tm_op_apply_dev on a caller's stream is ordered against
                     the operator's own stream (they share its workspace)
  alloc_in_capture   a buffer cannot be freed or reallocated inside a capture
  tune_families      autotuner choices of different kernel families never
                     share a cache entry
  verify_rejects     the explicit FP64 residual check rejects a wrong pair
  refine_unmet       an FP64 refinement that meets neither its residual nor
                     its error bound falls back to FP64 Krylov-Schur
  batch              concurrent solves from host threads (tm_solve_batch)

The oracle (oracle/selftest.c over every operator kind) runs under the same
sanitizers; its ThreadSanitizer build uses LLVM's OpenMP runtime with the
Archer tool so that the runtime's synchronisation is visible."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")
SCENARIOS = ["graph_reuse", "nonsym_left", "apply_dev_streams", "alloc_in_capture",
             "tune_families", "verify_rejects", "refine_unmet", "batch", "block"]
BUILDS = ["driver_test", "driver_test_asan", "driver_test_tsan"]


@pytest.fixture(scope="module")
def native():
    subprocess.run(["make", "-s", "-j3", "-C", NATIVE], check=True, capture_output=True)
    return NATIVE


@pytest.mark.parametrize("build", BUILDS)
def test_driver_scenarios(native, build):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(native, build), "--scenarios"], capture_output=True,
                       text=True, timeout=600, env=env)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    ok = {ln.split()[1] for ln in r.stdout.splitlines() if ln.startswith("ok ")}
    assert ok == set(SCENARIOS)
    assert "Sanitizer" not in r.stderr


@pytest.fixture(scope="module")
def oracle_selftests():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, capture_output=True)
    return ORACLE


@pytest.mark.parametrize("build", ["selftest", "selftest_asan", "selftest_tsan"])
def test_oracle_sanitized(oracle_selftests, build):
    env = dict(os.environ, OMP_NUM_THREADS="4", TSAN_OPTIONS=(
        "halt_on_error=1 suppressions=" + os.path.join(ORACLE, "tsan.supp")))
    if build == "selftest_tsan":
        env["OMP_TOOL_LIBRARIES"] = "/opt/rocm/llvm/lib/libarcher.so"
    r = subprocess.run([os.path.join(oracle_selftests, build)], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert r.stdout.strip().endswith("selftest ok")
    assert "Sanitizer" not in r.stderr
