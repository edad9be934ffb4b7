// CPU driver of the library's host code (tm_solver.cpp, tm_runtime.cpp,
// tm_api.cpp, tm_dense.cpp) over dense matrices, on the deferred-execution
// runtime stand-in (fake_runtime.cpp), for tests/test_driver_cpu.py.
//
//   driver_test FILE [v]   solve the dense matrix in FILE and print the results
//                          (text: n sym nev ncv prec want_left basis32, then
//                          n*n column-major)
//   driver_test --scenarios [NAME]
//                          run the built-in scenarios (or one of them) and print
//                          "ok NAME" or "FAIL NAME: why" per scenario
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <functional>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../latttmat_amd/csrc/tm_internal.hpp"

namespace ltm {
tm_status solve_impl(tm_op *op, const tm_solve_opts &o, tm_solve_out *out);
}

namespace {

tm_op *dense_op(const std::vector<double> &A, int n, bool sym)
{
  tm_op *op = new tm_op();
  op->ncol = op->nstate = n;
  op->symmetric = sym;
  TM_HIP(hipStreamCreate(&op->stream));
  op->tab64.alloc(sizeof(double) * A.size());
  ltm::copy_sync(op->tab64.get(), A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice,
                 op->stream);
  return op;
}

// Symmetric n x n matrix H diag(d) H with a random Householder reflection
// H = I - 2 u u^T / (u^T u): dense, with the known spectrum d_i =
// sign_i 10 q^i (transfer-matrix-like geometric decay, one negative value).
std::vector<double> random_sym(int n, uint64_t seed, std::vector<double> *spec, double q = 0.8,
                               const std::vector<double> *given = nullptr)
{
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> v(n), d(n), A((size_t)n * n);
  double uu = 0;
  for (auto &x : v) {
    x = u(g);
    uu += x * x;
  }
  for (int i = 0; i < n; ++i) d[i] = (i == 1 ? -10.0 : 10.0) * std::pow(q, i);
  if (given) d = *given;
  // A = (I - c v v^T) D (I - c v v^T), c = 2 / uu
  const double c = 2.0 / uu;
  std::vector<double> Dv(n);
  double vDv = 0;
  for (int i = 0; i < n; ++i) {
    Dv[i] = d[i] * v[i];
    vDv += v[i] * Dv[i];
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      A[i + (size_t)j * n] = (i == j ? d[i] : 0.0) - c * (Dv[i] * v[j] + v[i] * Dv[j]) +
                             c * c * vDv * v[i] * v[j];
  if (spec) {
    *spec = d;
    std::sort(spec->begin(), spec->end(),
              [](double a, double b) { return std::fabs(a) > std::fabs(b); });
  }
  return A;
}

std::vector<double> random_pos(int n, uint64_t seed)
{
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(0.1, 1.0);
  std::vector<double> A((size_t)n * n);
  for (auto &v : A) v = u(g);
  return A;
}

double lead_gen_ref(const std::vector<double> &A, int n)
{
  std::vector<double> wr(n), wi(n), Zr((size_t)n * n), Zi((size_t)n * n);
  ltm::gen_eig(n, A.data(), wr.data(), wi.data(), Zr.data(), Zi.data());
  double best = 0, lam = 0;
  for (int i = 0; i < n; ++i)
    if (std::hypot(wr[i], wi[i]) > best) best = std::hypot(wr[i], wi[i]), lam = wr[i];
  return lam;
}

struct Solve {
  tm_status st;
  std::vector<double> er, ei, rs, vecs, vl;
  tm_solve_out out;
  std::string err;
};

tm_solve_opts opts(int nev, int ncv = 0, tm_precision prec = TM_PREC_F64)
{
  tm_solve_opts o;
  tm_solve_opts_default(&o);
  o.nev = nev;
  o.ncv = ncv;
  o.precision = prec;
  o.mixed_tol = 1e-6;
  return o;
}

Solve run(tm_op *op, const tm_solve_opts &o)
{
  Solve s;
  const int nev = o.nev, n = (int)op->ncol;
  s.er.assign(nev, 0);
  s.ei.assign(nev, 0);
  s.rs.assign(nev, 0);
  s.vecs.assign((size_t)2 * n * nev, 0);
  s.vl.assign(n, 0);
  std::memset(&s.out, 0, sizeof s.out);
  s.out.evals_re = s.er.data();
  s.out.evals_im = s.ei.data();
  s.out.resid = s.rs.data();
  s.out.vecs = o.want_vectors ? s.vecs.data() : nullptr;
  s.out.vecl = o.want_left ? s.vl.data() : nullptr;
  s.st = tm_solve(op, &o, &s.out);
  if (s.st != TM_OK) s.err = tm_last_error();
  return s;
}

struct Fail {
  std::string why;
};
void expect(bool c, const std::string &why)
{
  if (!c) throw Fail{why};
}
std::string fmt(const char *f, double a, double b = 0)
{
  char buf[256];
  std::snprintf(buf, sizeof buf, f, a, b);
  return buf;
}

void expect_evals(const Solve &s, const std::vector<double> &ref, int nev, double tol,
                  const char *what)
{
  expect(s.st == TM_OK, std::string(what) + ": status " + std::to_string((int)s.st) + " " + s.err);
  for (int i = 0; i < nev; ++i) {
    const double rel = std::fabs(std::fabs(s.er[i]) / std::fabs(ref[i]) - 1);
    expect(rel <= tol, std::string(what) + ": eigenvalue " + std::to_string(i) +
                           fmt(" is %.17g, expected %.17g", s.er[i], ref[i]));
  }
}

// ---- scenarios ------------------------------------------------------------

// Repeated solves of one operator: graph-replayed restart cycles cached on
// the operator, a larger basis (every cached buffer reallocated), mixed
// precision, then FP64 again; each against the dense spectrum.
void sc_graph_reuse()
{
  const int n = 240;
  std::vector<double> w;
  auto A = random_sym(n, 11, &w, 0.93);
  tm_op *op = dense_op(A, n, true);
  const long g0 = fake_graph_launches();
  Solve a = run(op, opts(3));
  expect_evals(a, w, 3, 1e-10, "first solve");
  expect(a.out.restarts > 0, "the solve needs restarts for this check");
  expect(std::isfinite(a.out.resid_check) && a.out.resid_check <= 1e-6, "explicit residual");
  Solve b = run(op, opts(3));
  expect_evals(b, w, 3, 1e-10, "second solve (cached graphs)");
  for (int i = 0; i < 3; ++i) expect(a.er[i] == b.er[i], "replayed solve differs from the first");
  expect(fake_graph_launches() > g0, "no graph was replayed");
  Solve c = run(op, opts(3, 15));
  expect_evals(c, w, 3, 1e-10, "larger basis");
  Solve d = run(op, opts(3, 0, TM_PREC_MIXED));
  expect_evals(d, w, 3, 1e-10, "mixed");
  Solve e = run(op, opts(3));
  expect_evals(e, w, 3, 1e-10, "FP64 after mixed");
  for (int i = 0; i < 3; ++i) expect(a.er[i] == e.er[i], "FP64 solve changed after other solves");
  tm_op_destroy(op);
}

// Non-symmetric operator with the left vector.
void sc_nonsym_left()
{
  const int n = 300;
  auto A = random_pos(n, 5);
  const double l0 = lead_gen_ref(A, n);
  tm_op *op = dense_op(A, n, false);
  tm_solve_opts o = opts(2);
  o.want_left = 1;
  o.want_vectors = 1;
  Solve s = run(op, o);
  expect(s.st == TM_OK, "status " + s.err);
  expect(std::fabs(s.er[0] / l0 - 1) < 1e-10, fmt("lambda0 %.17g vs %.17g", s.er[0], l0));
  expect(std::fabs(s.out.eval_left / l0 - 1) < 1e-9, "left eigenvalue");
  tm_op_destroy(op);
}

// tm_op_apply_dev on a caller's stream, then an apply on the operator's own
// stream before the caller waits: both use the operator's workspace, so the
// second must be ordered after the first.
void sc_apply_dev_streams()
{
  const int n = 200;
  auto A = random_sym(n, 3, nullptr);
  tm_op *op = dense_op(A, n, true);
  hipStream_t cs;
  TM_HIP(hipStreamCreate(&cs));
  std::vector<double> x1(n), x2(n), y1(n), y2(n);
  for (int i = 0; i < n; ++i) {
    x1[i] = 1.0 + i;
    x2[i] = -3.0 + 0.5 * i;
  }
  ltm::DevBuf dx(sizeof(double) * n), dy(sizeof(double) * n);
  ltm::copy_sync(dx.get(), x1.data(), sizeof(double) * n, hipMemcpyHostToDevice, cs);
  expect(tm_op_apply_dev(op, dx.get(), dy.get(), n, 1, TM_F64, TM_F64, 0, cs) == TM_OK,
         "apply_dev");
  expect(tm_op_apply(op, x2.data(), y2.data(), 0) == TM_OK, "apply");
  ltm::copy_sync(y1.data(), dy.get(), sizeof(double) * n, hipMemcpyDeviceToHost, cs);
  double e1 = 0, e2 = 0;
  for (int i = 0; i < n; ++i) {
    double r1 = 0, r2 = 0;
    for (int j = 0; j < n; ++j) {
      r1 += A[i + (size_t)j * n] * x1[j];
      r2 += A[i + (size_t)j * n] * x2[j];
    }
    e1 = std::max(e1, std::fabs(y1[i] - r1) / (std::fabs(r1) + 1));
    e2 = std::max(e2, std::fabs(y2[i] - r2) / (std::fabs(r2) + 1));
  }
  expect(e1 < 1e-12, fmt("apply_dev result wrong (%.3e)", e1));
  expect(e2 < 1e-12, fmt("apply result wrong (%.3e)", e2));
  TM_HIP(hipStreamDestroy(cs));
  tm_op_destroy(op);
}

// A buffer cannot change inside a graph capture: the allocation fails and
// leaves the buffer as it was.
void sc_alloc_in_capture()
{
  hipStream_t s;
  TM_HIP(hipStreamCreate(&s));
  ltm::DevBuf b(64);
  void *p0 = b.get();
  bool threw = false;
  {
    ltm::CaptureGuard g;
    expect(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess, "capture");
    try {
      b.ensure(128);
    } catch (const ltm::Error &) {
      threw = true;
    }
    hipGraph_t gr = nullptr;
    (void)hipStreamEndCapture(s, &gr);
    if (gr) (void)hipGraphDestroy(gr);
  }
  expect(threw, "allocation inside a capture did not fail");
  expect(b.get() == p0 && b.size() == 64, "the buffer changed inside a capture");
  b.ensure(128);  // outside the capture it grows
  expect(b.size() == 128, "growth after the capture");
  TM_HIP(hipStreamDestroy(s));
}

// Tile choices of different kernel families never share a cache entry.
void sc_tune_families()
{
  hipStream_t s;
  TM_HIP(hipStreamCreate(&s));
  // the fake clock counts executed operations: a candidate's "time" is the
  // number of operations it queues
  auto runner = [&](int fast) {
    return [=](int t) {
      for (int k = 0; k < (t == fast ? 1 : 9); ++k) fake_enqueue(s, [] {});
    };
  };
  const int a = ltm::autotune_tile(ltm::TUNE_KRON, 16, 16, 8, {1, 2}, true, s, runner(2));
  const int b = ltm::autotune_tile(ltm::TUNE_SWEEP, 16, 16, 8, {1, 2}, true, s, runner(1));
  expect(a == 2, "kron choice");
  expect(b == 1, "sweep choice taken from another family's cache entry");
  expect(ltm::autotune_cached(ltm::TUNE_KRON, 16, 16, 8) == 2 &&
             ltm::autotune_cached(ltm::TUNE_SWEEP, 16, 16, 8) == 1,
         "cache entries");
  TM_HIP(hipStreamSynchronize(s));
  TM_HIP(hipStreamDestroy(s));
}

// Operators whose apply starts returning a different matrix after a given
// number of matvecs: the explicit residual check of the leading pair must
// reject the result the Krylov recurrence accepted.
extern "C" int g_fake_perturb_after;
void sc_verify_rejects()
{
  const int n = 300;
  auto A = random_sym(n, 17, nullptr);
  tm_op *op = dense_op(A, n, true);
  Solve a = run(op, opts(1));
  expect(a.st == TM_OK, "clean solve");
  const int64_t mv = a.out.matvecs_f64 - 1;  // the Krylov matvecs (the check adds one)
  tm_op_destroy(op);
  tm_op *op2 = dense_op(A, n, true);
  g_fake_perturb_after = (int)mv;
  Solve b = run(op2, opts(1));
  g_fake_perturb_after = -1;
  expect(b.st == TM_EHIP && b.err.find("explicit residual") != std::string::npos,
         "a perturbed operator passed the explicit residual check: " + b.err);
  tm_op_destroy(op2);
}

// A mixed-precision solve whose wanted pairs sit next to a dense cluster:
// the FP64 refinement of the FP32 Ritz block cannot meet its residual or
// error-bound test in its iterations, and the solve must then finish with
// FP64 Krylov-Schur instead of returning the unrefined block.
void sc_refine_unmet()
{
  const int n = 400;
  // one dominant value above a semicircle bulk (the spectrum of a dense
  // random symmetric matrix): quantiles of the semicircle law on [-11.4, 11.4]
  std::vector<double> d(n);
  d[0] = 200.0;
  for (int i = 1; i < n; ++i) {
    const double p = (i - 0.5) / (n - 1);  // solve F(x) = p for the semicircle CDF
    double lo = -1, hi = 1;
    for (int it = 0; it < 60; ++it) {
      const double x = 0.5 * (lo + hi);
      const double F = 0.5 + (x * std::sqrt(1 - x * x) + std::asin(x)) / M_PI;
      (F < p ? lo : hi) = x;
    }
    d[i] = 11.4 * 0.5 * (lo + hi);
  }
  std::vector<double> w;
  auto A = random_sym(n, 23, &w, 0.0, &d);
  tm_op *op = dense_op(A, n, true);
  Solve s = run(op, opts(3, 0, TM_PREC_MIXED));
  expect_evals(s, w, 3, 1e-10, "mixed next to a cluster");
  tm_op_destroy(op);
}

// Block Krylov-Schur (b vectors per step), FP64 and mixed precision, with
// repeated solves, against the known spectrum; and a start block that spans
// an invariant subspace early (breakdown: a random replacement block).
void sc_block()
{
  const int n = 240;
  std::vector<double> w;
  auto A = random_sym(n, 31, &w, 0.9);
  tm_op *op = dense_op(A, n, true);
  for (int b : {2, 4}) {
    tm_solve_opts o = opts(4, 12);
    o.block = b;
    Solve a = run(op, o);
    expect_evals(a, w, 4, 1e-10, ("block FP64 b=" + std::to_string(b)).c_str());
    expect(a.out.resid_check <= 1e-6, "block explicit residual");
    tm_solve_opts om = opts(4, 12, TM_PREC_MIXED);
    om.block = b;
    Solve c = run(op, om);
    expect_evals(c, w, 4, 1e-10, ("block mixed b=" + std::to_string(b)).c_str());
  }
  tm_op_destroy(op);
  // a matrix of rank 6: the block Krylov space is invariant after two steps
  std::vector<double> d(n, 0.0);
  for (int i = 0; i < 6; ++i) d[i] = 10.0 - i;
  std::vector<double> w2;
  auto B = random_sym(n, 37, &w2, 0.0, &d);
  tm_op *op2 = dense_op(B, n, true);
  tm_solve_opts o = opts(2, 10);
  o.block = 4;
  o.verbose = std::getenv("TM_SCENARIO_VERBOSE") ? 1 : 0;
  Solve r = run(op2, o);
  expect_evals(r, w2, 2, 1e-10, "block with breakdown");
  tm_op_destroy(op2);
}

// Concurrent solves of several operators from host threads (the batch
// driver), each against its dense spectrum.
void sc_batch()
{
  const int nop = 8;
  std::vector<std::vector<double>> As, ws(nop);
  std::vector<tm_op *> ops;
  std::vector<tm_solve_opts> os;
  std::vector<tm_solve_out> outs(nop);
  std::vector<std::vector<double>> er(nop, std::vector<double>(2)), ei(nop, std::vector<double>(2)),
      rs(nop, std::vector<double>(2));
  for (int i = 0; i < nop; ++i) {
    const int n = 60 + 20 * i;
    As.push_back(random_sym(n, 100 + i, &ws[i]));
    ops.push_back(dense_op(As.back(), n, true));
    tm_solve_opts o = opts(2, 0, i % 2 ? TM_PREC_MIXED : TM_PREC_F64);
    os.push_back(o);
    std::memset(&outs[i], 0, sizeof outs[i]);
    outs[i].evals_re = er[i].data();
    outs[i].evals_im = ei[i].data();
    outs[i].resid = rs[i].data();
  }
  const tm_status st = tm_solve_batch(ops.data(), nop, os.data(), outs.data(), 4);
  expect(st == TM_OK, std::string("batch status ") + tm_last_error());
  for (int i = 0; i < nop; ++i) {
    const auto &w = ws[i];
    for (int k = 0; k < 2; ++k)
      expect(std::fabs(std::fabs(er[i][k]) / std::fabs(w[k]) - 1) < 1e-10,
             "batch instance " + std::to_string(i));
    tm_op_destroy(ops[i]);
  }
}

struct Scenario {
  const char *name;
  void (*fn)();
};
const Scenario SCENARIOS[] = {
    {"graph_reuse", sc_graph_reuse},     {"nonsym_left", sc_nonsym_left},
    {"apply_dev_streams", sc_apply_dev_streams}, {"alloc_in_capture", sc_alloc_in_capture},
    {"tune_families", sc_tune_families}, {"verify_rejects", sc_verify_rejects},
    {"refine_unmet", sc_refine_unmet}, {"block", sc_block}, {"batch", sc_batch},
};

int scenarios(const char *only)
{
  int bad = 0;
  for (const Scenario &s : SCENARIOS) {
    if (only && std::strcmp(only, s.name) != 0) continue;
    try {
      s.fn();
      std::printf("ok %s\n", s.name);
    } catch (const Fail &f) {
      std::printf("FAIL %s: %s\n", s.name, f.why.c_str());
      ++bad;
    } catch (const std::exception &e) {
      std::printf("FAIL %s: exception %s\n", s.name, e.what());
      ++bad;
    }
    std::fflush(stdout);
  }
  return bad ? 1 : 0;
}

int solve_file(const char *path, bool verbose)
{
  FILE *f = std::fopen(path, "r");
  if (!f) return 2;
  int n, sym, nev, ncv, prec, left;
  int basis32 = 0;
  if (std::fscanf(f, "%d %d %d %d %d %d %d", &n, &sym, &nev, &ncv, &prec, &left, &basis32) != 7)
    return 2;
  std::vector<double> A((size_t)n * n);
  for (auto &v : A)
    if (std::fscanf(f, "%lf", &v) != 1) return 3;
  std::fclose(f);
  tm_op *op = dense_op(A, n, sym != 0);
  tm_solve_opts o = opts(nev, ncv, (tm_precision)prec);
  o.want_left = left;
  o.want_vectors = 1;
  o.verbose = verbose ? 3 : 0;
  o.refine = 0;
  o.basis_f32 = basis32;
  o.refine_extra = 1;
  Solve s = run(op, o);
  if (s.st != TM_OK && s.st != TM_ENOCONV) {
    std::printf("error %d %s\n", (int)s.st, s.err.c_str());
    tm_op_destroy(op);
    return 1;
  }
  std::printf("status %d nconv %d matvecs %lld %lld restarts %lld\n", (int)s.st, s.out.nconv,
              (long long)s.out.matvecs_f64, (long long)s.out.matvecs_f32,
              (long long)s.out.restarts);
  for (int i = 0; i < nev; ++i) std::printf("eval %.17g %.17g %.3e\n", s.er[i], s.ei[i], s.rs[i]);
  std::printf("left %.17g\n", s.out.eval_left);
  std::printf("check %.3e\n", s.out.resid_check);
  tm_op_destroy(op);
  return 0;
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 2) return 2;
  if (std::strcmp(argv[1], "--scenarios") == 0) return scenarios(argc > 2 ? argv[2] : nullptr);
  return solve_file(argv[1], argc > 2);
}
