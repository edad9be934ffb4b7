// Runs the eigen driver (tm_solver.cpp) on dense matrices read from a file
// and prints eigenvalues, for tests/test_driver_cpu.py.
// File format (text): n sym nev ncv prec want_left, then n*n column-major.
#include <cstdio>
#include <vector>

#include "../../latttmat_amd/csrc/tm_internal.hpp"

namespace ltm { tm_status solve_impl(tm_op *op, const tm_solve_opts &o, tm_solve_out *out); }

int main(int argc, char **argv)
{
  FILE *f = std::fopen(argv[1], "r");
  int n, sym, nev, ncv, prec, left;
  int basis32 = 0;
  if (std::fscanf(f, "%d %d %d %d %d %d %d", &n, &sym, &nev, &ncv, &prec, &left, &basis32) != 7)
    return 2;
  tm_op op;
  op.ncol = n;
  op.symmetric = sym;
  op.tab64.alloc(sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i)
    if (std::fscanf(f, "%lf", op.tab64.as<double>() + i) != 1) return 3;
  tm_solve_opts o;
  std::memset(&o, 0, sizeof o);
  o.nev = nev; o.ncv = ncv; o.tol = 1e-10; o.maxit = 1000; o.maxtrial = 3;
  o.precision = (tm_precision)prec; o.seed = 1; o.want_left = left; o.want_vectors = 1;
  o.mixed_tol = 1e-6; o.verbose = argc > 2 ? 3 : 0; o.refine = 0; o.basis_f32 = basis32; o.refine_extra = 1;
  std::vector<double> er(nev), ei(nev), rs(nev), vecs((size_t)2 * n * nev), vl(n);
  tm_solve_out out;
  std::memset(&out, 0, sizeof out);
  out.evals_re = er.data(); out.evals_im = ei.data(); out.resid = rs.data();
  out.vecs = vecs.data(); out.vecl = vl.data();
  try {
    tm_status st = ltm::solve_impl(&op, o, &out);
    std::printf("status %d nconv %d matvecs %lld %lld restarts %lld\n", (int)st, out.nconv,
                (long long)out.matvecs_f64, (long long)out.matvecs_f32, (long long)out.restarts);
  } catch (const ltm::Error &e) {
    std::printf("error %d %s\n", (int)e.status, e.what());
    return 1;
  }
  for (int i = 0; i < nev; ++i) std::printf("eval %.17g %.17g %.3e\n", er[i], ei[i], rs[i]);
  std::printf("left %.17g\n", out.eval_left);
  return 0;
}
