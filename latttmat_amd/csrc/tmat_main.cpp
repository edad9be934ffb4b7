// tmat_main.cpp -- the `tmat` program: reads a tmat input file, builds the
// operator through the C ABI, solves on the GPU and writes the outputs of the
// reference program.
//
//   tmat [input.dat] [--mixed] [--verbose] [--check-input]
//
// --mixed solves with FP32 T.x and FP64 refinement; --check-input only
// parses the input and exits with the parser's code.
//
// Follows TMat/main.cpp:31-58 (driver, timer, iteration count),
// TMat/read_input.cpp:14-92 (positional "key = value" parser, error codes
// 2/3/4), TMat/pfunc.cpp:16-106 (dispatch, through tm_op_create) and
// TMat/pfunc.hpp:79-202 (solve with retries, writers X_val.dat / X_vec.dat /
// X_vecl.dat / X_phys.dat).
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/latttmat.hpp"
#include "read_input.hpp"

namespace {

// Right-aligned column of values at a common width, no trailing newline:
// the layout of streaming an Eigen vector at setprecision(16)
// (TMat/pfunc.hpp:117-118).
template <typename T>
void write_column(std::ostream &os, const std::vector<T> &v)
{
  std::vector<std::string> s(v.size());
  size_t w = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    std::ostringstream o;
    o << std::setprecision(16) << v[i];
    s[i] = o.str();
    w = std::max(w, s[i].size());
  }
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) os << '\n';
    os << std::setw((int)w) << s[i];
  }
}

// Phys output X_phys.dat (TMat/pfunc.hpp:167-200): first line the
// TMAT_NMOMENTS=4 magnetisation moments, second line the operator's extra
// observables, both as tm_op_moments reduces them on the GPU (which
// observables each operator kind reports is documented there).
void write_phys(const std::string &fn, const tm_op *op, const double *ql, const double *qr)
{
  double v[8];
  int nv = 0;
  latttmat::check(tm_op_moments(op, ql, qr, v, &nv));
  std::ofstream f(fn);
  f << std::setprecision(10);
  for (int k = 0; k < 4; ++k) f << (k ? "\t" : "") << v[k];
  for (int k = 4; k < nv; ++k) f << (k > 4 ? '\t' : '\n') << v[k];
}


}  // namespace

int main(int argc, char **argv)
{
  const char *infile = "input.dat";
  bool mixed = false, verbose = false, check_input = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--mixed")) mixed = true;
    else if (!std::strcmp(argv[i], "--verbose")) verbose = true;
    else if (!std::strcmp(argv[i], "--check-input")) check_input = true;
    else infile = argv[i];
  }
  tmcli::ReadInput in;
  const int err = in.read(infile);
  // The reference program prints the reason and still exits with status 0
  // (TMat/main.cpp:42-43, 63-64); --check-input returns read()'s code
  // (2 cannot open, 3 read failure, 4 invalid dim/model/size) instead.
  if (check_input) return err;
  if (err) return 0;

  tm_params p;
  std::memset(&p, 0, sizeof p);
  p.model = in.model;
  p.dim = in.dim;
  p.axialdir = in.axialdir;
  p.neigs = in.neigs;
  p.size1 = in.size1;
  p.size2 = in.size2;
  p.betaJ = in.betaJ;
  p.kappa = in.kappa;
  p.kappa2 = in.kappa2;
  p.h = in.h;

  const auto t1 = std::chrono::steady_clock::now();
  int64_t matvecs = 0;
  try {
    latttmat::TransferMatrix op(p);
    const int64_t n = op.cols();
    const bool sym = op.is_symmetric();
    const int nev = in.neigs;
    tm_solve_opts o;
    tm_solve_opts_default(&o);
    o.nev = nev;
    o.precision = mixed ? TM_PREC_MIXED : TM_PREC_F64;
    o.want_vectors = 1;
    o.want_left = (in.savevec[1] && !sym) ? 1 : 0;
    o.verbose = verbose;
    std::vector<double> er(nev), ei(nev), rs(nev), vecs(sym ? (size_t)n * nev : (size_t)2 * n * nev),
        vl(n);
    tm_solve_out out;
    std::memset(&out, 0, sizeof out);
    out.evals_re = er.data();
    out.evals_im = ei.data();
    out.resid = rs.data();
    out.vecs = vecs.data();
    out.vecl = o.want_left ? vl.data() : nullptr;
    const tm_status st = tm_solve(op.handle(), &o, &out);
    matvecs = out.matvecs_f64 + out.matvecs_f32;
    if (st != TM_OK && st != TM_ENOCONV) latttmat::check(st);
    for (int t = 1; t < out.trials; ++t)
      std::cout << "eigensolver not converged (" << t << "/3)" << std::endl;
    const int ncv = out.nconv;
    if (ncv == 0) {
      std::cerr << "Number of converged eigenvalue is zero" << std::endl;
      return -1;
    }
    const double nf = op.getnormfactor();
    {  // X_val.dat (pfunc.hpp:116-119)
      std::ofstream f(in.outprefix + "_val.dat");
      if (sym) {
        std::vector<double> v(ncv);
        for (int i = 0; i < ncv; ++i) v[i] = er[i] / nf;
        write_column(f, v);
      } else {
        std::vector<std::complex<double>> v(ncv);
        for (int i = 0; i < ncv; ++i) v[i] = std::complex<double>(er[i], ei[i]) / nf;
        write_column(f, v);
      }
    }
    if (in.savevec[0]) {  // X_vec.dat (pfunc.hpp:124-142)
      std::ofstream f(in.outprefix + "_vec.dat", std::ios::binary);
      if (sym) {
        f.write(reinterpret_cast<const char *>(vecs.data()), sizeof(double) * (size_t)ncv * n);
      } else if (ncv == 1) {
        for (int64_t i = 0; i < n; ++i)
          f.write(reinterpret_cast<const char *>(&vecs[2 * i]), sizeof(double));
      } else {
        f.write(reinterpret_cast<const char *>(vecs.data()), sizeof(double) * 2 * (size_t)ncv * n);
      }
    }
    if (in.savevec[1]) {
      std::vector<double> qr(n), ql;
      for (int64_t i = 0; i < n; ++i) qr[i] = sym ? vecs[i] : vecs[2 * i];
      const bool have_left = !sym && out.eval_left == out.eval_left;
      if (have_left) {
        if (in.savevec[0]) {  // X_vecl.dat (pfunc.hpp:155-165)
          std::ofstream f(in.outprefix + "_vecl.dat", std::ios::binary);
          f.write(reinterpret_cast<const char *>(vl.data()), sizeof(double) * n);
        }
        std::cout.precision(10);
        std::cout << "Eigensolver check: " << std::abs(er[0] / out.eval_left) << std::endl;
        std::cout.precision(6);
        ql = vl;
      } else {
        ql = qr;
      }
      // X_phys.dat
      write_phys(in.outprefix + "_phys.dat", op.handle(), ql.data(), qr.data());
    }
    if (verbose)
      std::cerr << "[tmat] matvecs fp64 " << out.matvecs_f64 << " fp32 " << out.matvecs_f32
                << " restarts " << out.restarts << " solve " << out.seconds << " s" << std::endl;
  } catch (const latttmat::Error &e) {
    std::cerr << "Catched latttmat::Error (status " << (int)e.status() << ")\nwhat()  " << e.what()
              << std::endl;
    return 1;
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  if (dt >= 1.0) std::cout << "Time consumed: " << dt << " s" << std::endl;
  if (matvecs > 0) std::cout << "Number of iteration: " << matvecs << std::endl;
  return 0;
}
