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

namespace {

struct ReadInput {
  char model = 0;
  int dim = 0, axialdir = 0, neigs = 0;
  int savevec[2] = {0, 0};
  std::string outprefix;
  int size = 0, size1 = 0, size2 = 0;
  double betaJ = 0, kappa = 0, kappa2 = 0, h = 0;

  // TMat/read_input.cpp:14-92: every value is read after the next '=';
  // key names are not checked.
  int read(const char *inputf)
  {
    std::ifstream in(inputf);
    if (!in) {
      std::cout << "Can't open " << inputf << " for input." << std::endl;
      return 2;
    }
    std::cout << "Reading input from file " << inputf << std::endl;
    auto next = [&]() {
      in.ignore(std::numeric_limits<std::streamsize>::max(), '=');
    };
    next(); in >> model;
    next(); in >> dim;
    next(); in >> axialdir;
    next(); in >> neigs;
    next(); in >> savevec[0] >> savevec[1];
    next(); in >> outprefix;
    next();
    if (dim == 2) {
      in >> size;
      size1 = size;
    } else if (dim == 3) {
      in >> size1 >> size2;
      size = size1 << 16 | size2;
    } else {
      std::cout << "Invalid dim " << dim << std::endl;
      return 4;
    }
    next(); in >> betaJ;
    next();
    std::string kline;
    std::getline(in, kline);
    std::istringstream ks(kline);
    ks >> kappa;
    if (!(ks >> kappa2)) kappa2 = 0.0;
    next(); in >> h;
    int error = 0;
    if (in.fail()) {
      std::cout << "Error reading input file " << inputf << std::endl;
      error = 3;
    }
    std::cout << "     MODEL : " << dim << "D_" << model << axialdir << "\n";
    std::cout << "Save & flg : " << outprefix << " " << savevec[0] << " " << savevec[1] << "\n";
    if (dim == 2) std::cout << "      size : " << size << "\n";
    else std::cout << "      size : " << size1 << " " << size2 << "\n";
    std::cout << "     betaJ : " << betaJ << "\n";
    std::cout << "     kappa : " << kappa << std::endl;
    if (model == 'B') std::cout << "    kappa2 : " << kappa2 << std::endl;
    std::cout << "         h : " << h << std::endl;
    if (model != 'A' && model != 'B' && model != 'D') {
      std::cout << "Invalid model " << model << std::endl;
      error = 4;
    } else if (size < 4) {
      std::cout << "Invalid size (<4) " << size << std::endl;
      error = 4;
    }
    return error;
  }
};

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

// Phys output X_phys.dat: first line the TMAT_NMOMENTS=4 magnetisation
// moments <m^k>/L (m = 2 popcount(s) - L), second line the extra
// observables.  Symmetric operators weight state s by q_s^2 and add
// <node>, <node^2> (TMat/tmatphys.cpp:13-57; reduced operators sum over
// orbit representatives, TMat/tmatreduce.cpp:565-613; the 3D reduced
// operator reports NaN for the node counts, TMat/tmat3d.cpp:165-177).
// Non-symmetric operators weight by ql_s qr_s / sum(ql qr) and have no extra
// observables (TMat/tmatphys.cpp:106-132, 98-104; the reduced ZBM likewise
// over its representatives, TMat/tmatreduce.cpp:1748-1781).  The reduced DNNI
// operator reports <|m_st|>/L, <|m|>/L, <m_st^2>/L^2, <m^2>/L^2 instead, m_st
// the magnetisation with every odd site flipped (TMat/tmatreduce.cpp:1289-1317).
// The axial two-layer operators use the generic moments over all 4^L states
// with the low layer's magnetisation (TMat/tmatphys.cpp:71-98); Y2AT writes
// zeros (its moments are a stub, TMat/tmat.cpp:304-309).
// The reduced axial operators weight by the representative's current (high)
// layer (TMat/tmatreduce.cpp:882-909).
enum class PhysKind { Sym2D, Sym3D, Gen, Dnni, GenLayers, GenHigh, Zero };

int nnode(uint64_t s, int L, int bc)  // TMat/utility.cpp:33-47
{
  uint64_t sh = s >> 1;
  if (!bc) sh |= (s & 1) << (L - 1);
  else sh |= (s & (uint64_t)(L - 1)) << 1;  // mask as written there
  return __builtin_popcountll(s ^ sh);
}

void write_phys(const std::string &fn, PhysKind pk, int L, int bc,
                const std::vector<uint64_t> &states, const std::vector<double> &ql,
                const std::vector<double> &qr)
{
  double mom[4] = {0, 0, 0, 0}, dnorm = 0;
  std::vector<double> ext;
  if (pk == PhysKind::Sym2D) ext.assign(2, 0.0);
  if (pk == PhysKind::Sym3D) ext.assign(2, NAN);
  if (pk == PhysKind::Dnni) ext.assign(4, 0.0);
  const uint64_t mask = L >= 64 ? ~0ull : (1ull << L) - 1;
  const size_t n = qr.size();
  for (size_t i = 0; i < n; ++i) {
    const uint64_t s = states.empty() ? (uint64_t)i : states[i];
    const uint64_t sl = pk == PhysKind::GenLayers ? (s & mask) : pk == PhysKind::GenHigh ? (s >> L) : s;
    const double m = 2.0 * __builtin_popcountll(sl) - L;
    const bool gen = pk == PhysKind::Gen || pk == PhysKind::GenLayers ||
                     pk == PhysKind::GenHigh || pk == PhysKind::Zero;
    const double w = gen ? ql[i] * qr[i] : ql[i] * ql[i];
    dnorm += w;
    double mm = m;
    for (int k = 0; k < 4; ++k) {
      mom[k] += mm * w;
      mm *= m;
    }
    if (pk == PhysKind::Sym2D) {
      const double nd = nnode(s, L, bc);
      ext[0] += nd * w;
      ext[1] += nd * nd * w;
    }
    if (pk == PhysKind::Dnni) {
      const double ms = 2.0 * __builtin_popcountll((s ^ 0xaaaaaaaaaaaaaaaaull) & mask) - L;
      ext[0] += std::fabs(ms) * w;
      ext[1] += std::fabs(m) * w;
      ext[2] += ms * ms * w;
      ext[3] += m * m * w;
    }
  }
  if (pk == PhysKind::Dnni) {
    ext[0] /= L;
    ext[1] /= L;
    ext[2] /= (double)L * L;
    ext[3] /= (double)L * L;
  }
  const double div = (pk == PhysKind::Gen || pk == PhysKind::GenLayers || pk == PhysKind::GenHigh)
                         ? L * dnorm
                         : (double)L;
  if (pk == PhysKind::Zero) mom[0] = mom[1] = mom[2] = mom[3] = 0.0;
  std::ofstream f(fn);
  f << std::setprecision(10);
  for (int k = 0; k < 4; ++k) f << (k ? "\t" : "") << mom[k] / div;
  for (size_t k = 0; k < ext.size(); ++k) f << (k ? '\t' : '\n') << ext[k];
}

// Boundary flag the operator's physics uses for node counting: the XAT
// boundary code (TMat/tmat.cpp:39), ext & 1 for the reduced operator
// (TMat/tmatreduce.cpp:338); XDT keeps the periodic default
// (TMat/tmat.cpp:322-323).
int phys_bc(const ReadInput &in)
{
  if (in.model != 'A' || in.dim != 2) return 0;
  switch (in.axialdir) {
    case 9: case 13: return 1;
    case 17: return 5;
    case 18: return 9;
    case 19: return 13;
    case 20: return 17;
    default: return 0;
  }
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
  ReadInput in;
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
      const tm_kind kind = tm_op_kind(op.handle());
      std::vector<uint64_t> states;
      if (kind == TM_KIND_XAM2D || kind == TM_KIND_XAM3D || kind == TM_KIND_XDM ||
          kind == TM_KIND_ZBM || kind == TM_KIND_YAM || kind == TM_KIND_YBM) {
        std::vector<uint16_t> dups(n);
        states.resize(n);
        latttmat::check(tm_op_orbits(op.handle(), states.data(), dups.data()));
      }
      const PhysKind pk = kind == TM_KIND_YAT    ? PhysKind::GenLayers
                          : (kind == TM_KIND_YAM || kind == TM_KIND_YBM) ? PhysKind::GenHigh
                          : kind == TM_KIND_Y2AT ? PhysKind::Zero
                          : !sym                 ? PhysKind::Gen
                          : kind == TM_KIND_XDM ? PhysKind::Dnni
                          : (kind == TM_KIND_XAM3D || kind == TM_KIND_XAT3D) ? PhysKind::Sym3D
                                                                             : PhysKind::Sym2D;
      write_phys(in.outprefix + "_phys.dat", pk, op.getLayerSize(), phys_bc(in), states, ql, qr);
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
