// dumpconfig_main.cpp -- the `dumpconfig` program: plants equilibrium
// configurations from the eigenvector a tmat run saved, through the C ABI
// (tm_sample_configs), as the reference's DumpConfig/main.cpp:30-144 does.
//
//   dumpconfig [input.dat] [N] [length] [--seed S]
//
// Reads the same input file as tmat and <prefix>_vec.dat, writes N files
// <prefix>_config_<i>.dat of `length` lines each, one layer state per line as
// its bits from site 0 (DumpConfig/dumpconfig.hpp:32-40).  The draws come
// from std::mt19937 and uniform_real_distribution<double>, seeded with
// time(NULL) as DumpConfig/utilityb.cpp:19-22 does, or with S.
// The models dumpconfig dispatches (DumpConfig/main.cpp:63-129): A 1 (XAT),
// A 2 (YAT), A 5 (XAM), A 6 (YAM), B 6 (YBM), B 7 (ZBM), D 5 (XDM); the
// non-symmetric ones also read <prefix>_vecl.dat.
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/latttmat.hpp"
#include "read_input.hpp"

namespace {

// DumpConfig<S>::readfile (DumpConfig/dumpconfig.hpp:56-76): at least n
// doubles; for the non-symmetric models (S complex) a file of at least n
// complex values is read as interleaved pairs, keeping the real parts.  The
// sampler takes the absolute values.
std::vector<double> read_vec(const std::string &fn, int64_t n, bool complex_ok)
{
  struct stat sb;
  if (stat(fn.c_str(), &sb)) throw std::runtime_error("Open \"" + fn + "\" failed");
  if ((uint64_t)sb.st_size < sizeof(double) * (uint64_t)n)
    throw std::runtime_error("File \"" + fn + "\" size incorrect");
  const bool cplx = complex_ok && (uint64_t)sb.st_size >= 2 * sizeof(double) * (uint64_t)n;
  std::ifstream f(fn, std::ios::binary);
  std::vector<double> v(n);
  if (!cplx) {
    f.read(reinterpret_cast<char *>(v.data()), sizeof(double) * n);
  } else {
    std::vector<double> t(2 * n);
    f.read(reinterpret_cast<char *>(t.data()), sizeof(double) * 2 * n);
    for (int64_t i = 0; i < n; ++i) v[i] = t[2 * i];
  }
  if (!f) throw std::runtime_error("Read from \"" + fn + "\" failed");
  return v;
}

}  // namespace

int main(int argc, char **argv)
{
  const char *infile = "input.dat";
  int ndump = 1, length = 100, npos = 0;
  bool seeded = false;
  unsigned seed = 0;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) {
      seeded = true;
      seed = (unsigned)std::strtoul(argv[++i], nullptr, 10);
    } else if (npos == 0) {
      infile = argv[i];
      ++npos;
    } else if (npos == 1) {
      ndump = std::atoi(argv[i]);
      ++npos;
    } else if (npos == 2) {
      length = std::atoi(argv[i]);
      ++npos;
    }
  }
  std::mt19937 rg;
  rg.seed(seeded ? seed : (unsigned)time(NULL));
  std::uniform_real_distribution<> rrand;
  tmcli::ReadInput in;
  const int rerr = in.read(infile);
  if (rerr) return rerr;  // DumpConfig/main.cpp:57-58
  const auto t1 = std::chrono::steady_clock::now();
  // DumpConfig/main.cpp:63-129: the models dumpconfig dispatches
  const int ad = in.axialdir;
  const bool ok = in.dim == 2 && ((in.model == 'A' && (ad == 1 || ad == 2 || ad == 5 || ad == 6)) ||
                                  (in.model == 'B' && (ad == 6 || ad == 7)) ||
                                  (in.model == 'D' && ad == 5));
  if (in.dim != 2) {
    std::cerr << "Error: dimension not supported yet" << std::endl;
    return 0;
  }
  if (!ok) {
    std::cerr << "Error: axial direction not supported yet" << std::endl;
    return 0;
  }
  if (ndump < 1 || length < 1) return 0;
  try {
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
    latttmat::TransferMatrix op(p);
    const int64_t n = op.cols();
    const tm_kind kind = tm_op_kind(op.handle());
    const bool sym = op.is_symmetric();
    const std::vector<double> vec = read_vec(in.outprefix + "_vec.dat", n, !sym);
    std::vector<double> vecl;
    if (!sym) vecl = read_vec(in.outprefix + "_vecl.dat", n, true);
    const bool onestep = kind == TM_KIND_XAT || kind == TM_KIND_YAT;
    const int64_t per = onestep ? length : 2 * (int64_t)length - 1;
    std::vector<double> u((size_t)per * ndump);
    for (double &x : u) x = rrand(rg);
    std::vector<uint64_t> states((size_t)ndump * length);
    latttmat::check(tm_sample_configs(op.handle(), vec.data(), sym ? nullptr : vecl.data(),
                                      u.data(), ndump, length, states.data()));
    const int L = op.getLayerSize();
    const uint64_t lm = (1ull << L) - 1;
    auto line = [&](std::ofstream &f, uint64_t s) {  // DumpConfig_Base::output
      for (int b = 0; b < L; ++b, s >>= 1) f << (s & 1);
      f << std::endl;
    };
    for (int c = 0; c < ndump; ++c) {
      std::ofstream f(in.outprefix + "_config_" + std::to_string(c) + ".dat");
      const uint64_t *st = states.data() + (size_t)c * length;
      if (kind == TM_KIND_YAT) {  // dumpconfig.cpp:187-220: low layer, then the high layers
        line(f, st[0] & lm);
        for (int j = 0; j < length; ++j) line(f, st[j] >> L);
      } else if (kind == TM_KIND_YAM || kind == TM_KIND_YBM) {  // the low layers
        for (int j = 0; j < length; ++j) line(f, st[j] & lm);
      } else {
        for (int j = 0; j < length; ++j) line(f, st[j]);
      }
    }
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  if (dt >= 1.0) std::cout << "Time consumed: " << dt << " s" << std::endl;
  return 0;
}
