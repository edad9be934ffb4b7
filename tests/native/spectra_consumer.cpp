// A Spectra-style consumer of the operator concept (TMat/pfunc.hpp:52-54,
// 84-98): generic eigensolvers templated on an OpType that use only rows(),
// cols(), perform_op(x, y) and get_t_instance(), the way SymEigsSolver /
// GenEigsSolver use the reference's operator classes.  Compiled with g++
// against include/latttmat.hpp and liblatttmat.so; tests/test_gpu_spectra.py
// compares its eigenvalues with tm_solve.  Test infrastructure only.
//
//   spectra_consumer sym L      XAT (A, axialdir 1) width L, betaJ 0.5, kappa 0.6
//   spectra_consumer gen L      ZBT (B, axialdir 3) width L, betaJ 0.5,
//                               kappa 0.5, kappa2 0.3: lambda0 of T and of T^T
// prints "lambda0 <value>" (and "lambda0_T <value>") with 17 digits.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "latttmat.hpp"

namespace {

// Eigenvalues of a small symmetric matrix (row-major n x n) by cyclic Jacobi.
std::vector<double> jacobi_eigenvalues(std::vector<double> A, int n)
{
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        const double t = (th >= 0 ? 1 : -1) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
      }
  }
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
  return w;
}

// Largest-magnitude eigenvalue of a symmetric OpType: Lanczos with full
// reorthogonalisation, m steps from a uniform(0,1) start vector.
template <class OpType>
double sym_leading(const OpType &op, int m)
{
  const int64_t n = op.cols();
  std::vector<std::vector<double>> V;
  std::vector<double> v(n), w(n);
  unsigned long long seed = 1;
  double nn = 0;
  for (auto &x : v) {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    x = ((seed >> 11) + 0.5) / 9007199254740992.0;
    nn += x * x;
  }
  for (auto &x : v) x /= std::sqrt(nn);
  std::vector<double> H((size_t)m * m, 0.0);
  int k = 0;
  for (; k < m; ++k) {
    V.push_back(v);
    op.perform_op(V[k].data(), w.data());
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j <= k; ++j) {
        double d = 0;
        for (int64_t i = 0; i < n; ++i) d += V[j][i] * w[i];
        for (int64_t i = 0; i < n; ++i) w[i] -= d * V[j][i];
        H[j * m + k] += d;
      }
    double b = 0;
    for (double x : w) b += x * x;
    b = std::sqrt(b);
    if (k + 1 == m || b < 1e-12) {
      ++k;
      break;
    }
    H[(k + 1) * m + k] = b;
    for (int64_t i = 0; i < n; ++i) v[i] = w[i] / b;
  }
  std::vector<double> T((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) T[i * k + j] = 0.5 * (H[i * m + j] + H[j * m + i]);
  double best = 0;
  for (double x : jacobi_eigenvalues(T, k))
    if (std::fabs(x) > std::fabs(best)) best = x;
  return best;
}

// Perron root of a positive OpType by the power iteration (the dominant
// eigenvalue of a positive matrix is real and simple).
template <class OpType>
double power_leading(const OpType &op, int maxit)
{
  const int64_t n = op.cols();
  std::vector<double> x(n, 1.0), y(n);
  double lam = 0;
  for (int it = 0; it < maxit; ++it) {
    op.perform_op(x.data(), y.data());
    double num = 0, den = 0, yy = 0;
    for (int64_t i = 0; i < n; ++i) {
      num += x[i] * y[i];
      den += x[i] * x[i];
      yy += y[i] * y[i];
    }
    const double l = num / den;
    const double ny = std::sqrt(yy);
    for (int64_t i = 0; i < n; ++i) x[i] = y[i] / ny;
    if (it > 10 && std::fabs(l - lam) <= 1e-15 * std::fabs(l)) return l;
    lam = l;
  }
  return lam;
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 3) return 2;
  const std::string mode = argv[1];
  const int L = std::atoi(argv[2]);
  tm_params p;
  std::memset(&p, 0, sizeof p);
  p.dim = 2;
  p.neigs = 1;
  p.size1 = L;
  p.betaJ = 0.5;
  try {
    if (mode == "sym") {
      p.model = 'A';
      p.axialdir = 1;
      p.kappa = 0.6;
      latttmat::TransferMatrix op(p);
      if (op.rows() != op.cols() || !op.is_symmetric()) return 3;
      std::printf("ncv %d normfactor %.17g\n", op.suggested_nev(1), op.getnormfactor());
      std::printf("lambda0 %.17g\n", sym_leading(op, 60));
    } else {
      p.model = 'B';
      p.axialdir = 3;
      p.kappa = 0.5;
      p.kappa2 = 0.3;
      latttmat::TransferMatrix op(p);
      auto opt = op.get_t_instance();
      std::printf("lambda0 %.17g\n", power_leading(op, 20000));
      std::printf("lambda0_T %.17g\n", power_leading(opt, 20000));
    }
  } catch (const latttmat::Error &e) {
    std::printf("error %d %s\n", (int)e.status(), e.what());
    return 1;
  }
  return 0;
}
