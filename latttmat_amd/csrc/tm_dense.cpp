// tm_dense.cpp -- small dense eigensolvers for the projected (Ritz) problem.
// The matrices are at most a few dozen rows (the Krylov basis size), so
// plain O(n^3) algorithms are used: cyclic Jacobi for the symmetric case,
// Householder reduction + shifted complex QR (Schur form) + triangular
// back-substitution for the general case.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "tm_internal.hpp"

namespace ltm {

void sym_eig(int n, const double *A, double *w, double *Z)
{
  std::vector<double> a(A, A + (size_t)n * n);
  auto at = [&](int i, int j) -> double & { return a[i + (size_t)j * n]; };
  auto zt = [&](int i, int j) -> double & { return Z[i + (size_t)j * n]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) zt(i, j) = (i == j) ? 1.0 : 0.0;
  // symmetrise
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) {
      const double v = 0.5 * (at(i, j) + at(j, i));
      at(i, j) = at(j, i) = v;
    }
  double fro = 0;
  for (double v : a) fro += v * v;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        if (i != j) off += at(i, j) * at(i, j);
    if (off <= 1e-32 * fro || off == 0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = at(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (at(q, q) - at(p, p)) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = at(k, p), akq = at(k, q);
          at(k, p) = c * akp - s * akq;
          at(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = at(p, k), aqk = at(q, k);
          at(p, k) = c * apk - s * aqk;
          at(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double zkp = zt(k, p), zkq = zt(k, q);
          zt(k, p) = c * zkp - s * zkq;
          zt(k, q) = s * zkp + c * zkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = at(i, i);
}

void gen_eig(int n, const double *A, double *wr, double *wi, double *Zr, double *Zi)
{
  using cd = std::complex<double>;
  std::vector<cd> H((size_t)n * n), Q((size_t)n * n);
  auto h = [&](int i, int j) -> cd & { return H[i + (size_t)j * n]; };
  auto q = [&](int i, int j) -> cd & { return Q[i + (size_t)j * n]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      h(i, j) = A[i + (size_t)j * n];
      q(i, j) = (i == j) ? 1.0 : 0.0;
    }
  // Householder reduction to upper Hessenberg form: H = Q^H A Q
  std::vector<cd> v(n);
  for (int k = 0; k < n - 2; ++k) {
    double alpha2 = 0;
    for (int i = k + 1; i < n; ++i) alpha2 += std::norm(h(i, k));
    const double alpha = std::sqrt(alpha2);
    if (alpha < 1e-300) continue;
    const cd x0 = h(k + 1, k);
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0);
    for (int i = 0; i < n; ++i) v[i] = 0;
    v[k + 1] = x0 + ph * alpha;
    for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
    double vn2 = 0;
    for (int i = k + 1; i < n; ++i) vn2 += std::norm(v[i]);
    if (vn2 < 1e-300) continue;
    // H = (I - 2vv^H/vn2) H (I - 2vv^H/vn2)
    for (int j = 0; j < n; ++j) {
      cd s = 0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * h(i, j);
      s *= 2.0 / vn2;
      for (int i = k + 1; i < n; ++i) h(i, j) -= v[i] * s;
    }
    for (int i = 0; i < n; ++i) {
      cd s = 0;
      for (int j = k + 1; j < n; ++j) s += h(i, j) * v[j];
      s *= 2.0 / vn2;
      for (int j = k + 1; j < n; ++j) h(i, j) -= s * std::conj(v[j]);
    }
    for (int i = 0; i < n; ++i) {
      cd s = 0;
      for (int j = k + 1; j < n; ++j) s += q(i, j) * v[j];
      s *= 2.0 / vn2;
      for (int j = k + 1; j < n; ++j) q(i, j) -= s * std::conj(v[j]);
    }
    for (int i = k + 2; i < n; ++i) h(i, k) = 0;
  }
  // shifted QR to Schur form: H = Q T Q^H
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int iter = 0;
  std::vector<cd> gc(n), gs(n);
  while (hi > 0 && iter < 100 * n) {
    // deflation
    int lo = hi;
    while (lo > 0) {
      const double sc = std::abs(h(lo, lo)) + std::abs(h(lo - 1, lo - 1));
      if (std::abs(h(lo, lo - 1)) <= eps * (sc > 0 ? sc : 1.0)) {
        h(lo, lo - 1) = 0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      --hi;
      iter = 0;
      continue;
    }
    ++iter;
    // Wilkinson shift from the trailing 2x2 of the active window
    const cd a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
    const cd tr = a + d, det = a * d - b * c;
    const cd disc = std::sqrt(tr * tr / 4.0 - det);
    cd mu1 = tr / 2.0 + disc, mu2 = tr / 2.0 - disc;
    cd mu = std::abs(mu1 - d) < std::abs(mu2 - d) ? mu1 : mu2;
    if (iter % 11 == 10) mu = d + std::abs(h(hi, hi - 1));  // exceptional shift
    for (int i = lo; i <= hi; ++i) h(i, i) -= mu;
    for (int k = lo; k < hi; ++k) {
      const cd x = h(k, k), y = h(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      cd cc, ss;
      if (r < 1e-300) { cc = 1; ss = 0; }
      else { cc = x / r; ss = y / r; }
      gc[k] = cc;
      gs[k] = ss;
      // rows k, k+1: [conj(c) conj(s); -s c]
      for (int j = k; j < n; ++j) {
        const cd t1 = h(k, j), t2 = h(k + 1, j);
        h(k, j) = std::conj(cc) * t1 + std::conj(ss) * t2;
        h(k + 1, j) = -ss * t1 + cc * t2;
      }
    }
    for (int k = lo; k < hi; ++k) {
      const cd cc = gc[k], ss = gs[k];
      const int imax = std::min(k + 2, hi);
      for (int i = 0; i <= imax; ++i) {
        const cd t1 = h(i, k), t2 = h(i, k + 1);
        h(i, k) = t1 * cc + t2 * ss;
        h(i, k + 1) = -t1 * std::conj(ss) + t2 * std::conj(cc);
      }
      for (int i = 0; i < n; ++i) {
        const cd t1 = q(i, k), t2 = q(i, k + 1);
        q(i, k) = t1 * cc + t2 * ss;
        q(i, k + 1) = -t1 * std::conj(ss) + t2 * std::conj(cc);
      }
    }
    for (int i = lo; i <= hi; ++i) h(i, i) += mu;
  }
  // eigenvectors of the triangular T, then x = Q y
  std::vector<cd> y(n);
  double tnorm = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, std::abs(h(i, j)));
  const double small = eps * (tnorm > 0 ? tnorm : 1.0);
  for (int k = 0; k < n; ++k) {
    const cd lam = h(k, k);
    for (int i = 0; i < n; ++i) y[i] = 0;
    y[k] = 1;
    for (int i = k - 1; i >= 0; --i) {
      cd s = 0;
      for (int l = i + 1; l <= k; ++l) s += h(i, l) * y[l];
      cd den = h(i, i) - lam;
      if (std::abs(den) < small) den = small;
      y[i] = -s / den;
    }
    double nrm = 0;
    std::vector<cd> x(n);
    for (int i = 0; i < n; ++i) {
      cd s = 0;
      for (int l = 0; l <= k; ++l) s += q(i, l) * y[l];
      x[i] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    // fix the phase: largest component real positive
    int imax = 0;
    for (int i = 1; i < n; ++i)
      if (std::abs(x[i]) > std::abs(x[imax])) imax = i;
    const cd ph = std::abs(x[imax]) > 0 ? std::conj(x[imax]) / std::abs(x[imax]) : cd(1.0);
    for (int i = 0; i < n; ++i) {
      const cd z = x[i] * ph / nrm;
      Zr[i + (size_t)k * n] = z.real();
      Zi[i + (size_t)k * n] = z.imag();
    }
    wr[k] = lam.real();
    wi[k] = lam.imag();
  }
  // A real matrix has real eigenvalues or conjugate pairs; eigenvalues whose
  // imaginary part is at rounding level are real: make them exactly real
  // with real (phase-fixed) eigenvectors.
  const double thr = 64 * eps * (tnorm > 0 ? tnorm : 1.0);
  for (int k = 0; k < n; ++k)
    if (std::fabs(wi[k]) <= thr) {
      wi[k] = 0;
      double nr = 0;
      for (int i = 0; i < n; ++i) nr += Zr[i + (size_t)k * n] * Zr[i + (size_t)k * n];
      nr = std::sqrt(nr);
      for (int i = 0; i < n; ++i) {
        Zr[i + (size_t)k * n] /= nr;
        Zi[i + (size_t)k * n] = 0;
      }
    }
}

}  // namespace ltm
