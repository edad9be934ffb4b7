// tm_dense.cpp -- small dense eigensolvers for the projected (Ritz) problem.
// The matrices are at most a few dozen rows (the Krylov basis size), so
// plain O(n^3) algorithms are used: Householder tridiagonalisation + implicit
// symmetric QR for the symmetric case,
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
  // Householder reduction to tridiagonal form T = Q^T A Q, then implicit
  // symmetric QR steps with Wilkinson shifts on T (Golub & Van Loan, Matrix
  // Computations, algorithms 8.3.1 and 8.3.2), the rotations accumulated
  // into Z = Q.  The symmetric part of A is used.
  std::vector<double> a(A, A + (size_t)n * n), v(n), p(n);
  auto at = [&](int i, int j) -> double & { return a[i + (size_t)j * n]; };
  auto zt = [&](int i, int j) -> double & { return Z[i + (size_t)j * n]; };
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) {
      const double x = 0.5 * (at(i, j) + at(j, i));
      at(i, j) = at(j, i) = x;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) zt(i, j) = (i == j) ? 1.0 : 0.0;
  for (int k = 0; k + 2 < n; ++k) {
    // reflector I - tau v v^T mapping A[k+1.., k] onto alpha e_1
    const int r0 = k + 1, len = n - r0;
    double xn = 0;
    for (int i = 0; i < len; ++i) xn += at(r0 + i, k) * at(r0 + i, k);
    xn = std::sqrt(xn);
    if (xn == 0) continue;
    const double alpha = at(r0, k) > 0 ? -xn : xn;
    double vn = 0;
    for (int i = 0; i < len; ++i) {
      v[i] = at(r0 + i, k) - (i == 0 ? alpha : 0.0);
      vn += v[i] * v[i];
    }
    if (vn == 0) continue;
    const double tau = 2.0 / vn;
    // trailing block: A22 <- A22 - v q^T - q v^T, p = tau A22 v,
    // q = p - (tau / 2) (v^T p) v
    double vp = 0;
    for (int i = 0; i < len; ++i) {
      double t = 0;
      for (int j = 0; j < len; ++j) t += at(r0 + i, r0 + j) * v[j];
      p[i] = tau * t;
      vp += v[i] * p[i];
    }
    const double kk = 0.5 * tau * vp;
    for (int i = 0; i < len; ++i) p[i] -= kk * v[i];
    for (int j = 0; j < len; ++j)
      for (int i = 0; i < len; ++i) at(r0 + i, r0 + j) -= v[i] * p[j] + p[i] * v[j];
    at(r0, k) = at(k, r0) = alpha;
    for (int i = 1; i < len; ++i) at(r0 + i, k) = at(k, r0 + i) = 0.0;
    // Z[:, r0..] <- Z[:, r0..] (I - tau v v^T)
    for (int i = 0; i < n; ++i) {
      double t = 0;
      for (int j = 0; j < len; ++j) t += zt(i, r0 + j) * v[j];
      t *= tau;
      for (int j = 0; j < len; ++j) zt(i, r0 + j) -= t * v[j];
    }
  }
  // T: diagonal d = w, off-diagonal e[i] = T(i, i+1)
  std::vector<double> e(n > 1 ? n - 1 : 1, 0.0);
  for (int i = 0; i < n; ++i) w[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);
  const double eps = 2.220446049250313e-16;
  for (int iter = 0; iter < 60 * n; ++iter) {
    // deflate: the bottom hi of the last unreduced block [lo, hi]
    int hi = n - 1;
    while (hi > 0 && std::fabs(e[hi - 1]) <= eps * (std::fabs(w[hi - 1]) + std::fabs(w[hi]))) {
      e[hi - 1] = 0;
      --hi;
    }
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0) {
      if (std::fabs(e[lo - 1]) <= eps * (std::fabs(w[lo - 1]) + std::fabs(w[lo]))) {
        e[lo - 1] = 0;
        break;
      }
      --lo;
    }
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 closer to w[hi]
    const double dl = 0.5 * (w[hi - 1] - w[hi]), bb = e[hi - 1];
    const double mu = w[hi] - bb * bb / (dl + std::copysign(std::sqrt(dl * dl + bb * bb), dl));
    // one implicit QR step on [lo, hi]: rotation k on (k, k+1) is chosen
    // from the shifted first column (k = lo) or to remove the bulge at
    // (k-1, k+1), and leaves a bulge at (k, k+2)
    double x = w[lo] - mu, z = e[lo];
    for (int k = lo; k < hi; ++k) {
      const double r = std::sqrt(x * x + z * z);
      const double c = r > 0 ? x / r : 1.0, sn = r > 0 ? z / r : 0.0;
      if (k > lo) e[k - 1] = r;
      const double dk = w[k], ek = e[k], dk1 = w[k + 1];
      w[k] = c * c * dk + 2 * c * sn * ek + sn * sn * dk1;
      w[k + 1] = sn * sn * dk - 2 * c * sn * ek + c * c * dk1;
      e[k] = c * sn * (dk1 - dk) + (c * c - sn * sn) * ek;
      if (k + 1 < hi) {
        z = sn * e[k + 1];
        e[k + 1] *= c;
        x = e[k];
      }
      for (int i = 0; i < n; ++i) {
        const double zk = zt(i, k), zk1 = zt(i, k + 1);
        zt(i, k) = c * zk + sn * zk1;
        zt(i, k + 1) = c * zk1 - sn * zk;
      }
    }
  }
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
