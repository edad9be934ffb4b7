// tm_consts.hpp -- host-side operator constants (normfactor, the three
// popcount-indexed diagonal tables, the Boltzmann factors bzwj1..3), written
// once over a scalar type S: S = double builds the operator, S = HDual (value
// and derivative with respect to one model parameter) builds the
// derivative tables the Hellmann-Feynman kernels need.
#pragma once

#include <cmath>

namespace ltm {

// Forward-mode dual number on the host.
struct HDual {
  double v = 0, d = 0;
  HDual() = default;
  HDual(double a) : v(a), d(0) {}
  HDual(double a, double b) : v(a), d(b) {}
};
inline HDual operator+(HDual a, HDual b) { return {a.v + b.v, a.d + b.d}; }
inline HDual operator-(HDual a, HDual b) { return {a.v - b.v, a.d - b.d}; }
inline HDual operator-(HDual a) { return {-a.v, -a.d}; }
inline HDual operator*(HDual a, HDual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline HDual operator/(HDual a, HDual b)
{
  return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
inline HDual exp(HDual a)
{
  const double e = std::exp(a.v);
  return {e, e * a.d};
}
inline HDual sqrt(HDual a)
{
  const double r = std::sqrt(a.v);
  return {r, 0.5 * a.d / r};
}
inline double val(double a) { return a; }
inline double val(HDual a) { return a.v; }
inline double der(double) { return 0.0; }
inline double der(HDual a) { return a.d; }

template <class S>
struct OpConsts {
  S tab[192];          // diagonal tables: [0,64) i1, [64,128) i2, [128,192) i0
  S nf;                // normfactor
  S b1[2], b2[2], b3[2];  // bzwj1 = e^{+-bJ}, bzwj2 = e^{-+bJ k}, bzwj3 = e^{-+bJ k2}
  S outfac;            // factor applied to the output (XDT: nf; else 1)
};

template <class S>
void consts_bz(OpConsts<S> &c, S bJ, S k, S k2)
{
  using std::exp;
  c.b1[0] = exp(bJ);
  c.b1[1] = S(1.0) / c.b1[0];
  c.b2[0] = exp(-(bJ * k));
  c.b2[1] = S(1.0) / c.b2[0];
  c.b3[0] = exp(-(bJ * k2));
  c.b3[1] = S(1.0) / c.b3[0];
  c.outfac = S(1.0);
}

// normfactor exponent of the X-type 2D operators in units of betaJ
// (TMat/tmat.cpp:48-55, TMat/tmatreduce.cpp:349-359)
template <class S>
S nfexp_x2d(int L, S k)
{
  return val(k) < 0.5 ? S((double)L) * (k - S(2.0)) : -(S((double)L) * (k + S(1.0)));
}
// 3D (TMat/tmat3d.cpp:117-127)
template <class S>
S nfexp_x3d(int size, S k)
{
  return val(k) < 0.5 ? S((double)size) * (k - S(3.0)) : -(S((double)size) * (k + S(2.0)));
}

// XAT: Tv^1/2 with betaJ/2 (genTvec, TMat/tmat.hpp:208-310), sqrt(nf) folded
// into the first table.  bc = 0, 1, 5, 9, 13, 17.
template <class S>
OpConsts<S> consts_xat(int L, int bc, S bJ, S k, S h)
{
  using std::exp;
  using std::sqrt;
  OpConsts<S> c;
  consts_bz(c, bJ, k, S(0.0));
  c.nf = exp(bJ * nfexp_x2d(L, k));
  const S sqnf = sqrt(c.nf);
  const S J1 = S(-0.5) * bJ, J2 = S(0.5) * bJ * k, hJ = S(-0.5) * bJ * h;
  int c1 = L, c2 = L;
  switch (bc) {
    case 1: c1 = L - 1; c2 = L - 2; break;
    case 5: case 9: c1 = L + 1; c2 = L; break;
    case 13: case 17: c1 = L + 1; c2 = L + 2; break;
    default: break;
  }
  for (int n = 0; n < 64; ++n) {
    c.tab[n] = sqnf * exp(J1 * S(2.0 * n - c1));
    c.tab[64 + n] = exp(J2 * S(2.0 * n - c2));
    c.tab[128 + n] = exp(hJ * S(2.0 * n - L));
  }
  return c;
}

// XDT (TMat/tmat.cpp:333-340 normfactor; genTvec with betaJ/2,
// TMat/tmat.hpp:437-504); the normfactor is applied to the output.
template <class S>
OpConsts<S> consts_xdt(int L, int bc, S bJ, S k, S h)
{
  using std::exp;
  OpConsts<S> c;
  consts_bz(c, bJ, k, S(0.0));
  const S Ld((double)L);
  c.nf = val(k) < 0.5 ? exp(Ld * bJ * S(2.0) * (k - S(1.0))) : exp(-(Ld * bJ * S(2.0) * k));
  c.outfac = c.nf;
  const S J1 = S(-0.5) * bJ, hJ = S(-0.5) * bJ * h;
  const int c1 = bc == 0 ? L : L + 1;
  for (int n = 0; n < 64; ++n) {
    c.tab[n] = exp(J1 * S(2.0 * n - c1));
    c.tab[64 + n] = S(1.0);
    c.tab[128 + n] = exp(hJ * S(2.0 * n - L));
  }
  return c;
}

// ZBT (TMat/tmat.cpp:491-511 normfactor; genTvec with betaJ/2,
// TMat/tmat.hpp:537-556); the half normfactor is folded into the first table.
template <class S>
OpConsts<S> consts_zbt(int L, S bJ, S k, S k2, S h)
{
  using std::exp;
  OpConsts<S> c;
  consts_bz(c, bJ, k, k2);
  const S Ld((double)L);
  S nfe;
  if (val(k) + val(k2) < 0.5) nfe = Ld * S(2.0) * (k + k2 - S(1.0));
  else if (val(k) < 0.5 && val(k2) - val(k) < 0.5) nfe = -Ld;
  else if (val(k2) < 2 * val(k)) nfe = -(Ld * S(2.0) * k);
  else nfe = Ld * S(2.0) * (k - k2);
  const S half = exp(S(0.5) * bJ * nfe);
  c.nf = half * half;
  const S J1 = S(-0.5) * bJ, J3 = S(0.5) * k2 * bJ, hJ = S(-0.5) * bJ * h;
  for (int n = 0; n < 64; ++n) {
    c.tab[n] = half * exp(J1 * S(2.0 * n - L));
    c.tab[64 + n] = exp(J3 * S(2.0 * n - L));
    c.tab[128 + n] = exp(hJ * S(2.0 * n - L));
  }
  return c;
}

// Full-space 3D layer (energies of TMat/tmat3d.cpp:188-199), sqrt(nf) folded.
template <class S>
OpConsts<S> consts_xat3d(int size, S bJ, S k, S h)
{
  using std::exp;
  OpConsts<S> c;
  consts_bz(c, bJ, k, S(0.0));
  const S e = nfexp_x3d(size, k);
  c.nf = exp(bJ * e);
  const S hb = S(0.5) * bJ, Jn = hb * e;
  for (int n = 0; n < 64; ++n) {
    c.tab[n] = exp(S(-2.0) * hb * S((double)(n - size)) + Jn);
    c.tab[64 + n] = exp(hb * k * S(2.0 * n - size));
    c.tab[128 + n] = exp(-(hb * h) * S(2.0 * n - size));
  }
  return c;
}

}  // namespace ltm
