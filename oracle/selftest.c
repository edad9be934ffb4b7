/*
 * selftest.c -- runs every oracle operator kind at small sizes through
 * creation, T.x, T^T.x, the observables, the orbit enumerations and the
 * configuration sampler, for the AddressSanitizer / UndefinedBehaviorSanitizer
 * and ThreadSanitizer builds of the oracle (oracle/Makefile).  Checks that
 * every result is finite and that <y, T x> = <T^T y, x>.  Test
 * infrastructure only.  Exit status 0 when every case passes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "tm_oracle.h"

static double urand(unsigned long long *s)
{
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return ((*s >> 11) + 0.5) / 9007199254740992.0;
}

struct Case {
  int kind, s1, s2, bc;
  double bJ, k, k2, h;
  int sampler, gen;
};

static int run_case(const struct Case *c)
{
  or_op *op = or_op_create(c->kind, c->s1, c->s2, c->bc, c->bJ, c->k, c->k2, c->h);
  if (!op) {
    printf("FAIL kind %d size %d: rejected\n", c->kind, c->s1);
    return 1;
  }
  const int64_t n = or_op_ncol(op);
  double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
  double *tx = malloc(sizeof(double) * n), *ty = malloc(sizeof(double) * n);
  unsigned long long seed = 12345 + (unsigned long long)c->kind * 7 + c->s1;
  for (int64_t i = 0; i < n; ++i) {
    x[i] = urand(&seed);
    y[i] = urand(&seed);
  }
  or_op_apply(op, x, tx, 0);
  or_op_apply(op, y, ty, 1);
  double a = 0, b = 0, m = 0;
  int bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!isfinite(tx[i]) || !isfinite(ty[i])) bad = 1;
    a += y[i] * tx[i];
    b += ty[i] * x[i];
    m = fmax(m, fabs(tx[i]));
  }
  /* Y2AT's transposed instance is the operator itself, as the reference's
   * (TMat/tmat.hpp:382-384), not its transpose: no adjoint identity */
  const int adj = c->kind != OR_Y2AT;
  if (bad || (adj && !(fabs(a - b) <= 1e-12 * fabs(a)))) {
    printf("FAIL kind %d size %d: adjoint %.17g vs %.17g (finite %d)\n", c->kind, c->s1, a, b,
           !bad);
    bad = 1;
  }
  double out[8];
  int nout = 0;
  or_op_moments(op, c->gen ? y : NULL, x, out, &nout);
  for (int k = 0; k < nout && k < 4; ++k)
    if (!isfinite(out[k])) {
      printf("FAIL kind %d size %d: moment %d\n", c->kind, c->s1, k);
      bad = 1;
    }
  if (c->sampler) {
    const int nchain = 3, length = 4;
    double u[3 * 8];
    for (int i = 0; i < 3 * 8; ++i) u[i] = urand(&seed);
    uint64_t st[3 * 4];
    if (or_sample_configs(op, x, c->gen ? y : NULL, u, nchain, length, st) != 0) {
      printf("FAIL kind %d size %d: sampler\n", c->kind, c->s1);
      bad = 1;
    }
  }
  or_op_destroy(op);
  free(x);
  free(y);
  free(tx);
  free(ty);
  if (!bad) printf("ok kind %d size %d x %d bc %d (n %lld)\n", c->kind, c->s1, c->s2, c->bc,
                   (long long)n);
  return bad;
}

int main(void)
{
  static const struct Case cases[] = {
      {OR_XAT, 8, 0, 0, 0.5, 0.6, 0, 0.1, 1, 0},   {OR_XAT, 7, 0, 1, 0.5, 0.6, 0, 0.1, 0, 0},
      {OR_XAT, 8, 0, 5, 0.5, 0.3, 0, 0.0, 0, 0},   {OR_XAT, 8, 0, 17, 0.5, 0.3, 0, 0.1, 0, 0},
      {OR_XDT, 8, 0, 0, 0.6, 0.7, 0, 0.1, 0, 0},   {OR_XDT, 8, 0, 9, 0.6, 0.7, 0, 0.1, 0, 0},
      {OR_ZBT, 8, 0, 0, 0.5, 0.5, 0.3, 0.1, 0, 1}, {OR_XAM2D, 10, 0, 0, 0.4, 0.5, 0, 0, 1, 0},
      {OR_XAM2D, 9, 0, 1, 0.4, 0.5, 0, 0, 1, 0},   {OR_XAM2D, 10, 0, 2, 0.4, 0.5, 0, 0.1, 1, 0},
      {OR_XAM3D, 3, 3, 0, 0.25, 0.4, 0, 0, 0, 0},  {OR_XAM3D, 3, 3, 0, 0.25, 0.0, 0, 0, 0, 0},
      {OR_XAT3D, 3, 3, 0, 0.25, 0.4, 0, 0.05, 0, 0}, {OR_XDM, 10, 0, 0, 0.45, 0.35, 0, 0, 1, 0},
      {OR_XDM, 10, 0, 2, 0.45, 0.35, 0, 0.1, 1, 0},  {OR_ZBM, 10, 0, 0, 0.4, 0.5, 0.2, 0, 1, 1},
      {OR_ZBM, 10, 0, 2, 0.4, 0.5, 0.2, 0.1, 1, 1},  {OR_YAT, 4, 0, 0, 0.5, 0.3, 0, 0.05, 1, 1},
      {OR_YAT, 5, 0, 0, 0.5, 0.3, 0, 0.05, 1, 1},    {OR_Y2AT, 4, 0, 0, 0.5, 0.3, 0, 0.05, 0, 1},
      {OR_YAM, 5, 0, 0, 0.45, 0.3, 0, 0, 1, 1},      {OR_YAM, 5, 0, 2, 0.45, 0.3, 0, 0.1, 1, 1},
      {OR_YBM, 5, 0, 0, 0.45, 0.3, 0.2, 0, 1, 1},
  };
  int bad = 0;
  for (size_t i = 0; i < sizeof cases / sizeof cases[0]; ++i) bad += run_case(&cases[i]);
  /* orbit enumerations at sizes beyond the cases above */
  const int64_t m = or_orbits_x2d(14, 0, 1, NULL, NULL);
  uint64_t *r = malloc(sizeof(uint64_t) * m);
  uint16_t *d = malloc(sizeof(uint16_t) * m);
  if (or_orbits_x2d(14, 0, 1, r, d) != m || m != 362) {
    printf("FAIL orbits x2d L=14: %lld\n", (long long)m);
    ++bad;
  }
  free(r);
  free(d);
  printf(bad ? "selftest FAILED (%d)\n" : "selftest ok\n", bad);
  return bad ? 1 : 0;
}
