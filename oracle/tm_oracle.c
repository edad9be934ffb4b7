/*
 * tm_oracle.c -- CPU restatement of the LattTMat transfer-matrix operators.
 *
 * TEST INFRASTRUCTURE ONLY (see tm_oracle.h): the checker for the HIP product
 * path, never part of it.  Every function cites the reference file:line whose
 * algorithm it restates; loop structure and floating-point operation order
 * follow the reference so that results agree to rounding.
 */
#include "tm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define POP(x) __builtin_popcountll(x)

int or_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* bit primitives                                                            */
/* ------------------------------------------------------------------------ */

/* TMat/utility.hpp:91-94: rotate right by d within n bits */
uint64_t or_offset_state(uint64_t s, int n, int d)
{
  return (s >> d) | ((s & ((1ULL << d) - 1)) << (n - d));
}

/* TMat/utility.hpp:99-102: rotate left by d within n bits */
uint64_t or_lffset_state(uint64_t s, int n, int d)
{
  return ((s << d) & ((1ULL << n) - 1)) | (s >> (n - d));
}

/* TMat/utility.cpp:17-31: reverse the low n bits */
uint64_t or_bitreversal(uint64_t s, int n)
{
  s = ((s >> 1) & 0x5555555555555555ULL) | ((s & 0x5555555555555555ULL) << 1);
  s = ((s >> 2) & 0x3333333333333333ULL) | ((s & 0x3333333333333333ULL) << 2);
  s = ((s >> 4) & 0x0f0f0f0f0f0f0f0fULL) | ((s & 0x0f0f0f0f0f0f0f0fULL) << 4);
  s = ((s >> 8) & 0x00ff00ff00ff00ffULL) | ((s & 0x00ff00ff00ff00ffULL) << 8);
  s = ((s >> 16) & 0x0000ffff0000ffffULL) | ((s & 0x0000ffff0000ffffULL) << 16);
  s = (s >> 32) | (s << 32);
  return s >> (64 - n);
}

/* TMat/utility.hpp:83-86 */
int or_spinbzexp(uint64_t a, uint64_t b, int n)
{
  return 2 * POP(a ^ b) - n;
}

/* TMat/utility.hpp:36-55: a[0]^(-b) by repeated squaring of a[0] or a[1] */
double or_powdi(const double a[2], int b)
{
  double result = 1.0;
  const double *p = a;
  if (b < 0) b = -b; else ++p;
  double t = *p;
  while (b) {
    if (b & 1) result *= t;
    t *= t;
    b >>= 1;
  }
  return result;
}

/* TMat/utility.cpp:52-63: rotate each of the s1 rows (s2 bits) right by d */
uint64_t or_rffset3d(uint64_t s, int s1, int s2, int d)
{
  uint64_t mask = (1ULL << s2) - 1, r = 0;
  for (int rp = 0; rp < s1; ++rp) {
    uint64_t t = or_offset_state(s & mask, s2, d);
    r |= t << (rp * s2);
    s >>= s2;
  }
  return r;
}

/* TMat/utility.cpp:68-73: rotate the rows down by d (whole-word rotate by d*s2) */
uint64_t or_dffset3d(uint64_t s, int s1, int s2, int d)
{
  uint64_t mask = (1ULL << (s2 * d)) - 1, r = s >> (s2 * d);
  r |= (s & mask) << ((s1 - d) * s2);
  return r;
}

/* TMat/utility.cpp:78-91: rotate an n x n layer by 90 degrees */
uint64_t or_rotate3d(uint64_t s, int n)
{
  uint64_t r = 0;
  for (int rp = 0; rp < n; ++rp)
    for (int rq = 0; rq < n; ++rq) {
      int loc = rp * n + rq;
      if (s & (1ULL << loc)) r |= 1ULL << ((rq + 1) * n - (rp + 1));
    }
  return r;
}

/* TMat/utility.cpp:97-120: which=0 reverses each row, which=1 reverses row order */
uint64_t or_reflect3d(uint64_t s, int s1, int s2, int which)
{
  uint64_t mask = (1ULL << s2) - 1, r = 0;
  if (!which) {
    for (int rp = 0; rp < s1; ++rp) {
      uint64_t t = or_bitreversal(s & mask, s2);
      r |= t << (rp * s2);
      s >>= s2;
    }
  } else {
    for (int rp = 0; rp < s1; ++rp) {
      r <<= s2;
      r |= s & mask;
      s >>= s2;
    }
  }
  return r;
}

/* TMat/utility.hpp:125-128 */
static inline double bzexpn(int nspin, int nanti, double J)
{
  return J * (2.0 * nanti - nspin);
}

/* ------------------------------------------------------------------------ */
/* genTvec variants                                                          */
/* ------------------------------------------------------------------------ */

/* TMat/tmat.hpp:208-310 (XATMatVecProd2D::genTvec) */
void or_tvec_xat(int L, int bc, double betaJ, double kappa, double h, double *V)
{
  uint64_t nstate = 1ULL << L, mask = nstate - 1;
  double J1 = -betaJ, J2 = betaJ * kappa, hJ = -betaJ * h;
#pragma omp parallel for schedule(static)
  for (int64_t is = 0; is < (int64_t)nstate; ++is) {
    uint64_t s = (uint64_t)is, s1s, s2s;
    int n1, n2, n0;
    double b1, b2, b0;
    switch (bc) {
      case 0:
        s1s = or_offset_state(s, L, 1);
        s2s = or_offset_state(s, L, 2);
        n1 = POP(s ^ s1s); n2 = POP(s ^ s2s); n0 = POP(s);
        b1 = bzexpn(L, n1, J1); b2 = bzexpn(L, n2, J2); b0 = bzexpn(L, n0, hJ);
        break;
      case 1:
        s1s = s >> 1; s2s = s >> 2;
        n1 = POP((s & (mask >> 1)) ^ s1s); n2 = POP((s & (mask >> 2)) ^ s2s); n0 = POP(s);
        b1 = bzexpn(L - 1, n1, J1); b2 = bzexpn(L - 2, n2, J2); b0 = bzexpn(L, n0, hJ);
        break;
      case 5:
      case 9:
        s1s = (bc == 5) ? (s << 1) : ((s << 1) | 1);
        s2s = s >> 1;
        n1 = POP(s ^ s1s); n2 = POP((s1s & mask) ^ s2s); n0 = POP(s);
        b1 = bzexpn(L + 1, n1, J1); b2 = bzexpn(L, n2, J2); b0 = bzexpn(L, n0, hJ);
        break;
      case 13:
      case 17:
        s1s = (bc == 13) ? (s << 1) : ((s << 1) | 1);
        s2s = (bc == 13) ? (s << 2) : ((s << 2) | 3);
        n1 = POP(s ^ s1s); n2 = POP(s ^ s2s); n0 = POP(s);
        b1 = bzexpn(L + 1, n1, J1); b2 = bzexpn(L + 2, n2, J2); b0 = bzexpn(L, n0, hJ);
        break;
      default:
        b1 = b2 = b0 = NAN;
    }
    V[s] = exp(b1 + b2 + b0);
  }
}

/* TMat/tmat.hpp:437-504 (XDTMatVecProd2D::genTvec) */
void or_tvec_xdt(int L, int bc, double betaJ, double h, double *V)
{
  uint64_t nstate = 1ULL << L;
  double J1 = -betaJ, hJ = -betaJ * h;
#pragma omp parallel for schedule(static)
  for (int64_t is = 0; is < (int64_t)nstate; ++is) {
    uint64_t s = (uint64_t)is;
    int n1, n0 = POP(s);
    double b1;
    switch (bc) {
      case 0: n1 = POP(s ^ or_offset_state(s, L, 1)); b1 = bzexpn(L, n1, J1); break;
      case 1: n1 = POP((s & ((nstate - 1) >> 1)) ^ (s >> 1)); b1 = bzexpn(L - 1, n1, J1); break;
      case 5: n1 = POP((s << 1) ^ s); b1 = bzexpn(L + 1, n1, J1); break;
      case 9: n1 = POP(((s << 1) | 1) ^ s); b1 = bzexpn(L + 1, n1, J1); break;
      default: b1 = NAN;
    }
    V[s] = exp(b1 + bzexpn(L, n0, hJ));
  }
}

/* TMat/tmat.hpp:537-556 (ZBTMatVecProd2D::genTvec) */
void or_tvec_zbt(int L, double betaJ, double kappa2, double h, double *V)
{
  uint64_t nstate = 1ULL << L;
  double J1 = -betaJ, J3 = kappa2 * betaJ, hJ = -betaJ * h;
#pragma omp parallel for schedule(static)
  for (int64_t is = 0; is < (int64_t)nstate; ++is) {
    uint64_t s = (uint64_t)is;
    int n1 = POP(s ^ or_offset_state(s, L, 1));
    int n2 = POP(s ^ or_offset_state(s, L, 2));
    int n0 = POP(s);
    V[s] = exp(bzexpn(L, n1, J1) + bzexpn(L, n2, J3) + bzexpn(L, n0, hJ));
  }
}

/* ------------------------------------------------------------------------ */
/* orbit enumeration                                                         */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t *w; } bitset;
static int bs_init(bitset *b, int nbits)
{
  uint64_t nchunk = nbits <= 6 ? 1 : (1ULL << (nbits - 6));
  b->w = (uint64_t *)calloc(nchunk, sizeof(uint64_t));
  return b->w != NULL;
}
static inline int bs_get(const bitset *b, uint64_t s) { return (b->w[s >> 6] >> (s & 63)) & 1; }
static inline void bs_set(bitset *b, uint64_t s) { b->w[s >> 6] |= 1ULL << (s & 63); }

static void push_rep(uint64_t *reps, uint16_t *dups, int64_t idx, uint64_t s)
{
  if (reps) reps[idx] = s;
  if (dups) dups[idx] = 0;
}

/* TMat/tmatreduce.cpp:19-111 (genX2DDuplist) */
int64_t or_orbits_x2d(int L, int bc, int censym, uint64_t *reps, uint16_t *dups)
{
  uint64_t nstate = 1ULL << L, mask = nstate - 1;
  bitset buf;
  if (!bs_init(&buf, L)) return -1;
  int64_t m = 0;
  for (uint64_t now = 0; now < nstate; ++now) {
    if (bs_get(&buf, now)) continue;
    push_rep(reps, dups, m, now);
    uint16_t d = 0;
    if (bc == 0) {
      uint64_t nx;
      for (nx = now; !bs_get(&buf, nx); nx = (nx >> 1) | ((nx & 1) << (L - 1))) {
        ++d; bs_set(&buf, nx);
        if (censym) {
          uint64_t fm = ~nx & mask;
          if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
        }
      }
      for (nx = or_bitreversal(now, L); !bs_get(&buf, nx); nx = (nx >> 1) | ((nx & 1) << (L - 1))) {
        ++d; bs_set(&buf, nx);
        if (censym) {
          uint64_t fm = ~nx & mask;
          if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
        }
      }
    } else {
      bs_set(&buf, now); ++d;
      uint64_t nx = or_bitreversal(now, L);
      if (nx != now) { ++d; bs_set(&buf, nx); }
      if (censym) {
        uint64_t fm = ~now & mask;
        if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
        fm = ~nx & mask;
        if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
      }
    }
    if (dups) dups[m] = d;
    ++m;
  }
  free(buf.w);
  return m;
}

/* TMat/tmat3d.cpp:17-98 (genX3DDuplist) */
int64_t or_orbits_x3d(int s1, int s2, int censym, int rotate, uint64_t *reps, uint16_t *dups)
{
  int size = s1 * s2;
  uint64_t nstate = 1ULL << size, mask = nstate - 1;
  bitset buf;
  if (!bs_init(&buf, size)) return -1;
  int64_t m = 0;
  for (uint64_t now = 0; now < nstate; ++now) {
    if (bs_get(&buf, now)) continue;
    push_rep(reps, dups, m, now);
    uint16_t d = 0;
    uint64_t nr = now;
    int rot_status = 0;
    do {
      uint64_t nc = nr;
      int cen_status = 0;
      do {
        uint64_t nm1 = nc;
        int m1_status = 0;
        do {
          uint64_t nm2 = nm1;
          int m2_status = 0;
          do {
            for (uint64_t o1 = nm2; !bs_get(&buf, o1); o1 = or_rffset3d(o1, s1, s2, 1))
              for (uint64_t o2 = o1; !bs_get(&buf, o2); o2 = or_dffset3d(o2, s1, s2, 1)) {
                bs_set(&buf, o2);
                ++d;
              }
            if (0 == m2_status) { m2_status = 1; nm2 = or_reflect3d(nm2, s1, s2, 1); }
            else break;
          } while (!bs_get(&buf, nm2));
          if (0 == m1_status) { m1_status = 1; nm1 = or_reflect3d(nm1, s1, s2, 0); }
          else break;
        } while (!bs_get(&buf, nm1));
        if (0 == cen_status && censym) { cen_status = 1; nc = ~nc & mask; }
        else break;
      } while (!bs_get(&buf, nc));
      if (0 == rot_status && rotate) { rot_status = 1; nr = or_rotate3d(nr, s1); }
      else break;
    } while (!bs_get(&buf, nr));
    if (dups) dups[m] = d;
    ++m;
  }
  free(buf.w);
  return m;
}

/* TMat/tmatreduce.cpp:242-334 (genZ2DDuplist), periodic boundary: orbits of
 * the translations by two sites and of the mirror image
 * bitreversal(rotr(s, 1)) (i -> -i), times the spin flip if censym. */
int64_t or_orbits_z2d(int L, int censym, uint64_t *reps, uint16_t *dups)
{
  uint64_t nstate = 1ULL << L, mask = nstate - 1;
  bitset buf;
  if (!bs_init(&buf, L)) return -1;
  int64_t m = 0;
  for (uint64_t now = 0; now < nstate; ++now) {
    if (bs_get(&buf, now)) continue;
    push_rep(reps, dups, m, now);
    uint16_t d = 0;
    uint64_t nx;
    for (nx = now; !bs_get(&buf, nx); nx = (nx >> 2) | ((nx & 3) << (L - 2))) {
      ++d; bs_set(&buf, nx);
      if (censym) {
        uint64_t fm = ~nx & mask;
        if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
      }
    }
    for (nx = or_bitreversal((now >> 1) | ((now & 1) << (L - 1)), L); !bs_get(&buf, nx);
         nx = (nx >> 2) | ((nx & 3) << (L - 2))) {
      ++d; bs_set(&buf, nx);
      if (censym) {
        uint64_t fm = ~nx & mask;
        if (!bs_get(&buf, fm)) { ++d; bs_set(&buf, fm); }
      }
    }
    if (dups) dups[m] = d;
    ++m;
  }
  free(buf.w);
  return m;
}

/* TMat/utility.hpp:108-118: rotations of nowstate (right by one) until it
 * equals target; nspin if never */
static int findcyclic(uint64_t nowstate, uint64_t target, int nspin)
{
  uint64_t nx = nowstate;
  int rp;
  for (rp = 0; rp < nspin; ++rp) {
    if (nx == target) break;
    nx = (nx >> 1) | ((nx & 1) << (nspin - 1));
  }
  return rp;
}

typedef struct {
  uint64_t *reps; uint16_t *dups; int64_t *swp; unsigned char (*lead)[4];
  int64_t m, nlead, cap, capl;
} ylists;

static int ygrow(ylists *y)
{
  if (y->m < y->cap) return 1;
  int64_t c = y->cap ? 2 * y->cap : 1024;
  uint64_t *r = (uint64_t *)realloc(y->reps, sizeof(uint64_t) * c);
  if (!r) return 0;
  y->reps = r;
  uint16_t *d = (uint16_t *)realloc(y->dups, sizeof(uint16_t) * c);
  if (!d) return 0;
  y->dups = d;
  int64_t *w = (int64_t *)realloc(y->swp, sizeof(int64_t) * c);
  if (!w) return 0;
  y->swp = w;
  y->cap = c;
  return 1;
}

/* one rotation chain of genY2DDuplist: both layers rotated right together */
static void ychain(uint64_t start, int L, bitset *eqb, uint16_t *d, uint64_t *smallest)
{
  uint64_t lm = (1ULL << L) - 1, lf = start >> L, rt = start & lm, nx;
  for (nx = start; !bs_get(eqb, nx);) {
    uint64_t tmp = (rt << L) | lf;
    if (tmp < *smallest) *smallest = tmp;
    ++*d;
    bs_set(eqb, nx);
    lf = (lf >> 1) | ((lf & 1) << (L - 1));
    rt = (rt >> 1) | ((rt & 1) << (L - 1));
    nx = (lf << L) | rt;
  }
}

/* TMat/tmatreduce.cpp:114-239 (genY2DDuplist): orbits of two-layer states
 * under rotations and the reflection of both layers together (and the flip
 * of both); swp pairs an orbit with the orbit of its layer-swapped
 * representative; lead holds the periods of each distinct high layer. */
static int orbits_y2d(int L, int censym, ylists *y)
{
  uint64_t nstate = 1ULL << (2 * L), lm = (1ULL << L) - 1, mask = nstate - 1;
  bitset eqb;
  if (!bs_init(&eqb, 2 * L)) return 0;
  memset(y, 0, sizeof *y);
  for (uint64_t now = 0; now < nstate; ++now) {
    if (bs_get(&eqb, now)) continue;
    uint64_t lf = now >> L;
    if (y->m == 0 || (y->reps[y->m - 1] >> L) != lf) {
      if (y->nlead == y->capl) {
        int64_t c = y->capl ? 2 * y->capl : 256;
        unsigned char(*t)[4] = (unsigned char(*)[4])realloc(y->lead, 4 * (size_t)c);
        if (!t) goto fail;
        y->lead = t;
        y->capl = c;
      }
      unsigned char *ld = y->lead[y->nlead++];
      ld[0] = ld[1] = ld[2] = ld[3] = (unsigned char)L;
      int count = findcyclic((lf >> 1) | ((lf & 1) << (L - 1)), lf, L);
      if (count < L) ld[0] = (unsigned char)(++count);
      ld[1] = (unsigned char)findcyclic(or_bitreversal(lf, L), lf, L);
      if (censym && POP(lf) * 2 == L) {
        ld[2] = (unsigned char)findcyclic(~lf & lm, lf, L);
        ld[3] = (unsigned char)findcyclic(or_bitreversal(~lf & lm, L), lf, L);
      }
    }
    if (!ygrow(y)) goto fail;
    int64_t me = y->m++;
    y->swp[me] = me;
    y->reps[me] = now;
    uint16_t d = 0;
    uint64_t smallest = ((now & lm) << L) | lf;
    ychain(now, L, &eqb, &d, &smallest);
    ychain((or_bitreversal(now >> L, L) << L) | or_bitreversal(now & lm, L), L, &eqb, &d,
           &smallest);
    if (censym) {
      uint64_t ni = ~now & mask;
      ychain(ni, L, &eqb, &d, &smallest);
      ychain((or_bitreversal(ni >> L, L) << L) | or_bitreversal(ni & lm, L), L, &eqb, &d,
             &smallest);
    }
    y->dups[me] = d;
    if (smallest < now) {
      int64_t lo = 0, hi = me;  /* lower_bound over the ascending reps */
      while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (y->reps[mid] < smallest) lo = mid + 1; else hi = mid;
      }
      if (lo >= me || y->reps[lo] != smallest) goto fail;  /* tmatreduce.cpp:232-236 */
      y->swp[me] = lo;
      y->swp[lo] = me;
    }
  }
  free(eqb.w);
  return 1;
fail:
  free(eqb.w);
  return 0;
}

int64_t or_orbits_y2d(int L, int censym, uint64_t *reps, uint16_t *dups, int64_t *swp)
{
  ylists y;
  if (L < 1 || L > 15 || !orbits_y2d(L, censym, &y)) return -1;
  if (reps) memcpy(reps, y.reps, sizeof(uint64_t) * y.m);
  if (dups) memcpy(dups, y.dups, sizeof(uint16_t) * y.m);
  if (swp) memcpy(swp, y.swp, sizeof(int64_t) * y.m);
  int64_t m = y.m;
  free(y.reps); free(y.dups); free(y.swp); free(y.lead);
  return m;
}

/* ------------------------------------------------------------------------ */
/* operator construction                                                     */
/* ------------------------------------------------------------------------ */

static double nf_x2d(int L, double betaJ, double kappa, double *nfexp)
{
  /* TMat/tmat.cpp:48-55, TMat/tmatreduce.cpp:349-359 */
  double e = (kappa < 0.5) ? L * (kappa - 2.0) : -L * (kappa + 1.0);
  if (nfexp) *nfexp = e;
  return exp(betaJ * e);
}

static double nf_x3d(int size, double betaJ, double kappa, double *nfexp)
{
  /* TMat/tmat3d.cpp:117-127 */
  double e = (kappa < 0.5) ? size * (kappa - 3.0) : -size * (kappa + 2.0);
  if (nfexp) *nfexp = e;
  return exp(betaJ * e);
}

/* Energy exponent of a 3D layer, TMat/tmat3d.cpp:188-199 (without Jnormf) */
static double e3d_half(uint64_t s, int s1, int s2, double betaJ, double kappa, double h)
{
  int size = s1 * s2;
  double hb = 0.5 * betaJ;
  double J1 = -hb, J2 = hb * kappa, hJ = -hb * h;
  int n1a = POP(s ^ or_rffset3d(s, s1, s2, 1));
  int n1b = POP(s ^ or_dffset3d(s, s1, s2, 1));
  int n2 = POP(s ^ or_rffset3d(s, s1, s2, 2));
  int n0 = POP(s);
  return 2 * J1 * (n1a + n1b - size) + J2 * (2 * n2 - size) + hJ * (2 * n0 - size);
}

or_op *or_op_create(int kind, int size1, int size2, int bc,
                    double betaJ, double kappa, double kappa2, double h)
{
  or_op *op = (or_op *)calloc(1, sizeof(or_op));
  if (!op) return NULL;
  op->kind = kind;
  op->betaJ = betaJ; op->kappa = kappa; op->kappa2 = kappa2; op->h = h;
  op->bc = bc;
  op->bzwj1[0] = exp(betaJ); op->bzwj1[1] = 1.0 / op->bzwj1[0];
  op->bzwj2[0] = exp(-betaJ * kappa); op->bzwj2[1] = 1.0 / op->bzwj2[0];
  op->bzwj3[0] = exp(-betaJ * kappa2); op->bzwj3[1] = 1.0 / op->bzwj3[0];
  switch (kind) {
    case OR_XAT: {
      /* TMat/tmat.cpp:39-56 */
      int L = size1;
      if (L < 2 || L > 30) goto fail;
      if (!(bc == 0 || bc == 1 || bc == 5 || bc == 9 || bc == 13 || bc == 17)) goto fail;
      op->L = L; op->nstate = op->ncol = 1LL << L;
      op->nf = nf_x2d(L, betaJ, kappa, NULL);
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!op->tvec) goto fail;
      or_tvec_xat(L, bc, 0.5 * betaJ, kappa, h, op->tvec);
      break;
    }
    case OR_XDT: {
      /* TMat/tmat.cpp:322-341 */
      int L = size1;
      if (L < 3 || L > 30 || !(bc == 0 || bc == 5 || bc == 9)) goto fail;
      op->L = L; op->nstate = op->ncol = 1LL << L;
      op->nf = (kappa < 0.5) ? exp(L * betaJ * 2.0 * (kappa - 1.0)) : exp(-L * betaJ * 2.0 * kappa);
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!op->tvec) goto fail;
      or_tvec_xdt(L, bc, 0.5 * betaJ, h, op->tvec);
      break;
    }
    case OR_ZBT: {
      /* TMat/tmat.cpp:480-518 */
      int L = size1;
      if (L < 4 || L > 30 || (L & 1)) goto fail;
      op->L = L; op->nstate = op->ncol = 1LL << L;
      double nfe;
      if (kappa + kappa2 < 0.5) nfe = L * 2.0 * (kappa + kappa2 - 1.0);
      else if (kappa < 0.5 && kappa2 - kappa < 0.5) nfe = -L;
      else if (kappa2 < 2 * kappa) nfe = -L * 2.0 * kappa;
      else nfe = L * 2.0 * (kappa - kappa2);
      double half = exp(0.5 * betaJ * nfe);
      op->nf = half * half;
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!op->tvec) goto fail;
      or_tvec_zbt(L, 0.5 * betaJ, kappa2, h, op->tvec);
      for (int64_t i = 0; i < op->ncol; ++i) op->tvec[i] *= half;
      break;
    }
    case OR_XAM2D: {
      /* TMat/tmatreduce.cpp:338-365 and genTvec :367-404 */
      int L = size1;
      if (L < 2 || L > 30) goto fail;
      int extend = bc;
      op->bc = extend & 1;
      op->L = L; op->nstate = 1LL << L;
      op->nleadbit = (L < 10 || op->bc) ? 1 : 2 + (int)log2((double)L);
      op->censym = (h == 0 && (extend & 2) == 0);
      double nfe;
      op->nf = nf_x2d(L, betaJ, kappa, &nfe);
      int64_t m = or_orbits_x2d(L, op->bc, op->censym, NULL, NULL);
      if (m <= 0) goto fail;
      op->ncol = m;
      op->reps = (uint64_t *)malloc(sizeof(uint64_t) * m);
      op->dups = (uint16_t *)malloc(sizeof(uint16_t) * m);
      op->tvec = (double *)malloc(sizeof(double) * m);
      if (!op->reps || !op->dups || !op->tvec) goto fail;
      or_orbits_x2d(L, op->bc, op->censym, op->reps, op->dups);
      double hb = 0.5 * betaJ;
      double J1 = -hb, J2 = hb * kappa, hJ = -hb * h, Jn = hb * nfe;
      uint64_t mask = op->nstate - 1;
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t s = op->reps[rp];
        int n1, n2, n0 = POP(s);
        if (!op->bc) {
          n1 = POP(s ^ or_offset_state(s, L, 1));
          n2 = POP(s ^ or_offset_state(s, L, 2));
        } else {
          n1 = POP((s & (mask >> 1)) ^ (s >> 1));
          n2 = POP((s & (mask >> 2)) ^ (s >> 2));
        }
        double e = J1 * (2 * n1 - L) + J2 * (2 * n2 - L) + hJ * (2 * n0 - L) + Jn;
        op->tvec[rp] = exp(e) / sqrt((double)op->dups[rp]);
      }
      break;
    }
    case OR_XDM: {
      /* TMat/tmatreduce.cpp:1024-1053 and genTvec :1367-1398 */
      int L = size1;
      if (L < 4 || L > 30) goto fail;
      int extend = bc;
      op->bc = 0;  /* periodic only (tmatreduce.cpp:1047-1049) */
      op->L = L; op->nstate = 1LL << L;
      op->nleadbit = L < 10 ? 1 : 2 + (int)log2((double)L);
      op->censym = (h == 0 && (extend & 2) == 0);
      double nfe = (kappa < 0.5) ? L * 2.0 * (kappa - 1.0) : -L * 2.0 * kappa;
      op->nf = exp(betaJ * nfe);
      int64_t m = or_orbits_x2d(L, 0, op->censym, NULL, NULL);
      if (m <= 0) goto fail;
      op->ncol = m;
      op->reps = (uint64_t *)malloc(sizeof(uint64_t) * m);
      op->dups = (uint16_t *)malloc(sizeof(uint16_t) * m);
      op->tvec = (double *)malloc(sizeof(double) * m);
      if (!op->reps || !op->dups || !op->tvec) goto fail;
      or_orbits_x2d(L, 0, op->censym, op->reps, op->dups);
      double hb = 0.5 * betaJ;
      double J1 = -hb, hJ = -hb * h, Jn = hb * nfe;
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t st = op->reps[rp];
        int n1 = POP(st ^ or_offset_state(st, L, 1)), n0 = POP(st);
        double e = J1 * (2 * n1 - L) + hJ * (2 * n0 - L) + Jn;
        op->tvec[rp] = exp(e) / sqrt((double)op->dups[rp]);
      }
      break;
    }
    case OR_ZBM: {
      /* TMat/tmatreduce.cpp:1403-1454 and genTvec :1813-1851 */
      int L = size1;
      if (L < 4 || L > 30 || (L & 1)) goto fail;
      int extend = bc;
      op->bc = 0;
      op->L = L; op->nstate = 1LL << L;
      op->nleadbit = L < 10 ? 2 : 2 + ((int)log2((double)L) & ~1);
      op->censym = (h == 0 && (extend & 2) == 0);
      op->bzwj2b[0] = exp(-2 * betaJ * kappa);
      op->bzwj2b[3] = 1.0 / op->bzwj2b[0];
      op->bzwj2b[1] = op->bzwj2b[2] = 1.0;
      for (int rp = 0; rp < 16; ++rp) {
        uint64_t ra = (uint64_t)rp >> 2, rb = (uint64_t)rp & 3;
        op->bzwj2c[rp] = exp(betaJ * (kappa * (2 * POP(ra) - 2) + kappa2 * (2 * POP(rb) - 2)));
      }
      double nfe;
      if (kappa + kappa2 < 0.5) nfe = L * 2.0 * (kappa + kappa2 - 1.0);
      else if (kappa < 0.5 && kappa2 - kappa < 0.5) nfe = -L;
      else if (kappa2 < 2 * kappa) nfe = -L * 2.0 * kappa;
      else nfe = L * 2.0 * (kappa - kappa2);
      op->nf = exp(betaJ * nfe);
      int64_t m = or_orbits_z2d(L, op->censym, NULL, NULL);
      if (m <= 0) goto fail;
      op->ncol = m;
      op->reps = (uint64_t *)malloc(sizeof(uint64_t) * m);
      op->dups = (uint16_t *)malloc(sizeof(uint16_t) * m);
      op->tvec = (double *)malloc(sizeof(double) * m);
      if (!op->reps || !op->dups || !op->tvec) goto fail;
      or_orbits_z2d(L, op->censym, op->reps, op->dups);
      double hb = 0.5 * betaJ;
      double J1 = -hb, J3 = hb * kappa2, hJ = -hb * h, Jn = hb * nfe;
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t st = op->reps[rp];
        int n1 = POP(st ^ or_offset_state(st, L, 1));
        int n2 = POP(st ^ or_offset_state(st, L, 2));
        int n0 = POP(st);
        double e = bzexpn(L, n1, J1) + bzexpn(L, n2, J3) + bzexpn(L, n0, hJ);
        op->tvec[rp] = exp(e + Jn) / sqrt((double)op->dups[rp]);
      }
      break;
    }
    case OR_YAT: {
      /* TMat/tmat.cpp:187-205; Tvec = genXATmat2D({betaJ, 0, 0, h, L})
       * (tmat.hpp:361-366, tmat.cpp:668-684) flattened column-major:
       * Tvec[(j << L) | i] = V(i) exp(-betaJ (2 popcount(i ^ j) - L)),
       * V = XATMatVecProd2D::genTvec with the full betaJ and kappa 0. */
      int L = size1;
      if (L < 2 || L > 13) goto fail;
      op->L = L; op->nstate = op->ncol = 1LL << (2 * L);
      op->nf = (kappa < 0.5) ? exp(L * betaJ * (kappa - 2.0)) : exp(-L * betaJ * (kappa + 1.0));
      uint64_t n1 = 1ULL << L;
      double *V = (double *)malloc(sizeof(double) * n1);
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!V || !op->tvec) { free(V); goto fail; }
      or_tvec_xat(L, 0, betaJ, 0.0, h, V);
      double J1 = -betaJ;
#pragma omp parallel for schedule(static)
      for (int64_t j = 0; j < (int64_t)n1; ++j)
        for (uint64_t i = 0; i < n1; ++i)
          op->tvec[((uint64_t)j << L) | i] = V[i] * exp(bzexpn(L, POP(i ^ (uint64_t)j), J1));
      free(V);
      break;
    }
    case OR_Y2AT: {
      /* TMat/tmat.cpp:246-263 and genTvec tmat.hpp:396-416 (built with
       * betaJ/2: Tvec is Tv^(1/2)) */
      int L = size1;
      if (L < 2 || L > 13) goto fail;
      op->L = L; op->nstate = op->ncol = 1LL << (2 * L);
      op->bzwj2[0] = exp(betaJ * (-2.0 * kappa));
      op->bzwj2[1] = 1.0 / op->bzwj2[0];
      op->nf = (kappa < 0.5) ? exp(2 * L * betaJ * (kappa - 2.0)) : exp(-2 * L * betaJ * (kappa + 1.0));
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!op->tvec) goto fail;
      uint64_t mask = (1ULL << L) - 1;
      double J1 = -0.5 * betaJ, hJ = -0.5 * betaJ * h;
#pragma omp parallel for schedule(static)
      for (int64_t is = 0; is < op->ncol; ++is) {
        uint64_t st = (uint64_t)is, lf = st & mask, rt = st >> L;
        int n1 = POP(lf ^ or_offset_state(lf, L, 1)), n2 = POP(rt ^ or_offset_state(rt, L, 1));
        int n0 = POP(lf ^ rt);
        int fa = 2 * (n1 + n2 + n0) - 3 * L, fb = 2 * (POP(st) - L);
        op->tvec[st] = exp(J1 * fa + hJ * fb);
      }
      break;
    }
    case OR_YAM:
    case OR_YBM: {
      /* TMat/tmatreduce.cpp:661-714, genTvec :717-745 (layertype 0) and
       * genTvec1 :748-780 (layertype 1 = YBM, tmatreduce.cpp:1017-1021) */
      int L = size1;
      if (L < 4 || L > 12) goto fail;
      int extend = bc, lt1 = kind == OR_YBM;
      op->L = L; op->nstate = 1LL << (2 * L);
      op->censym = (h == 0.0 && (extend & 2) == 0);
      double nfe;
      if (!lt1) {
        nfe = (kappa < 0.5) ? L * (kappa - 2.0) : -L * (kappa + 1.0);
      } else if (kappa + kappa2 < 0.5) {
        nfe = L * 2.0 * (kappa + kappa2 - 1.0);
      } else if (kappa < 0.5 && kappa2 - kappa < 0.5) {
        nfe = -L;
      } else if (kappa2 < 2 * kappa) {
        nfe = -L * 2.0 * kappa;
      } else {
        nfe = L * 2.0 * (kappa - kappa2);
      }
      op->nf = exp(betaJ * nfe);
      ylists y;
      if (!orbits_y2d(L, op->censym, &y)) goto fail;
      op->reps = y.reps; op->dups = y.dups; op->swp = y.swp; op->lead = y.lead;
      op->nlead = y.nlead;
      op->ncol = y.m;
      op->tvec = (double *)malloc(sizeof(double) * y.m);
      if (!op->tvec) goto fail;
      uint64_t lm = (1ULL << L) - 1;
      double hb = 0.5 * betaJ, J1 = -hb, hJ = -hb * h, Jn = hb * nfe;
      double k1J = hb * kappa, k2J = hb * kappa2;
      for (int64_t rp = 0; rp < y.m; ++rp) {
        uint64_t st = op->reps[rp], lf = st >> L, rt = st & lm;
        int n1 = POP(lf ^ or_offset_state(lf, L, 1)), n2 = POP(lf ^ rt), n0 = POP(lf);
        double e;
        if (!lt1) {
          e = J1 * (2 * (n1 + n2 - L)) + hJ * (2 * n0 - L) + Jn;
        } else {
          int b1 = POP(lf ^ or_offset_state(lf, L, 2));
          int d1 = POP(lf ^ or_offset_state(rt, L, 1)), d2 = POP(lf ^ or_lffset_state(rt, L, 1));
          e = J1 * (2 * (n1 + n2 - L)) + k1J * (2 * b1 - L) + k2J * (2 * (d1 + d2 - L)) +
              hJ * (2 * n0 - L) + Jn;
        }
        op->tvec[rp] = exp(e) / sqrt((double)op->dups[rp]);
      }
      break;
    }
    case OR_XAM3D: {
      /* TMat/tmat3d.cpp:102-134 and genTvec :179-201 */
      int size = size1 * size2;
      if (size1 < 2 || size2 < 3 || size > 30) goto fail;
      op->size1 = size1; op->size2 = size2; op->L = size;
      op->bc = bc & 1;
      op->nstate = 1LL << size;
      op->nleadbit = size < 10 ? 1 : 2 + (int)log2((double)size);
      op->censym = (h == 0);
      op->ising = (kappa == 0);
      op->rotate = op->ising && size1 == size2;
      double nfe;
      op->nf = nf_x3d(size, betaJ, kappa, &nfe);
      int64_t m = or_orbits_x3d(size1, size2, op->censym, op->rotate, NULL, NULL);
      if (m <= 0) goto fail;
      op->ncol = m;
      op->reps = (uint64_t *)malloc(sizeof(uint64_t) * m);
      op->dups = (uint16_t *)malloc(sizeof(uint16_t) * m);
      op->tvec = (double *)malloc(sizeof(double) * m);
      if (!op->reps || !op->dups || !op->tvec) goto fail;
      or_orbits_x3d(size1, size2, op->censym, op->rotate, op->reps, op->dups);
      double Jn = 0.5 * betaJ * nfe;
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        double e = e3d_half(op->reps[rp], size1, size2, betaJ, kappa, h) + Jn;
        op->tvec[rp] = exp(e) / sqrt((double)op->dups[rp]);
      }
      break;
    }
    case OR_XAT3D: {
      /* Full-space counterpart of XAMMatVecProd3D: same layer energies
       * (TMat/tmat3d.cpp:188-199), same normfactor (:117-127), Th = (x)K. */
      int size = size1 * size2;
      if (size1 < 2 || size2 < 3 || size > 30) goto fail;
      op->size1 = size1; op->size2 = size2; op->L = size;
      op->nstate = op->ncol = 1LL << size;
      double nfe;
      op->nf = nf_x3d(size, betaJ, kappa, &nfe);
      op->tvec = (double *)malloc(sizeof(double) * op->ncol);
      if (!op->tvec) goto fail;
      double Jn = 0.5 * betaJ * nfe;
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < op->ncol; ++s)
        op->tvec[s] = exp(e3d_half((uint64_t)s, size1, size2, betaJ, kappa, h) + Jn);
      break;
    }
    default:
      goto fail;
  }
  return op;
fail:
  or_op_destroy(op);
  return NULL;
}

void or_op_destroy(or_op *op)
{
  if (!op) return;
  free(op->tvec);
  free(op->reps);
  free(op->dups);
  free(op->swp);
  free(op->lead);
  free(op);
}

/* TMat/utility.cpp:33-47 (open boundary: the mask nspin-1 as written there) */
static int or_nnode(uint64_t s, int L, int bc)
{
  uint64_t sh = s >> 1;
  if (!bc) sh |= (s & 1ull) << (L - 1);
  else sh |= (s & (uint64_t)(L - 1)) << 1;
  return POP(s ^ sh);
}

void or_op_moments(const or_op *op, const double *ql, const double *qr, double out[8], int *nout)
{
  const int L = op->L;
  const int64_t n = op->ncol;
  const uint64_t lmask = (L >= 64) ? ~0ull : (1ull << L) - 1;
  int k;
  if (!ql) ql = qr;
  for (k = 0; k < 8; ++k) out[k] = 0.0;
  *nout = 4;
  switch (op->kind) {
    case OR_Y2AT:
      return;  /* tmat.cpp:312-317 */
    case OR_XAT: case OR_XDT: case OR_XAM2D: case OR_XAT3D: case OR_XAM3D: case OR_XDM: {
      /* SymMatPhys::moments (tmatphys.cpp:13-37) / XAMMatVecProd2D::moments
       * (tmatreduce.cpp:565-590): weight |q|^2, divide by size */
      const int red = op->kind == OR_XAM2D || op->kind == OR_XAM3D || op->kind == OR_XDM;
      /* node boundary: XAT its own code (tmat.cpp:39), XAM ext & 1
       * (tmatreduce.cpp:338), XDT the MatPhys default 0 (tmat.cpp:320-323) */
      const int nbc = op->kind == OR_XAT ? op->bc : op->kind == OR_XAM2D ? (op->bc & 1) : 0;
      double ext[4] = {0, 0, 0, 0};
      int64_t i;
      for (i = 0; i < n; ++i) {
        const uint64_t s = red ? op->reps[i] : (uint64_t)i;
        const double dpara = 2.0 * POP(s) - L;
        const double w = ql[i] * ql[i];
        double dm = dpara;
        for (k = 0; k < 4; ++k) {
          out[k] += dm * w;
          dm *= dpara;
        }
        if (op->kind == OR_XDM) {  /* tmatreduce.cpp:1289-1320 */
          const double ms = 2.0 * POP((s ^ 0xaaaaaaaaaaaaaaaaull) & lmask) - L;
          ext[0] += fabs(ms) * w;
          ext[1] += fabs(dpara) * w;
          ext[2] += ms * ms * w;
          ext[3] += dpara * dpara * w;
        } else if (op->kind != OR_XAT3D && op->kind != OR_XAM3D) {  /* tmatphys.cpp:39-57 */
          const double nd = or_nnode(s, L, nbc);
          ext[0] += nd * w;
          ext[1] += nd * nd * w;
        }
      }
      for (k = 0; k < 4; ++k) out[k] /= L;
      if (op->kind == OR_XDM) {
        out[4] = ext[0] / L;
        out[5] = ext[1] / L;
        out[6] = ext[2] / ((double)L * L);
        out[7] = ext[3] / ((double)L * L);
        *nout = 8;
      } else if (op->kind == OR_XAT3D || op->kind == OR_XAM3D) {
        out[4] = out[5] = NAN;  /* tmat3d.cpp:165-177 */
        *nout = 6;
      } else {
        out[4] = ext[0];
        out[5] = ext[1];
        *nout = 6;
      }
      return;
    }
    default: {
      /* GenMatPhys (tmatphys.cpp:71-98: low layer of a 4^L state),
       * ZMatVecProd2D (:106-132), ZBM (tmatreduce.cpp:1748-1775), YAM/YBM
       * (tmatreduce.cpp:882-909: the representative's high layer):
       * weight ql qr, divide by size * sum ql qr; no extra observables */
      double dnorm = 0;
      int64_t i;
      for (i = 0; i < n; ++i) {
        uint64_t s = op->reps ? op->reps[i] : (uint64_t)i;
        if (op->kind == OR_YAT) s &= lmask;
        if (op->kind == OR_YAM || op->kind == OR_YBM) s >>= L;
        const double dpara = 2.0 * POP(s) - L;
        const double w = ql[i] * qr[i];
        double dm = dpara;
        dnorm += w;
        for (k = 0; k < 4; ++k) {
          out[k] += dm * w;
          dm *= dpara;
        }
      }
      for (k = 0; k < 4; ++k) out[k] /= L * dnorm;
      return;
    }
  }
}

/* std::upper_bound over an ascending cumulative array: first i with
 * cum[i] > t; n if none (then the caller clamps). */
static int64_t or_upper_bound(const double *cum, int64_t n, double t)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (cum[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

/* the last index with a positive increment of cum (0 if none) */
static int64_t or_last_positive(const double *cum, int64_t n)
{
  int64_t i;
  for (i = n - 1; i > 0; --i)
    if (cum[i] > cum[i - 1]) return i;
  return 0;
}

static int64_t or_draw(const double *cum, int64_t n, double u)
{
  const int64_t i = or_upper_bound(cum, n, u * cum[n - 1]);
  return i < n ? i : or_last_positive(cum, n);
}

static int or_cmp_u64(const void *a, const void *b)
{
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* exponent of an orbit member's weight (a the current state, s the member);
 * *ok = 0 for a member the move cannot reach */
static double or_member_cross(const or_op *op, uint64_t a, uint64_t s, int *ok)
{
  const int L = op->L;
  const uint64_t lm = (1ull << L) - 1;
  const double i0 = op->betaJ, i1 = -op->kappa * op->betaJ, i2 = -op->kappa2 * op->betaJ;
  *ok = 1;
  if (op->kind == OR_YAM || op->kind == OR_YBM) {  /* dumpreduce.cpp:213-221 */
    if ((s & lm) != (a >> L)) {
      *ok = 0;
      return 0.0;
    }
    return i1 * (L - 2.0 * POP((a & lm) ^ (s >> L)));
  }
  if (op->kind == OR_ZBM) {  /* ZBM2DDumpConfig::crossintr, dumpreduce.cpp:389-407 */
    double x = i2 * (L - 2.0 * POP(a ^ s));
    uint64_t ra = a, rb;
    int rp, nanti = 0;
    x += i1 * (L - 2.0 * POP(a ^ or_offset_state(s, L, 2)));
    x += i1 * (L - 2.0 * POP(a ^ or_lffset_state(s, L, 2)));
    rb = s << 1 | s >> (L - 1);
    for (rp = 0; rp < L / 2; ++rp) {
      nanti += (int)((ra ^ rb) & 1);
      rb >>= 2;
      nanti += (int)((ra ^ rb) & 1);
      ra >>= 2;
    }
    return x + i0 * (L - 2.0 * nanti);
  }
  {
    double x = i0 * (L - 2.0 * POP(a ^ s));
    if (op->kind == OR_XDM)  /* dumpreduce.cpp:509-514 */
      x += i1 * (L - 2.0 * POP(a ^ or_offset_state(s, L, 1)))
           + i1 * (L - 2.0 * POP(a ^ or_lffset_state(s, L, 1)));
    return x;
  }
}

/* the members of the orbit of rep, ascending; returns their count */
static int or_orbit_members(const or_op *op, uint64_t rep, uint64_t *mem)
{
  const int L = op->L;
  const uint64_t lm = (1ull << L) - 1;
  int m = 0, k, r;
  if (op->kind == OR_YAM || op->kind == OR_YBM) {
    const uint64_t mir = or_bitreversal(rep >> L, L) << L | or_bitreversal(rep & lm, L);
    for (r = 0; r < L; ++r) {
      mem[m++] = or_offset_state(rep >> L, L, r) << L | or_offset_state(rep & lm, L, r);
      mem[m++] = or_offset_state(mir >> L, L, r) << L | or_offset_state(mir & lm, L, r);
    }
    if (op->censym) {
      const int m0 = m;
      const uint64_t full = (2 * L >= 64) ? ~0ull : (1ull << (2 * L)) - 1;
      for (k = 0; k < m0; ++k) mem[m++] = ~mem[k] & full;
    }
  } else {
    uint64_t mir;
    if (op->kind == OR_ZBM) {
      mir = or_bitreversal(or_offset_state(rep, L, 1), L);
      for (r = 0; r < L; r += 2) {
        mem[m++] = or_offset_state(rep, L, r);
        mem[m++] = or_offset_state(mir, L, r);
      }
    } else if (op->bc & 1) {
      mem[m++] = rep;
      mem[m++] = or_bitreversal(rep, L);
    } else {
      mir = or_bitreversal(rep, L);
      for (r = 0; r < L; ++r) {
        mem[m++] = or_offset_state(rep, L, r);
        mem[m++] = or_offset_state(mir, L, r);
      }
    }
    if (op->censym) {
      const int m0 = m;
      for (k = 0; k < m0; ++k) mem[m++] = ~mem[k] & lm;
    }
  }
  qsort(mem, (size_t)m, sizeof(uint64_t), or_cmp_u64);
  for (k = 1, r = 1; k < m; ++k)
    if (mem[k] != mem[r - 1]) mem[r++] = mem[k];
  return m ? r : 0;
}

int or_sample_configs(const or_op *op, const double *vec, const double *vecl, const double *u,
                      int nchain, int length, uint64_t *states)
{
  const int L = op->L;
  const int64_t n = op->ncol;
  const uint64_t lm = (1ull << L) - 1;
  const double j0 = op->betaJ, j1 = -op->kappa * op->betaJ, j2 = op->h * op->betaJ;
  const int red = op->kind != OR_XAT && op->kind != OR_YAT;
  const int gen = op->kind == OR_YAT || op->kind == OR_ZBM || op->kind == OR_YAM ||
                  op->kind == OR_YBM;
  double *v, *vl = NULL, *cum, *e = NULL, *y = NULL;
  int64_t i, ui = 0;
  int c, j;
  if (!(gen || op->kind == OR_XAT || op->kind == OR_XAM2D || op->kind == OR_XDM)) return -1;
  if (gen && !vecl) return -1;
  v = (double *)malloc(sizeof(double) * n);
  cum = (double *)malloc(sizeof(double) * n);
  for (i = 0; i < n; ++i) v[i] = fabs(vec[i]);
  if (gen) {
    vl = (double *)malloc(sizeof(double) * n);
    for (i = 0; i < n; ++i) vl[i] = fabs(vecl[i]);
  }
  if (red) {
    e = (double *)calloc((size_t)n, sizeof(double));
    y = (double *)malloc(sizeof(double) * n);
  }
  for (c = 0; c < nchain; ++c) {
    uint64_t *st = states + (int64_t)c * length;
    double d = 0;
    int64_t qa;
    for (i = 0; i < n; ++i) {  /* initq */
      d += gen ? v[i] * vl[i] : v[i] * v[i];
      cum[i] = d;
    }
    qa = or_draw(cum, n, u[ui++]);
    st[0] = red ? op->reps[qa] : (uint64_t)qa;
    for (j = 1; j < length; ++j) {
      const uint64_t ra = st[j - 1];
      if (op->kind == OR_XAT) {  /* dumpconfig.cpp:110-129 */
        const double intr1 = j2 * (L - 2.0 * POP(ra))
                             + j0 * (L - 2.0 * POP(ra ^ or_offset_state(ra, L, 1)))
                             + j1 * (L - 2.0 * POP(ra ^ or_offset_state(ra, L, 2)));
        d = 0;
        for (i = 0; i < n; ++i) {
          const double intr2 = j0 * (L - 2.0 * POP(ra ^ (uint64_t)i));
          d += exp(intr1 + intr2) * v[i];
          cum[i] = d;
        }
        st[j] = (uint64_t)or_draw(cum, n, u[ui++]);
        continue;
      }
      if (op->kind == OR_YAT) {  /* dumpconfig.cpp:189-218: (b, a) -> (c, b) */
        const uint64_t a = ra & lm, b = ra >> L;
        const int64_t m = (int64_t)1 << L;
        const double intr1 = j2 * (L - 2.0 * POP(a))
                             + j0 * (L - 2.0 * POP(a ^ or_offset_state(a, L, 1)))
                             + j0 * (L - 2.0 * POP(a ^ b));
        d = 0;
        for (i = 0; i < m; ++i) {
          const double intr2 = j1 * (L - 2.0 * POP(a ^ (uint64_t)i));
          d += exp(intr1 + intr2) * v[(uint64_t)i << L | b];
          cum[i] = d;
        }
        st[j] = (uint64_t)or_draw(cum, m, u[ui++]) << L | b;
        continue;
      }
      /* the column T e_qa, then an orbit member */
      e[qa] = 1.0;
      or_op_apply(op, e, y, 0);
      e[qa] = 0.0;
      d = 0;
      for (i = 0; i < n; ++i) {
        d += y[i] * v[i];
        cum[i] = d;
      }
      qa = or_draw(cum, n, u[ui++]);
      {
        uint64_t mem[512];
        double qc[512];
        const int m = or_orbit_members(op, op->reps[qa], mem);
        int k, ok;
        d = 0;
        for (k = 0; k < m; ++k) {
          const double x = or_member_cross(op, ra, mem[k], &ok);
          if (ok) d += exp(x);
          qc[k] = d;
        }
        st[j] = mem[or_draw(qc, m, u[ui++])];
      }
    }
  }
  free(v);
  free(vl);
  free(cum);
  free(e);
  free(y);
  return 0;
}

int64_t or_op_ncol(const or_op *op) { return op->ncol; }
double or_op_normfactor(const or_op *op) { return op->nf; }
const double *or_op_tvec(const or_op *op) { return op->tvec; }
const uint64_t *or_op_reps(const or_op *op) { return op->reps; }
const uint16_t *or_op_dups(const or_op *op) { return op->dups; }

/* ------------------------------------------------------------------------ */
/* operator application                                                      */
/* ------------------------------------------------------------------------ */

/* In-place Kronecker butterfly on nbits low bits: TMat/tmat.cpp:80-93,
 * TMat/tmatreduce.cpp:546-563, TMat/tmat3d.cpp:323-340. */
static void butterfly_low(double *y, int64_t n, int nbits, const double bz[2])
{
  for (int it = 0; it < nbits; ++it) {
    uint64_t ms = 1ULL << it;
#pragma omp parallel for schedule(static)
    for (int64_t rp = 0; rp < n; ++rp) {
      if ((uint64_t)rp & ms) {
        double t = y[rp];
        y[rp] = bz[0] * t + bz[1] * y[rp ^ ms];
        y[rp ^ ms] = bz[1] * t + bz[0] * y[rp ^ ms];
      }
    }
  }
}

/* TMat/tmat.cpp:71-102 (XATMatVecProd2D::perform_op) */
static void apply_xat(const or_op *op, const double *x, double *y)
{
  int64_t n = op->nstate;
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < n; ++rp) y[rp] = op->tvec[rp] * x[rp];
  butterfly_low(y, n, op->L, op->bzwj1);
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < n; ++rp) y[rp] = op->nf * op->tvec[rp] * y[rp];
}

/* Full 3D: same structure as XAT with sqrt(nf) folded into tvec. */
static void apply_xat3d(const or_op *op, const double *x, double *y)
{
  int64_t n = op->nstate;
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < n; ++rp) y[rp] = op->tvec[rp] * x[rp];
  butterfly_low(y, n, op->L, op->bzwj1);
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < n; ++rp) y[rp] = op->tvec[rp] * y[rp];
}

/* TMat/tmat.cpp:344-359 (XDT permute_mult_T1) and :361-376 (_last) */
static void xdt_step(const or_op *op, const double *xi, double *yo, int last, uint64_t lastbit)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate, nb = N << 1, mask = N - 1, maskb = N - 2;
  const double *b1 = op->bzwj1, *b2 = op->bzwj2;
#pragma omp parallel for schedule(static)
  for (int64_t irpb = 0; irpb < (int64_t)nb; ++irpb) {
    uint64_t rpb = (uint64_t)irpb;
    uint64_t auxp = rpb >> L;
    uint64_t rp = rpb & mask;
    uint64_t rpp = (auxp << L) | or_offset_state(rp, L, 1);
    uint64_t rq = (rp & maskb) | auxp;
    uint64_t sbit = rp & 1;
    uint64_t nb2 = last ? lastbit : ((rq >> 1) & 1);
    yo[rpp] = b1[sbit ^ auxp] * b2[nb2 ^ sbit] * (b2[sbit] * xi[rq] + b2[sbit ^ 1] * xi[rq | N]);
  }
}

/* TMat/tmat.cpp:379-476 (XDTMatVecProd2D::perform_op) */
static void apply_xdt(const or_op *op, const double *x, double *y)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate, nc = N >> 1;
  double *t1 = (double *)malloc(sizeof(double) * 2 * N);
  double *t2 = (double *)malloc(sizeof(double) * 2 * N);
  const double *tv = op->tvec;
  if (op->bc == 0) {
    for (uint64_t lastbit = 0; lastbit < 2; ++lastbit) {
      double *p1 = t1, *p2 = t2, *sw;
#pragma omp parallel for schedule(static)
      for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
        uint64_t rp = (uint64_t)irp;
        uint64_t rpb = ((rp & nc) << 1) | rp;
        p1[rpb] = ((rp & 1) == lastbit) ? tv[rp] * x[rp] : 0.0;
        p1[rpb ^ N] = 0.0;
      }
      for (int it = 0; it < L - 1; ++it) {
        xdt_step(op, p1, p2, 0, 0);
        sw = p1; p1 = p2; p2 = sw;
      }
      xdt_step(op, p1, p2, 1, lastbit);
#pragma omp parallel for schedule(static)
      for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
        uint64_t rp = (uint64_t)irp;
        double v = op->nf * tv[rp] * (p2[rp] + p2[rp ^ N]);
        if (lastbit == 0) y[rp] = v; else y[rp] += v;
      }
    }
  } else {
    double *p1 = t1, *p2 = t2, *sw;
    const double *b2 = op->bzwj2;
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp;
      if (op->bc == 5) {
        p1[rp] = tv[rp] * x[rp] * b2[rp & 1] * b2[rp >> (L - 1)];
        p1[rp | N] = 0.0;
      } else {
        p1[rp | N] = tv[rp] * x[rp] * b2[(rp & 1) ^ 1] * b2[rp >> (L - 1)];
        p1[rp] = 0.0;
      }
    }
    for (int it = 0; it < L - 1; ++it) {
      xdt_step(op, p1, p2, 0, 0);
      sw = p1; p1 = p2; p2 = sw;
    }
    xdt_step(op, p1, p2, 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp;
      y[rp] = op->nf * tv[rp] * (p2[rp] + p2[rp | N]);
    }
  }
  free(t1);
  free(t2);
}

/* TMat/tmat.cpp:521-553 (ZBT permute_mult_T1) and :556-588 (_last) */
static void zbt_step(const or_op *op, const double *xi, double *yo, int last, uint64_t lastbit, int dir)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate, nb = N << 2, mask = N - 1, maskb = N - 4;
  const double *b1 = op->bzwj1, *b2 = op->bzwj2, *b3 = op->bzwj3;
#pragma omp parallel for schedule(static)
  for (int64_t irpb = 0; irpb < (int64_t)nb; ++irpb) {
    uint64_t rpb = (uint64_t)irpb;
    uint64_t auxp = rpb >> L;
    uint64_t rp = rpb & mask;
    uint64_t rpp = (auxp << L) | or_offset_state(rp, L, 2);
    uint64_t rq = (rp & maskb) | auxp;
    uint64_t sl = (rp >> 1) & 1, sr = rp & 1, pl = auxp >> 1, pr = auxp & 1;
    uint64_t fl = last ? (lastbit >> 1) : ((rp >> 3) & 1);
    uint64_t fr = last ? (lastbit & 1) : ((rp >> 2) & 1);
    double w;
    if (!dir)
      w = b1[sr ^ pl] * b2[sr ^ fr] * b2[sl ^ fl] * b3[sr ^ pr] * b3[sl ^ pl];
    else
      w = b1[sl ^ fr] * b1[sl ^ pr] * b2[sr ^ fr] * b2[sl ^ fl] * b3[sr ^ pr] * b3[sl ^ pl];
    double acc = 0.0;
    for (uint64_t aq = 0; aq < 4; ++aq) {
      uint64_t al = aq >> 1, ar = aq & 1;
      uint64_t rqq = (aq << L) | rq;
      if (!dir) acc += xi[rqq] * b1[sr ^ al] * b2[sr ^ ar] * b2[sl ^ al];
      else acc += xi[rqq] * b2[sr ^ ar] * b2[sl ^ al];
    }
    w *= acc;
    yo[rpp] = w;
  }
}

/* TMat/tmat.cpp:596-649 (ZBTMatVecProd2D::perform_op_internal) */
static void apply_zbt(const or_op *op, const double *x, double *y, int dir)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate, nc = (N >> 1) | (N >> 2);
  double *t1 = (double *)malloc(sizeof(double) * 4 * N);
  double *t2 = (double *)malloc(sizeof(double) * 4 * N);
  const double *tv = op->tvec;
  for (uint64_t lastbit = 0; lastbit < 4; ++lastbit) {
    double *p1 = t1, *p2 = t2, *sw;
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp;
      uint64_t rpb = ((rp & nc) << 2) | rp;
      p1[rpb] = ((rp & 3) == lastbit) ? tv[rp] * x[rp] : 0.0;
      p1[rpb ^ N] = 0.0;
      p1[rpb ^ (N << 1)] = 0.0;
      p1[rpb ^ (nc << 2)] = 0.0;
    }
    for (int it = 0; it < L / 2 - 1; ++it) {
      zbt_step(op, p1, p2, 0, 0, dir);
      sw = p1; p1 = p2; p2 = sw;
    }
    zbt_step(op, p1, p2, 1, lastbit, dir);
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp;
      double v = tv[rp] * (p2[rp] + p2[rp ^ N] + p2[rp ^ (N << 1)] + p2[rp ^ (nc << 2)]);
      if (lastbit == 0) y[rp] = v; else y[rp] += v;
    }
  }
  free(t1);
  free(t2);
}

/* TMat/tmatreduce.cpp:406-529 (XAMMatVecProd2D::perform_op) */
static void apply_xam2d(const or_op *op, const double *x, double *y)
{
  int L = op->L, nlb = op->nleadbit;
  uint64_t N = (uint64_t)op->nstate;
  uint64_t nloop = 1ULL << nlb, vecsize = N >> nlb;
  uint64_t maskb = vecsize - 1, maskc = nloop - 1, maskbinv = ~maskb & (N - 1);
  int nsub = L - nlb;
  int64_t m = op->ncol;
  double *yt = (double *)malloc(sizeof(double) * vecsize);
  const uint64_t *idx = op->reps;
  const double *tv = op->tvec;
  int bc = op->bc;
  if (op->censym) {
    nloop >>= 1;
    for (uint64_t rdup = 0; rdup < nloop; ++rdup) {
      uint64_t rh = rdup << nsub, rr = or_bitreversal(rdup, nlb);
#pragma omp parallel for schedule(dynamic, 4096)
      for (int64_t rp = 0; rp < m; ++rp) {
        double d = tv[rp] * x[rp];
        uint64_t now = idx[rp], nx = now;
        do {
          if ((nx & maskbinv) == rh) yt[nx & maskb] = d;
          else if ((~nx & maskbinv) == rh) yt[~nx & maskb] = d;
          if ((nx & maskc) == rr) yt[or_bitreversal(nx, L) & maskb] = d;
          else if ((~nx & maskc) == rr) yt[or_bitreversal(~nx, L) & maskb] = d;
          if (bc) break;
          nx = (nx >> 1) | ((nx & 1) << (L - 1));
        } while (nx != now);
      }
      butterfly_low(yt, (int64_t)vecsize, nsub, op->bzwj1);
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t now = idx[rp];
        double d = or_powdi(op->bzwj1, or_spinbzexp(now >> nsub, rdup, nlb));
        double v = d * yt[now & maskb] + yt[~now & maskb] / d;
        if (rdup) y[rp] += v; else y[rp] = v;
      }
    }
  } else {
    for (uint64_t rdup = 0; rdup < nloop; ++rdup) {
      uint64_t rh = rdup << nsub, rr = or_bitreversal(rdup, nlb);
#pragma omp parallel for schedule(dynamic, 4096)
      for (int64_t rp = 0; rp < m; ++rp) {
        double d = tv[rp] * x[rp];
        uint64_t now = idx[rp], nx = now;
        do {
          if ((nx & maskbinv) == rh) yt[nx & maskb] = d;
          if ((nx & maskc) == rr) yt[or_bitreversal(nx, L) & maskb] = d;
          if (bc) break;
          nx = (nx >> 1) | ((nx & 1) << (L - 1));
        } while (nx != now);
      }
      butterfly_low(yt, (int64_t)vecsize, nsub, op->bzwj1);
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t now = idx[rp];
        double d = or_powdi(op->bzwj1, or_spinbzexp(now >> nsub, rdup, nlb));
        double v = d * yt[now & maskb];
        if (rdup) y[rp] += v; else y[rp] = v;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < m; ++rp) y[rp] *= op->dups[rp] * tv[rp];
  free(yt);
}

/* TMat/tmat3d.cpp:203-321 (XAMMatVecProd3D::perform_op) */
static void apply_xam3d(const or_op *op, const double *x, double *y)
{
  int s1 = op->size1, s2 = op->size2, size = op->L, nlb = op->nleadbit;
  uint64_t N = (uint64_t)op->nstate, mask = N - 1;
  uint64_t nloop = 1ULL << nlb, vecsize = N >> nlb;
  uint64_t maskb = vecsize - 1, maskbinv = ~maskb & (N - 1);
  int nsub = size - nlb;
  int64_t m = op->ncol;
  double *yt = (double *)malloc(sizeof(double) * vecsize);
  if (op->censym) nloop >>= 1;
  int ncen = op->censym ? 2 : 1, nrot = op->rotate ? 2 : 1;
  for (uint64_t rdup = 0; rdup < nloop; ++rdup) {
    uint64_t rh = rdup << nsub;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t rp = 0; rp < m; ++rp) {
      double d = op->tvec[rp] * x[rp];
      uint64_t nr = op->reps[rp];
      for (int ir = 0; ir < nrot; ++ir) {
        uint64_t nc = nr;
        for (int ic = 0; ic < ncen; ++ic) {
          uint64_t nm1 = nc;
          for (int i1 = 0; i1 < 2; ++i1) {
            uint64_t nm2 = nm1;
            for (int i2 = 0; i2 < 2; ++i2) {
              uint64_t o1 = nm2;
              do {
                uint64_t o2 = o1;
                do {
                  if ((o2 & maskbinv) == rh) yt[o2 & maskb] = d;
                  o2 = or_dffset3d(o2, s1, s2, 1);
                } while (o2 != o1);
                o1 = or_rffset3d(o1, s1, s2, 1);
              } while (o1 != nm2);
              nm2 = or_reflect3d(nm2, s1, s2, 1);
            }
            nm1 = or_reflect3d(nm1, s1, s2, 0);
          }
          nc = ~nc & mask;
        }
        nr = or_rotate3d(nr, s1);
      }
    }
    butterfly_low(yt, (int64_t)vecsize, nsub, op->bzwj1);
#pragma omp parallel for schedule(static)
    for (int64_t rp = 0; rp < m; ++rp) {
      uint64_t now = op->reps[rp];
      double d = or_powdi(op->bzwj1, or_spinbzexp(now >> nsub, rdup, nlb));
      double v;
      if (op->censym) v = d * yt[now & maskb] + yt[~now & maskb] / d;
      else v = d * yt[now & maskb];
      if (rdup) y[rp] += v; else y[rp] = v;
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < m; ++rp) y[rp] *= op->dups[rp] * op->tvec[rp];
  free(yt);
}

/* TMat/tmatreduce.cpp:1230-1259 (XDMMatVecProd2D::permute_mult_TL): the
 * sub-lead bits one at a time, the auxiliary bit (index | 2^nsub) holding the
 * previous old-layer spin; outbit is the new-layer spin past the last one. */
static void xdm_permute_mult_tl(const or_op *op, double *yt, int outbit)
{
  uint64_t pn = (uint64_t)op->nstate >> op->nleadbit;
  int nloop = op->L - op->nleadbit;
  const double *b1 = op->bzwj1, *b2 = op->bzwj2;
  for (int it = 0; it < nloop; ++it) {
    uint64_t ms = 1ULL << it, bshift = ~ms & ((uint64_t)op->nstate - 1);
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)pn; ++irp) {
      uint64_t rp = (uint64_t)irp;
      if (!(rp & ms)) continue;
      int nb = outbit;
      if (it < nloop - 1) nb = (int)((rp >> (it + 1)) & 1);
      uint64_t i00 = rp & bshift, i01 = rp, i10 = (rp & bshift) | pn, i11 = rp | pn;
      double d00 = yt[i00], d01 = yt[i01], d10 = yt[i10], d11 = yt[i11];
      yt[i00] = b1[0] * b2[nb] * (b2[0] * d00 + b2[1] * d10);
      yt[i01] = b1[1] * b2[nb ^ 1] * (b2[1] * d00 + b2[0] * d10);
      yt[i10] = b1[1] * b2[nb] * (b2[0] * d01 + b2[1] * d11);
      yt[i11] = b1[0] * b2[nb ^ 1] * (b2[1] * d01 + b2[0] * d11);
    }
  }
}

/* TMat/tmatreduce.cpp:1054-1228 (XDMMatVecProd2D::perform_op), periodic */
static void apply_xdm(const or_op *op, const double *x, double *y)
{
  int L = op->L, nlb = op->nleadbit;
  uint64_t N = (uint64_t)op->nstate;
  uint64_t nloop = 1ULL << nlb, vecsize = N >> nlb;
  uint64_t mask = N - 1, maskb = vecsize - 1, maskc = nloop - 1, maskd = maskc >> 1;
  uint64_t maskbinv = ~maskb & (N - 1);
  int nsub = L - nlb;
  uint64_t nsubstate = 1ULL << nsub;
  int64_t m = op->ncol;
  double *xt = (double *)malloc(sizeof(double) * vecsize);
  double *yt = (double *)malloc(sizeof(double) * 2 * vecsize);
  const uint64_t *idx = op->reps;
  const double *tv = op->tvec, *b1 = op->bzwj1, *b2 = op->bzwj2;
  for (int64_t rp = 0; rp < m; ++rp) y[rp] = 0.0;
  int cen = op->censym;
  if (cen) nloop >>= 1;
  for (uint64_t rdup = 0; rdup < nloop; ++rdup) {
    uint64_t rh = rdup << nsub, rr = or_bitreversal(rdup, nlb);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t rp = 0; rp < m; ++rp) {
      double d = tv[rp] * x[rp];
      uint64_t now = idx[rp], nx = now;
      do {
        if (cen) {
          if ((nx & maskbinv) == rh) xt[nx & maskb] = d;
          else if ((~nx & maskbinv) == rh) xt[~nx & maskb] = d;
          if ((nx & maskc) == rr) xt[or_bitreversal(nx, L) & maskb] = d;
          else if ((~nx & maskc) == rr) xt[or_bitreversal(~nx, L) & maskb] = d;
        } else {
          if ((nx & maskbinv) == rh) xt[nx & maskb] = d;
          if ((nx & maskc) == rr) xt[or_bitreversal(nx, L) & maskb] = d;
        }
        nx = (nx >> 1) | ((nx & 1) << (L - 1));
      } while (nx != now);
    }
    /* fbit: highest bit of the new layer */
    for (int fbit = 0; fbit < 2; ++fbit) {
      /* the old layer's highest bit (the auxiliary bit); always 0 when folded */
      uint64_t gbit = cen ? 0 : (rdup >> (nlb - 1)) << nsub;
#pragma omp parallel for schedule(static)
      for (int64_t irp = 0; irp < (int64_t)vecsize; ++irp) {
        uint64_t rp = (uint64_t)irp;
        yt[rp | gbit] = xt[rp] * b2[(rp & 1) ^ (uint64_t)fbit];
        yt[rp | (gbit ^ nsubstate)] = 0.0;
      }
      xdm_permute_mult_tl(op, yt, (int)(rdup & 1));
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t now = idx[rp];
        if (cen) {
          if (fbit) now = ~now & mask;
        } else if ((now >> (L - 1)) ^ (uint64_t)fbit) {
          continue;
        }
        uint64_t rdupb = now >> nsub;
        double d = or_powdi(b1, or_spinbzexp(rdup, rdupb, nlb));
        d *= or_powdi(b2, (2 * POP(((rdup >> 1) ^ rdupb) & maskd) - (nlb - 1)) +
                              (2 * POP((rdup ^ (rdupb >> 1)) & maskd) - (nlb - 1)));
        y[rp] += d * (yt[now & maskb] * b2[rdupb & 1] + yt[(now & maskb) | vecsize] * b2[(rdupb & 1) ^ 1]);
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < m; ++rp) y[rp] *= op->dups[rp] * tv[rp];
  free(xt);
  free(yt);
}

/* TMat/tmatreduce.cpp:1615-1652 (ZBMMatVecProd2D::permute_mult_T1) and, with
 * last >= 0, :1654-1691 (permute_mult_T1_last): one two-site step over the
 * sub-lead bits, two auxiliary bits per entry.  dir 1 = transpose. */
static void zbm_step(const or_op *op, const double *xin, double *yout, int last, int dir)
{
  int nsub = op->L - op->nleadbit;
  uint64_t vecsize = (uint64_t)op->nstate >> op->nleadbit;
  uint64_t nstateb = vecsize << 2, maskb = vecsize - 4;
  const double *b1 = op->bzwj1, *b2b = op->bzwj2b, *b2c = op->bzwj2c;
#pragma omp parallel for schedule(static)
  for (int64_t irpb = 0; irpb < (int64_t)nstateb; ++irpb) {
    uint64_t rpb = (uint64_t)irpb;
    uint64_t auxp = rpb & 3, rp = rpb >> 2;
    uint64_t rpp = (or_offset_state(rp, nsub, 2) << 2) | auxp;
    uint64_t rq = ((rp & maskb) | auxp) << 2;
    unsigned sbit = (unsigned)(rp & 3);
    unsigned hi = last >= 0 ? ((unsigned)last << 2) : (unsigned)(rp & 0xc);
    unsigned j2 = (hi | (unsigned)auxp) ^ (sbit << 2 | sbit);
    double v, d;
    if (!dir) {
      unsigned srbit = (unsigned)(rp & 1), plbit = (unsigned)(auxp >> 1);
      v = b1[srbit ^ plbit] * b2c[j2];
      d = xin[rq] * b1[srbit] * b2b[sbit];
      d += xin[rq + 1] * b1[srbit] * b2b[sbit ^ 1];
      d += xin[rq + 2] * b1[srbit ^ 1] * b2b[sbit ^ 2];
      d += xin[rq + 3] * b1[srbit ^ 1] * b2b[sbit ^ 3];
    } else {
      unsigned slbit = (unsigned)((rp >> 1) & 1), prbit = (unsigned)(auxp & 1);
      unsigned frbit = last >= 0 ? ((unsigned)last & 1) : (unsigned)((rp >> 2) & 1);
      v = b1[slbit ^ frbit] * b1[slbit ^ prbit] * b2c[j2];
      d = xin[rq] * b2b[sbit];
      d += xin[rq + 1] * b2b[sbit ^ 1];
      d += xin[rq + 2] * b2b[sbit ^ 2];
      d += xin[rq + 3] * b2b[sbit ^ 3];
    }
    yout[rpp] = v * d;
  }
}

/* TMat/tmatreduce.cpp:1693-1746 (ZBMMatVecProd2D::leadingbz): the lead bits'
 * interactions of state now with the transferred sub-lead part yt. */
static double zbm_leadingbz(const or_op *op, const double *yt, uint64_t now, uint64_t lead,
                            uint64_t lastbit, int dir)
{
  const uint64_t EVEN = 0x5555555555555555ULL;
  int nlb = op->nleadbit, nsub = op->L - nlb;
  uint64_t vecsize = (uint64_t)op->nstate >> nlb;
  uint64_t maskb = vecsize - 1, maskc = (1ULL << nlb) - 1, maskd = maskc >> 2;
  uint64_t nl = now >> nsub, nsh = (now & maskb) << 2;
  const double *b1 = op->bzwj1, *b2 = op->bzwj2, *b2b = op->bzwj2b;
  double df1 = or_powdi(op->bzwj3, or_spinbzexp(nl, lead, nlb));
  df1 *= or_powdi(b2, or_spinbzexp(nl, (lead >> 2) | (lastbit << (nlb - 2)), nlb) +
                          or_spinbzexp(nl >> 2, lead & maskd, nlb - 2));
  uint64_t inow, inxt;
  int iint;
  double df2;
  unsigned q = (unsigned)(nl & 3);
  if (!dir) {
    inow = nl & EVEN;
    inow |= inow >> 1;
    inxt = (lead >> 1) & EVEN;
    inxt |= (inxt << 1) & (maskc >> 1);
    iint = or_spinbzexp(inow, inxt, nlb - 1);
    unsigned lb = (unsigned)(nl & 1);
    df2 = b1[lb] * (b2b[q] * yt[nsh] + b2b[q ^ 1] * yt[nsh + 1]) +
          b1[lb ^ 1] * (b2b[q ^ 2] * yt[nsh + 2] + b2b[q ^ 3] * yt[nsh + 3]);
  } else {
    inow = (nl >> 1) & EVEN;
    inow |= inow << 1;
    inxt = (lead | (lastbit << nlb)) & EVEN;
    inxt = (inxt & maskc) | (inxt >> 1);
    iint = or_spinbzexp(inow, inxt, nlb);
    df2 = b2b[q] * yt[nsh] + b2b[q ^ 1] * yt[nsh + 1] + b2b[q ^ 2] * yt[nsh + 2] +
          b2b[q ^ 3] * yt[nsh + 3];
  }
  df1 *= or_powdi(b1, iint);
  return df1 * df2;
}

/* TMat/tmatreduce.cpp:1460-1612 (ZBMMatVecProd2D::perform_op_internal) */
static void apply_zbm(const or_op *op, const double *x, double *y, int dir)
{
  int L = op->L, nlb = op->nleadbit;
  uint64_t N = (uint64_t)op->nstate;
  uint64_t nloop = 1ULL << nlb, vecsize = N >> nlb;
  uint64_t mask = N - 1, maskb = vecsize - 1, maskc = nloop - 1;
  uint64_t maskbinv = ~maskb & (N - 1);
  int nsub = L - nlb;
  int64_t m = op->ncol;
  double *xt = (double *)malloc(sizeof(double) * vecsize);
  double *p1 = (double *)malloc(sizeof(double) * 4 * vecsize);
  double *p2 = (double *)malloc(sizeof(double) * 4 * vecsize);
  const uint64_t *idx = op->reps;
  const double *tv = op->tvec;
  for (int64_t rp = 0; rp < m; ++rp) y[rp] = 0.0;
  int cen = op->censym;
  if (cen) nloop >>= 1;
  for (uint64_t rdup = 0; rdup < nloop; ++rdup) {
    uint64_t rh = rdup << nsub, rhinv = (~rdup & maskc) << nsub;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t rp = 0; rp < m; ++rp) {
      double d = tv[rp] * x[rp];
      uint64_t now = idx[rp], nx = now;
      do {
        uint64_t ni = or_bitreversal((nx >> 1) | ((nx & 1) << (L - 1)), L);
        if (cen) {
          if ((nx & maskbinv) == rh) xt[nx & maskb] = d;
          else if ((nx & maskbinv) == rhinv) xt[~nx & maskb] = d;
          if ((ni & maskbinv) == rh) xt[ni & maskb] = d;
          else if ((ni & maskbinv) == rhinv) xt[~ni & maskb] = d;
        } else {
          if ((nx & maskbinv) == rh) xt[nx & maskb] = d;
          if ((ni & maskbinv) == rh) xt[ni & maskb] = d;
        }
        nx = (nx >> 2) | ((nx & 3) << (L - 2));
      } while (nx != now);
    }
    uint64_t fbit = rdup >> (nlb - 2);
    for (uint64_t lastbit = 0; lastbit < 4; ++lastbit) {
#pragma omp parallel for schedule(static)
      for (int64_t irp = 0; irp < (int64_t)vecsize; ++irp) {
        uint64_t rp = (uint64_t)irp, rrp = (rp << 2) | fbit;
        p1[rrp] = ((rp & 3) == lastbit) ? xt[rp] : 0.0;
        p1[rrp ^ 1] = p1[rrp ^ 2] = p1[rrp ^ 3] = 0.0;
      }
      double *a = p1, *b = p2, *t;
      for (int r = 0; r < nsub / 2 - 1; ++r) {
        zbm_step(op, a, b, -1, dir);
        t = a; a = b; b = t;
      }
      zbm_step(op, a, b, (int)(rdup & 3), dir);
#pragma omp parallel for schedule(static)
      for (int64_t rp = 0; rp < m; ++rp) {
        uint64_t now = idx[rp];
        double d = zbm_leadingbz(op, b, now, rdup, lastbit, dir);
        if (cen) d += zbm_leadingbz(op, b, ~now & mask, rdup, lastbit, dir);
        y[rp] += d;
      }
      /* the reference swaps its own buffer pointers; keep p1 the input side */
      if (b == p1) { p1 = p2; p2 = b; }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < m; ++rp) y[rp] *= op->dups[rp] * tv[rp];
  free(xt);
  free(p1);
  free(p2);
}

/* TMat/tmat.cpp:207-242 (YATMatVecProd2D::permute_mult_T1, perform_op) and,
 * for the transpose, :126-183 (YATMatVecProd2D_T).  A state is two layers,
 * index (s1 << L) | s2.  The reference's final layer swap reads the last
 * step's buffer, which is y itself after an odd number of steps (odd L): the
 * swap then overwrites entries it still has to read.  Here the steps always
 * end in a scratch buffer, so the swap reads intact data for every L. */
static void apply_yat(const or_op *op, const double *x, double *y, int dir)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate, lm = (1ULL << L) - 1, hb = 1ULL << L;
  const double *b2 = op->bzwj2, *tv = op->tvec;
  double *t1 = (double *)malloc(sizeof(double) * N);
  double *t2 = (double *)malloc(sizeof(double) * N);
  double *p1 = t1, *p2 = t2, *sw;
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < (int64_t)N; ++rp) p1[rp] = dir ? tv[rp] * x[rp] : x[rp];
  for (int r = 0; r < L; ++r) {
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp;
      if (!dir) {
        /* Kronecker step on the high layer, then rotate it right by one */
        uint64_t s1 = rp >> L, s2 = rp & lm;
        uint64_t rq = (or_offset_state(s1, L, 1) << L) | s2;
        p2[rq] = b2[0] * p1[rp] + b2[1] * p1[rp ^ hb];
      } else {
        /* transposed step on the low layer (rotate left, flip the last bit) */
        uint64_t rq = rp, s1 = rq & lm, s2 = rq >> L;
        uint64_t src = (s2 << L) | or_lffset_state(s1, L, 1);
        uint64_t srcf = (s2 << L) | or_lffset_state(s1 ^ 1ULL, L, 1);
        p2[rq] = b2[0] * p1[src] + b2[1] * p1[srcf];
      }
    }
    sw = p1; p1 = p2; p2 = sw;
  }
#pragma omp parallel for schedule(static)
  for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
    uint64_t rp = (uint64_t)irp, s1 = rp >> L, s2 = rp & lm, rq = (s2 << L) | s1;
    if (!dir) y[rp] = op->nf * tv[rp] * p1[rq];
    else y[rq] = op->nf * p1[rp];
  }
  free(t1);
  free(t2);
}

/* TMat/tmat.cpp:265-302 (Y2ATMatVecProd2D::permute_mult_T1, perform_op) */
static void apply_y2at(const or_op *op, const double *x, double *y)
{
  int L = op->L;
  uint64_t N = (uint64_t)op->nstate;
  uint64_t highbit = 1ULL << L, bothbit = highbit | 1, mask = highbit - 1;
  const double *b1 = op->bzwj1, *b2 = op->bzwj2, *tv = op->tvec;
  double *t1 = (double *)malloc(sizeof(double) * N);
  double *t2 = (double *)malloc(sizeof(double) * N);
  double *p1 = t1, *p2 = t2, *sw;
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < (int64_t)N; ++rp) p1[rp] = tv[rp] * x[rp];
  for (int r = 0; r < L; ++r) {
#pragma omp parallel for schedule(static)
    for (int64_t irp = 0; irp < (int64_t)N; ++irp) {
      uint64_t rp = (uint64_t)irp, lf = rp & mask, rt = rp >> L;
      uint64_t rq = (or_offset_state(rt, L, 1) << L) | or_offset_state(lf, L, 1);
      unsigned jb = (unsigned)((lf ^ rt) & 1);
      p2[rq] = b1[jb] * (b2[0] * p1[rp] + p1[rp ^ highbit]) +
               b1[jb ^ 1] * (b2[1] * p1[rp ^ bothbit] + p1[rp ^ 1]);
    }
    sw = p1; p1 = p2; p2 = sw;
  }
#pragma omp parallel for schedule(static)
  for (int64_t rp = 0; rp < (int64_t)N; ++rp) y[rp] = op->nf * tv[rp] * p1[rp];
  free(t1);
  free(t2);
}

/* TMat/tmatreduce.cpp:811-873 (YAMMatVecProd2D::permute_mult_TL, perform_op)
 * and, dir = 1, :957-1015 (YAMMatVecProd2D_T::perform_op): per group of
 * orbits sharing a high layer, scatter Tvec x over the low layer of every
 * image with that high layer, transfer the low layer by (x)K2, and gather at
 * the layer-swapped orbit (dir 0) or at the orbit itself from the swapped
 * input (dir 1). */
static void apply_yam(const or_op *op, const double *x, double *y, int dir)
{
  int L = op->L;
  uint64_t ls = 1ULL << L, lm = ls - 1;
  int64_t m = op->ncol, rp, rq, leadrp = 0;
  const double *tv = op->tvec, *b2 = op->bzwj2;
  double *yt = (double *)malloc(sizeof(double) * ls);
  for (rp = 0; rp < m; rp = rq, ++leadrp) {
    uint64_t lf = op->reps[rp] >> L;
    const unsigned char *ld = op->lead[leadrp];
    for (rq = rp; rq < m && (op->reps[rq] >> L) == lf; ++rq) {
      uint64_t rt = op->reps[rq] & lm;
      double d;
      if (!dir) {
        d = tv[rq] * x[rq];
      } else {
        int64_t sw = op->swp[rq];
        d = tv[sw] * x[sw];
      }
      for (int ch = 0; ch < 4; ++ch) {
        unsigned char sh = ld[ch];
        uint64_t rn = rt;
        if (ch && sh == (unsigned char)L) continue;
        switch (ch) {
          case 1: rn = or_bitreversal(rt, L); break;
          case 2: rn = ~rt & lm; break;
          case 3: rn = or_bitreversal(~rt & lm, L); break;
          default: break;
        }
        if (ch) rn = or_offset_state(rn, L, sh);
        for (int sa = 0; sa < L; sa += ld[0]) {
          yt[rn] = d;
          rn = or_offset_state(rn, L, ld[0]);
        }
      }
    }
    for (int it = 0; it < L; ++it) {
      uint64_t ms = 1ULL << it;
      for (uint64_t r = 0; r < ls; ++r) {
        if (!(r & ms)) continue;
        double t = yt[r];
        yt[r] = b2[0] * t + b2[1] * yt[r ^ ms];
        yt[r ^ ms] = b2[1] * t + b2[0] * yt[r ^ ms];
      }
    }
    for (int64_t rr = rp; rr < rq; ++rr) {
      uint64_t rt = op->reps[rr] & lm;
      int64_t o = dir ? rr : op->swp[rr];
      y[o] = tv[o] * op->dups[o] * yt[rt];
    }
  }
  free(yt);
}

void or_op_apply(const or_op *op, const double *x, double *y, int transpose)
{
  switch (op->kind) {
    case OR_XAT: apply_xat(op, x, y); break;
    case OR_XAT3D: apply_xat3d(op, x, y); break;
    case OR_XDT: apply_xdt(op, x, y); break;
    case OR_ZBT: apply_zbt(op, x, y, transpose ? 1 : 0); break;
    case OR_XAM2D: apply_xam2d(op, x, y); break;
    case OR_XAM3D: apply_xam3d(op, x, y); break;
    case OR_XDM: apply_xdm(op, x, y); break;
    case OR_ZBM: apply_zbm(op, x, y, transpose ? 1 : 0); break;
    case OR_YAT: apply_yat(op, x, y, transpose ? 1 : 0); break;
    case OR_Y2AT: apply_y2at(op, x, y); break;  /* tmat.hpp:382-384: T^T := T */
    case OR_YAM:
    case OR_YBM: apply_yam(op, x, y, transpose ? 1 : 0); break;
    default: break;
  }
}
