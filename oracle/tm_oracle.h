/*
 * tm_oracle.h -- CPU restatement of the LattTMat transfer-matrix operators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the
 * HIP product path is compared.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library
 * (latttmat_amd/liblatttmat.so) never links, loads or calls it.
 *
 * Each function restates the reference algorithm it names (file:line in the
 * reference tree, paths relative to its root).  Plain C99 + OpenMP, FP64.
 *
 * Parity pinning (what checks each function independently of this file):
 *   bit primitives            exhaustive textbook checks (tests/test_oracle_bits.py)
 *   or_tvec_* / XAT apply     dense T from the Hamiltonian, Kaufman closed form
 *   XDT apply                 dense T from the DNNI Hamiltonian (L<=10), Kaufman at kappa=0
 *   ZBT apply / transpose     adjoint identity <y,Tx>=<T^T y,x>, dense T of the
 *                             BNNNI/3NN diagonal-TM Hamiltonian, Perron positivity
 *   orbit enumeration         brute-force closure, Burnside counts, min-rep invariant
 *   XAM 2D / 3D apply         projection of the dense full T onto orbit sums;
 *                             lambda0(reduced) == lambda0(full)
 *   YAT / Y2AT apply          dense T of the two-layer transfer built from the
 *                             axial ANNNI Hamiltonian (tests/physics.py)
 *   YAM / YBM apply (and T)   projection of the dense symmetrised two-layer
 *                             Hamiltonian matrix onto layer-wise orbit sums;
 *                             lambda0(YAM) == lambda0(YAT)
 *   XDM / ZBM apply           projection of the dense full XDT / ZBT onto orbit
 *                             sums (ZBM transposed too); lambda0(reduced) ==
 *                             lambda0(full); group generators commute with T
 *   full 3D operator          dense T from the 3D Hamiltonian (2x3, 3x3 layers)
 */
#ifndef TM_ORACLE_H
#define TM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- bit-state primitives: TMat/utility.hpp:36-128, TMat/utility.cpp:17-120 ---- */
uint64_t or_offset_state(uint64_t s, int n, int d);   /* rotate right by d, utility.hpp:91-94 */
uint64_t or_lffset_state(uint64_t s, int n, int d);   /* rotate left by d,  utility.hpp:99-102 */
uint64_t or_bitreversal(uint64_t s, int n);           /* utility.cpp:17-31 */
int      or_spinbzexp(uint64_t a, uint64_t b, int n); /* 2*popcount(a^b)-n, utility.hpp:83-86 */
double   or_powdi(const double a[2], int b);          /* utility.hpp:36-55 */
uint64_t or_rffset3d(uint64_t s, int s1, int s2, int d);   /* utility.cpp:52-63 */
uint64_t or_dffset3d(uint64_t s, int s1, int s2, int d);   /* utility.cpp:68-73 */
uint64_t or_rotate3d(uint64_t s, int n);                    /* utility.cpp:78-91 */
uint64_t or_reflect3d(uint64_t s, int s1, int s2, int which);/* utility.cpp:97-120 */

/* ---- operator kinds ---- */
enum {
  OR_XAT = 1,      /* TMat/tmat.cpp:39-102     ANNNI perpendicular TM, full 2^L      */
  OR_XDT = 2,      /* TMat/tmat.cpp:322-476    DNNI perpendicular TM, full 2^L       */
  OR_ZBT = 3,      /* TMat/tmat.cpp:480-662    BNNNI/3NN diagonal TM, full 2^L       */
  OR_XAM2D = 4,    /* TMat/tmatreduce.cpp:338-563  ANNNI reduced by D_L (x flip)     */
  OR_XAM3D = 5,    /* TMat/tmat3d.cpp:102-340  3D ANNNI reduced                       */
  OR_XAT3D = 6,    /* full 2^(s1*s2) 3D ANNNI operator (no reference counterpart);
                      Tv from TMat/tmat3d.cpp:179-201 energies, Th = (x)K            */
  OR_XDM = 7,      /* TMat/tmatreduce.cpp:1024-1400 DNNI reduced by D_L (x flip)     */
  OR_ZBM = 8,      /* TMat/tmatreduce.cpp:1403-1868 BNNNI/3NN diagonal TM reduced by
                      two-site translations and i -> -i (x flip)                     */
  OR_YAT = 9,      /* TMat/tmat.cpp:104-243   ANNNI axial TM, two layers, 4^L; the
                      transpose is YATMatVecProd2D_T (tmat.cpp:104-184)             */
  OR_Y2AT = 10,    /* TMat/tmat.cpp:246-310   ANNNI, layer pairs, 4^L; its transposed
                      instance is the operator itself (tmat.hpp:382-384)            */
  OR_YAM = 11,     /* TMat/tmatreduce.cpp:661-949 ANNNI axial TM reduced by the
                      layer-wise D_L (x flip); transpose YAMMatVecProd2D_T :951-1015 */
  OR_YBM = 12      /* TMat/tmatreduce.cpp:1017-1021 BNNNI axial TM (YAM, layertype 1) */
};

typedef struct or_op {
  int kind;
  int L;            /* layer size (bits per state)                 */
  int size1, size2; /* 3D cross-section, size1 rows x size2 cols   */
  int bc;           /* boundary / extend code as the reference      */
  int censym, ising, rotate;
  int nleadbit;
  double betaJ, kappa, kappa2, h;
  double nf;        /* normfactor                                    */
  double bzwj1[2], bzwj2[2], bzwj3[2];
  double bzwj2b[4], bzwj2c[16];  /* ZBM step tables, TMat/tmatreduce.cpp:1416-1424 */
  int64_t nstate;   /* 2^L                                           */
  int64_t ncol;     /* rows()/cols()                                 */
  double *tvec;     /* length ncol                                   */
  uint64_t *reps;   /* reduced only: orbit representatives, ascending */
  uint16_t *dups;   /* reduced only: orbit sizes                      */
  int64_t *swp;     /* YAM/YBM: orbit of the layer-swapped representative */
  unsigned char (*lead)[4];  /* YAM/YBM: lead-layer periods (leadprdlist) */
  int64_t nlead;
} or_op;

/* Build an operator.  bc follows the reference: XAT 0,1,5,9,13,17; XDT 0,5,9;
 * XAM2D extend_cond (bit0 open boundary, 2 = no spin-flip folding);
 * XDM, ZBM extend_cond (2 = no spin-flip folding; periodic only);
 * XAM3D extend_cond (bit0 ignored, as the reference warns).
 * Returns NULL on invalid input. */
or_op *or_op_create(int kind, int size1, int size2, int bc,
                    double betaJ, double kappa, double kappa2, double h);
void   or_op_destroy(or_op *op);
/* y = T x (transpose=0) or y = T^T x (transpose=1; only ZBT differs). */
void   or_op_apply(const or_op *op, const double *x, double *y, int transpose);
int64_t or_op_ncol(const or_op *op);
double  or_op_normfactor(const or_op *op);
const double *or_op_tvec(const or_op *op);
const uint64_t *or_op_reps(const or_op *op);
const uint16_t *or_op_dups(const or_op *op);

/* Orbit enumeration restated from TMat/tmatreduce.cpp:19-111 and
 * TMat/tmat3d.cpp:17-98.  Returns the orbit count; reps/dups may be NULL to
 * count only, else must hold at least the returned count (call twice). */
int64_t or_orbits_x2d(int L, int bc, int censym, uint64_t *reps, uint16_t *dups);
int64_t or_orbits_x3d(int s1, int s2, int censym, int rotate, uint64_t *reps, uint16_t *dups);
int64_t or_orbits_z2d(int L, int censym, uint64_t *reps, uint16_t *dups);  /* tmatreduce.cpp:242-334 */
/* tmatreduce.cpp:114-239: two-layer orbits; reps/dups/swp of length M
 * (call with NULL arrays to count) */
int64_t or_orbits_y2d(int L, int censym, uint64_t *reps, uint16_t *dups, int64_t *swp);

/* Diagonal Tv^(1/2) generators (genTvec) for the full operators, exactly as
 * the reference builds them (argument betaJ is the one passed to genTvec). */
void or_tvec_xat(int L, int bc, double betaJ, double kappa, double h, double *V);  /* tmat.hpp:208-310 */
void or_tvec_xdt(int L, int bc, double betaJ, double h, double *V);               /* tmat.hpp:437-504 */
void or_tvec_zbt(int L, double betaJ, double kappa2, double h, double *V);        /* tmat.hpp:537-556 */

/* Observables of an eigenvector pair (the X_phys.dat figures), sequential
 * sums in the reference's order: out[0..3] the moments (moments()), out[4..]
 * the extra observables (phys_extend()), *nout = 4 + their count.  ql may be
 * NULL (ql = qr, TMat/pfunc.hpp:170-172).
 *   XAT, XDT, XAM2D  TMat/tmatphys.cpp:13-57, TMat/tmatreduce.cpp:565-613
 *   XAT3D, XAM3D     TMat/tmatphys.cpp:13-37, TMat/tmat3d.cpp:137-177
 *   XDM              TMat/tmatreduce.cpp:565-590 (its moments), 1289-1320
 *   ZBT, ZBM         TMat/tmatphys.cpp:106-132, TMat/tmatreduce.cpp:1748-1781
 *   YAT              TMat/tmatphys.cpp:71-104
 *   YAM, YBM         TMat/tmatreduce.cpp:882-909
 *   Y2AT             TMat/tmat.cpp:312-317 (zeros) */
void or_op_moments(const or_op *op, const double *ql, const double *qr, double out[8], int *nout);

/* Configuration chains (DumpConfig), sequential, without the Q cache (which
 * only saves recomputation):
 *   XAT   XAT2DDumpConfig::dump  DumpConfig/dumpconfig.cpp:88-136
 *   YAT   YAT2DDumpConfig::dump  DumpConfig/dumpconfig.cpp:171-222
 *   XAM2D, XDM, ZBM, YAM, YBM    DumpConfig/dumpreduce.cpp:47-122, 516-590,
 *                                409-470, 156-233, 279-356
 * Start weights vec^2 (SymDumpConfig::initq, dumpconfig.cpp:24-35) or
 * vec vecl (GenDumpConfig::initq, :44-56; vecl non-NULL for YAT, ZBM, YAM,
 * YBM).  vec and vecl are taken in absolute value; u holds the rand() draws
 * in call order; states receives nchain x length states.  Orbit members are
 * taken in ascending order (the reference iterates an unordered_set,
 * TMat/tmatreduce.cpp:615-657, 918-943, 1783-1810).  Returns 0, or -1 for an
 * operator kind without a sampler or a missing vecl. */
int or_sample_configs(const or_op *op, const double *vec, const double *vecl, const double *u,
                      int nchain, int length, uint64_t *states);

int or_num_threads(void);
/* threads of the following calls (OpenMP); the CPU baseline times 1 and all */
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
