/*
 * latttmat.h -- C ABI of the MI355X-native transfer-matrix eigensolver.
 *
 * The ABI replaces the operator surface of the reference program (paths are
 * relative to the reference tree):
 *   - the Spectra "OpType" concept rows()/cols()/perform_op(x,y)
 *     (TMat/tmat.hpp:86-89) and the transpose instance get_t_instance()
 *     (TMat/tmat.hpp:534-535, TMat/tmatreduce.hpp:195-196);
 *   - operator construction by model/dim/axialdir (TMat/pfunc.cpp:16-106);
 *   - the eigen driver pfunc_inner / getlefteigvec (TMat/pfunc.hpp:39-202)
 *     whose restart loop runs on the GPU here;
 *   - the orbit tables idxlist/duplist (TMat/tmatreduce.hpp:51-52,
 *     TMat/tmat3d.hpp:43-44).
 * No C++ or HIP types cross this boundary: plain pointers, sizes and enums.
 * All functions are thread-compatible: an operator handle must not be used
 * from two threads at once without external locking.
 */
#ifndef LATTTMAT_H
#define LATTTMAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TM_OK = 0,
  TM_EINVAL = 1,        /* invalid argument / unsupported model path          */
  TM_ENOMEM = 2,        /* device or host allocation failed                   */
  TM_EHIP = 3,          /* a HIP runtime call failed                          */
  TM_ENOCONV = 4,       /* eigensolver did not converge (pfunc.hpp:110-114)   */
  TM_EUNSUPPORTED = 5   /* path exists in the reference but is not built here */
} tm_status;

typedef enum { TM_F64 = 0, TM_F32 = 1 } tm_dtype;

typedef enum {
  TM_PREC_F64 = 0,      /* plain FP64 everywhere (the oracle method)           */
  TM_PREC_MIXED = 1     /* FP32 T.x, FP64 orthogonalisation + FP64 refinement  */
} tm_precision;

/* Operator kinds as built (reported by tm_op_kind). */
typedef enum {
  TM_KIND_XAT = 1,      /* TMat/tmat.cpp:39-102       ANNNI perp-TM, 2^L        */
  TM_KIND_XDT = 2,      /* TMat/tmat.cpp:322-476      DNNI perp-TM, 2^L         */
  TM_KIND_ZBT = 3,      /* TMat/tmat.cpp:480-662      BNNNI/3NN diagonal TM     */
  TM_KIND_XAM2D = 4,    /* TMat/tmatreduce.cpp:338-563 ANNNI reduced            */
  TM_KIND_XAM3D = 5,    /* TMat/tmat3d.cpp:102-340    3D ANNNI reduced          */
  TM_KIND_XAT3D = 6,    /* full 2^(s1*s2) 3D ANNNI (no reference counterpart)   */
  TM_KIND_XDM = 7,      /* TMat/tmatreduce.cpp:1024-1400 DNNI reduced           */
  TM_KIND_ZBM = 8,      /* TMat/tmatreduce.cpp:1403-1868 BNNNI/3NN diagonal reduced */
  TM_KIND_YAT = 9,      /* TMat/tmat.cpp:104-243      ANNNI axial TM, 4^L       */
  TM_KIND_Y2AT = 10,    /* TMat/tmat.cpp:246-310      ANNNI axial TM, layer pairs, 4^L */
  TM_KIND_YAM = 11,     /* TMat/tmatreduce.cpp:661-1015 ANNNI axial TM reduced  */
  TM_KIND_YBM = 12      /* TMat/tmatreduce.cpp:1017-1021 BNNNI axial TM reduced */
} tm_kind;

/* Model parameters: the tmat input-file keys (TMat/read_input.cpp:29-60). */
typedef struct {
  char   model;         /* 'A', 'B' or 'D'                                     */
  int    dim;           /* 2 or 3                                              */
  int    axialdir;      /* as in the input file, dispatched like pfunc.cpp     */
  int    neigs;         /* >1 disables spin-flip folding (pfunc.cpp:22)        */
  int    size1;         /* 2D: L; 3D: rows                                     */
  int    size2;         /* 3D: columns (0 in 2D)                               */
  double betaJ;         /* "J"                                                 */
  double kappa;         /* kappa1                                              */
  double kappa2;        /* kappa2 (0 if absent)                                */
  double h;             /* external field                                      */
  int    flags;         /* TM_FLAG_*                                           */
} tm_params;

/* Use the full-space operator for 3D (2^(s1*s2) states) instead of the
 * reference's symmetry-reduced one.  No reference counterpart. */
#define TM_FLAG_FULLSPACE 1
/* Force the generic expand/butterfly/gather kernels for reduced operators
 * (disables the two-pass fused kernel; used for cross-checking). */
#define TM_FLAG_GENERIC_REDUCED 2
/* Keep the default tile sizes: no timing of candidate tile sizes on the first
 * apply of each (kernel, width, precision). */
#define TM_FLAG_NO_AUTOTUNE 4

typedef struct tm_op tm_op;

/* pfunc dispatch (TMat/pfunc.cpp:16-106) + operator constructors:
 * XATMatVecProd2D (tmat.cpp:39-56), XDTMatVecProd2D (:322-341),
 * ZBTMatVecProd2D (:480-518), XAMMatVecProd2D (tmatreduce.cpp:338-365),
 * XAMMatVecProd3D (tmat3d.cpp:102-134).  Builds tables on the device. */
tm_status tm_op_create(const tm_params *p, tm_op **out);
void      tm_op_destroy(tm_op *op);

/* rows()/cols() (TMat/tmat.hpp:86-87) */
int64_t   tm_op_rows(const tm_op *op);
int64_t   tm_op_cols(const tm_op *op);
/* getnormfactor() (TMat/tmat.hpp:116) */
double    tm_op_normfactor(const tm_op *op);
/* Scalar==double selects the symmetric solver (TMat/pfunc.hpp:83-86). */
int       tm_op_is_symmetric(const tm_op *op);
tm_kind   tm_op_kind(const tm_op *op);
/* suggested_nev (TMat/tmat.hpp:88) */
int       tm_op_suggested_ncv(const tm_op *op, int neigs);
/* getLayerSize() (TMat/tmat.hpp:129,141) */
int       tm_op_layer_size(const tm_op *op);
/* OpCounter::get() (TMat/utility.hpp:191-199): matvecs since creation. */
int64_t   tm_op_matvec_count(const tm_op *op);

/* perform_op(x_in, y_out) (TMat/tmat.hpp:89): host pointers of length
 * cols(), synchronous, FP64.  transpose=1 applies T^T (the transpose
 * instance, e.g. ZBTMatVecProd2D_T, tmat.cpp:652-662); for symmetric
 * operators it is the same as transpose=0. */
tm_status tm_op_apply(const tm_op *op, const double *x, double *y, int transpose);

/* Stream-ordered device apply of nvec column vectors (leading dimension ld
 * elements) of type dtype.  compute selects FP64 or FP32 arithmetic inside
 * the operator (TM_F32 = the mixed-precision T.x).  stream is a hipStream_t
 * (NULL = default stream). */
tm_status tm_op_apply_dev(const tm_op *op, const void *x, void *y, int64_t ld,
                          int nvec, tm_dtype dtype, tm_dtype compute,
                          int transpose, void *stream);

/* tm_op_apply_dev for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, kernels and results as
 * tm_op_apply_dev (with the tile autotuner off: the default tiles); it calls
 * no hipMalloc, hipFree or allocating library routine.  Every buffer beyond
 * x and y (the FP32 input scaling's partial sums and scales, the reduced
 * operators' permutation buffers, the operator workspace) comes from a
 * file-scope device array of 32 MiB that each GPU holds until the process
 * ends; an apply that needs more returns TM_EINVAL and names the size.  The
 * array's device address is looked up on the first call on each device (with
 * that device current).  Calls on one GPU share the array: calls that overlap
 * on one GPU race (serialise them, e.g. on one stream).  The apply is not
 * ordered against the operator's own stream (tm_op_apply, tm_solve). */
tm_status tm_op_apply_dev_noalloc(const tm_op *op, const void *x, void *y, int64_t ld,
                                  int nvec, tm_dtype dtype, tm_dtype compute,
                                  int transpose, void *stream);

/* Host-pointer apply with a chosen compute precision (TM_F32 = the
 * mixed-precision T.x with a power-of-two input scale).  No reference
 * counterpart; used by the parity tests. */
tm_status tm_op_apply_prec(const tm_op *op, const double *x, double *y, int transpose,
                           tm_dtype compute);

/* Time `reps` device applies (after `warmup` untimed ones) on the operator's
 * stream with HIP events; inputs are resident in device memory.  Returns the
 * mean milliseconds per apply.  No reference counterpart (the reference only
 * times the whole run, TMat/main.cpp:48-53). */
tm_status tm_op_time_apply(const tm_op *op, tm_dtype io, tm_dtype compute, int transpose,
                           int warmup, int reps, double *ms_per_apply);

/* Orbit tables idxlist/duplist (reduced operators only): reps ascending,
 * dups = orbit sizes.  Arrays of length cols(). */
tm_status tm_op_orbits(const tm_op *op, uint64_t *reps, uint16_t *dups);

/* Options of the eigen driver.  Defaults (tm_solve_opts_default): nev=1,
 * ncv=0 (=suggested_nev), tol=1e-10, maxit=1000, maxtrial=3 (pfunc.hpp:87),
 * precision FP64, seed 1, want_left=0, use_graph=1, basis_f32=1, refine_extra=1,
 * verify=1. */
typedef struct {
  int          nev;
  int          ncv;
  double       tol;
  int          maxit;
  int          maxtrial;
  tm_precision precision;
  uint64_t     seed;
  int          want_left;   /* also solve T^T for the left vector (pfunc.hpp:47-67) */
  int          want_vectors;/* copy right (and left) vectors to the host          */
  double       mixed_tol;   /* residual target of the FP32 phase (default 2e-7;
                               it also stops at the FP32 noise floor)     */
  int          verbose;
  int          refine;      /* mixed precision, FP64 phase: 1 = block
                               Rayleigh-Ritz on the FP32 Ritz vectors, accepted
                               when r^2/(gap |theta|) <= 1e-13 (symmetric only);
                               2 = Krylov-Schur warm-started from the FP32 Ritz
                               vectors, run to tol; 0 = 1 if symmetric else 2 */
  int          use_graph;   /* 1 (default): restart cycles after the first
                               are captured once into a hipGraph and
                               replayed; 0: launched eagerly            */
  int          basis_f32;   /* mixed precision, FP32 phase: 1 (default) stores
                               the Krylov basis in FP32 (half the bytes of
                               every orthogonalisation sweep; products and
                               sums still in FP64); 0 stores it in FP64   */
  int          refine_extra;/* refine 1: the FP64 Rayleigh-Ritz block holds
                               nev + refine_extra Ritz vectors (default 1);
                               a larger block widens the gap in the error
                               bound, a smaller one saves FP64 T.x; a block
                               whose bound is not met is iterated          */
  int          verify;      /* 1 (default): before returning TM_OK, check the
                               leading pair's explicit FP64 residual
                               ||Tx - lambda x|| / |lambda| <= 1e-6 (one FP64
                               T.x, or the refinement's explicit residual);
                               a failed check returns TM_EHIP              */
  int          block;       /* symmetric operators: 0 or 1 (default) = one
                               vector per Krylov step; b in [2, 8] = block
                               Krylov-Schur with b vectors per step (needs
                               ncv >= nev + b and ncv + 2b <= 64; the Gram
                               of the block Gram-Schmidt on FP64 MFMA);
                               operators with cols() < 8 (ncv + 2b) and
                               nonsymmetric ones run the one-vector form */
} tm_solve_opts;

/* Results.  Eigenvalues are those of the operator as applied (the reference
 * writes them divided by the normfactor, pfunc.hpp:117-118).  Vector buffers
 * are owned by the caller and may be NULL. */
typedef struct {
  int      nconv;
  double  *evals_re;        /* [nev]                                        */
  double  *evals_im;        /* [nev]                                        */
  double  *resid;           /* [nev] relative residual ||Tx-lx||/|l| (FP64)  */
  double  *vecs;            /* right vectors; symmetric: nev x cols() FP64
                               column-major; general: cols() x nev complex128
                               interleaved (re,im)                            */
  double  *vecl;            /* left leading vector, cols() FP64 (real part)  */
  double   eval_left;       /* leading eigenvalue of T^T (NaN if not solved) */
  int64_t  matvecs_f64;
  int64_t  matvecs_f32;
  int64_t  restarts;
  int      trials;
  double   seconds;         /* wall time of the solve (setup excluded)        */
  int64_t  graph_cycles;    /* restart cycles replayed from a hipGraph        */
  double   resid_check;     /* explicit FP64 relative residual of the leading
                               pair (NaN if opts.verify is 0)                 */
} tm_solve_out;

void      tm_solve_opts_default(tm_solve_opts *o);
/* pfunc_inner's solve + retry loop (TMat/pfunc.hpp:87-114) on the GPU. */
tm_status tm_solve(tm_op *op, const tm_solve_opts *opts, tm_solve_out *out);

/* Solve n independent operators (any kinds and sizes) concurrently:
 * instances are taken largest first by nworkers host threads (0 = 4), each
 * driving the instance's own HIP stream, so small solves run beside the
 * large ones; every instance's restart cycles are graph-replayed as in
 * tm_solve.  opts[i] / outs[i] as for tm_solve.  Returns the first non-OK
 * status (every instance is attempted; tm_last_error names the instance).
 * No reference counterpart (the reference solves one parameter point per
 * run, TMat/main.cpp:31-58). */
tm_status tm_solve_batch(tm_op *const *ops, int n, const tm_solve_opts *opts,
                         tm_solve_out *outs, int nworkers);

/* Hellmann-Feynman derivative of the leading eigenvalue of the operator as
 * applied, with respect to (betaJ, kappa, kappa2, h):
 *   dlam[i] = <l, dT/dtheta_i r> / <l, r>.
 * l and r are host vectors of length cols() (l = r for symmetric T).
 * Also returns d ln(normfactor)/d theta_i in dlnnf (may be NULL), so that
 * the physical eigenvalue lam/nf has d ln(lam/nf) = dlam/lam - dlnnf.
 * No reference counterpart (the reference has no gradients). */
tm_status tm_grad_hf(const tm_op *op, const double *l, const double *r,
                     double dlam[4], double dlnnf[4]);

/* Observables of the leading eigenvector pair, reduced on the device: the
 * figures the reference writes to X_phys.dat (TMat/pfunc.hpp:167-200).
 *   out[0..3]  the magnetisation moments <m^k>/L, k = 1..4
 *              (SymMatPhys::moments, TMat/tmatphys.cpp:13-37; GenMatPhys and
 *              ZMatVecProd2D, :71-98, :106-132, normalised by sum ql qr;
 *              the reduced operators over their representatives,
 *              TMat/tmatreduce.cpp:565-590, 882-909, 1748-1775,
 *              TMat/tmat3d.cpp:137-163; zeros for Y2AT, TMat/tmat.cpp:312-317);
 *   out[4..]   the operator's extra observables (phys_extend): <node>,
 *              <node^2> (TMat/tmatphys.cpp:39-57, TMat/tmatreduce.cpp:593-613;
 *              NaN for the 3D operators, TMat/tmat3d.cpp:165-177), or
 *              <|m_st|>/L, <|m|>/L, <m_st^2>/L^2, <m^2>/L^2 for the reduced
 *              DNNI operator (TMat/tmatreduce.cpp:1289-1320), or none;
 *   *nout      4 + the number of extra observables (out holds up to 8).
 * qr and ql are host vectors of length cols() as returned by tm_solve; ql
 * may be NULL (then ql = qr, as TMat/pfunc.hpp:170-172 does), and the
 * symmetric operators weight by ql^2. */
tm_status tm_op_moments(const tm_op *op, const double *ql, const double *qr, double out[8],
                        int *nout);

/* Plant equilibrium configurations from the leading eigenvector, as the
 * reference's dumpconfig program does (DumpConfig/main.cpp:30-144), as
 * Markov chains on the device:
 *   XAT (A, axialdir 1)       XAT2DDumpConfig::dump, DumpConfig/dumpconfig.cpp:88-136
 *   YAT (A, axialdir 2)       YAT2DDumpConfig::dump, DumpConfig/dumpconfig.cpp:171-222
 *   XAM2D (A, axialdir 5/13)  XAM2DDumpConfig::dump, DumpConfig/dumpreduce.cpp:47-122
 *   YAM (A 6), YBM (B 6)      YAM2DDumpConfig::dump, YBM2DDumpConfig::dump,
 *                             DumpConfig/dumpreduce.cpp:156-233, 279-356
 *   ZBM (B, axialdir 7)       ZBM2DDumpConfig::dump, DumpConfig/dumpreduce.cpp:409-470
 *   XDM (D, axialdir 5/6)     XDM2DDumpConfig::dump, DumpConfig/dumpreduce.cpp:516-590
 * vec: the right eigenvector as X_vec.dat holds it (cols() entries); vecl:
 * the left eigenvector (X_vecl.dat) for the non-symmetric kinds, NULL
 * otherwise; absolute values are taken (DumpConfig/dumpconfig.hpp:71).  The
 * start is drawn with weight vec^2 (symmetric) or vec vecl.  u: uniforms in
 * [0, 1) in the order MyUtility::rand() is called, chain after chain:
 * length per chain for XAT and YAT, 2 length - 1 for the reduced kinds.
 * states: nchain * length states, chain-major: layer states (XAT, XAM2D,
 * XDM, ZBM) or two-layer states (YAT, YAM, YBM; index (high << L) | low).
 * The reference's output lines are the states themselves, the low layers
 * (YAM, YBM), or the first state's low layer followed by every state's high
 * layer (YAT); DumpConfig_Base::output, dumpconfig.hpp:32-40.
 * Each draw picks the first index whose running weight exceeds u times the
 * total (std::upper_bound on the cumulative sums).  One difference: an
 * orbit's members are taken in ascending order, where the reference takes
 * the iteration order of an unordered_set (getEqvStates,
 * TMat/tmatreduce.cpp:615-657, 918-943, 1783-1810); the probabilities are
 * the same.  Other operator kinds return TM_EUNSUPPORTED. */
tm_status tm_sample_configs(const tm_op *op, const double *vec, const double *vecl,
                            const double *u, int nchain, int length, uint64_t *states);

/* Thread-local message for the last non-OK status. */
const char *tm_last_error(void);
/* Library version string. */
const char *tm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LATTTMAT_H */
