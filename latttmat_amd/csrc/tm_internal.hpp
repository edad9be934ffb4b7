// tm_internal.hpp -- shared declarations of the MI355X transfer-matrix library.
//
// Layering:
//   tm_api.cpp      C ABI (include/latttmat.h), pfunc-style dispatch
//   tm_ops.cpp      operator classes: tables on device, apply() launch plans
//   tm_kron.hip     full-space Kronecker (butterfly) passes: XAT, 3D full
//   tm_reduced.hip  orbit enumeration + reduced operators (XAM 2D/3D)
//   tm_sweep.hip    aux-bit sweeps: XDT (DNNI), ZBT/ZBT^T (BNNNI/3NN)
//   tm_krylov.hip   Krylov vector kernels (multi-dot, CGS2 update, restart)
//   tm_solver.cpp   thick-restart Krylov-Schur driver, mixed precision
//   tm_dense.cpp    small dense eigensolvers for the Ritz problem (host)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/latttmat.h"

namespace ltm {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  tm_status status;
  Error(tm_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line);

#define TM_HIP(expr)                                                   \
  do {                                                                 \
    hipError_t _e = (expr);                                            \
    if (_e != hipSuccess) ::ltm::throw_hip(_e, #expr, __FILE__, __LINE__); \
  } while (0)

#define TM_CHECK(cond, status, msg)                     \
  do {                                                  \
    if (!(cond)) throw ::ltm::Error((status), (msg));    \
  } while (0)

// Launch-error check after a kernel launch.
#define TM_LAUNCH_CHECK() TM_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Autotuner for tile sizes.  For a (family, kernel id, L, element size) key,
// the first call outside graph capture with tune == true runs every
// candidate through run(t) on stream s (one warm-up, then three timed with
// HIP events), caches the fastest and returns it; later calls with tune ==
// true return the cached choice.  With tune == false, or during a capture
// before a choice exists, returns cands[0] (the LDS-size default).  Each
// kernel family has its own id space: a candidate value means something
// different to each family's launcher.  run(t) must leave its input intact
// (callers pass tune == false for in-place applies).
enum TuneFamily : int { TUNE_KRON = 1, TUNE_KRON_DUAL = 2, TUNE_KRON_RANGE = 3, TUNE_SWEEP = 4 };
int autotune_tile(TuneFamily fam, uint32_t id, int L, int es, const std::vector<int> &cands,
                  bool tune, hipStream_t s, const std::function<void(int)> &run);
// the cached choice for a key, or -1 (for tests)
int autotune_cached(TuneFamily fam, uint32_t id, int L, int es);

// Order all work enqueued on `waiter` from now on after the work enqueued
// on `producer` so far (an event recorded on producer, waited on by
// waiter).  No-op when the two are the same stream.
void stream_wait(hipStream_t waiter, hipStream_t producer);
// Wait for everything queued on s by polling the stream from this thread
// (hipStreamQuery): a blocking wait returns only after the runtime's
// completion interrupt, tens of microseconds later, which the Krylov loop
// would pay once per restart cycle with the device idle.
void stream_sync_poll(hipStream_t s);

// Stream-ordered copy that returns once the data has arrived.  The library
// never uses the legacy null stream, whose implicit synchronisation with
// every blocking stream would break a graph capture running on another
// thread.
void copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
// The same for large host <-> device copies of pageable host memory: through
// a pinned double buffer, DMA overlapping a multi-threaded host copy.
void copy_large(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);

// Graph capture and device allocation exclude each other across threads:
// a capture holds this lock exclusively (and marks its thread), allocations
// and frees hold it shared.  An allocation on the capturing thread throws.
struct CaptureGuard {
  CaptureGuard();
  ~CaptureGuard();
  CaptureGuard(const CaptureGuard &) = delete;
  CaptureGuard &operator=(const CaptureGuard &) = delete;
};
// true on a thread that holds a CaptureGuard
bool capturing_here();
// hipGraphExecDestroy under the shared side of the same lock
void destroy_graph_exec(hipGraphExec_t g);

// page-locked host buffer (RAII) for the solver's small per-cycle transfers
class PinnedBuf {
 public:
  PinnedBuf() = default;
  explicit PinnedBuf(size_t bytes) { alloc(bytes); }
  // a buffer cannot be freed inside a graph capture: it is then leaked
  ~PinnedBuf()
  {
    try { release(); } catch (...) {}
  }
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  void alloc(size_t bytes);
  void ensure(size_t bytes) { if (bytes > n_) alloc(bytes); }
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p_); }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

// device buffer (RAII)
// ---------------------------------------------------------------------------
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  // a buffer cannot be freed inside a graph capture: it is then leaked
  ~DevBuf()
  {
    try { release(); } catch (...) {}
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf &operator=(DevBuf &&o) {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t bytes);
  void ensure(size_t bytes) { if (bytes > n_) alloc(bytes); }
  void release();
  template <typename T> T *as() const { return static_cast<T *>(p_); }
  void *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

// Device and pinned buffers plus captured restart-cycle graphs that the
// eigensolver keeps on an operator between solves (slot numbers are
// assigned by the eigensolver).  A graph is keyed by every pointer and size its
// launches embed, so a key match means identical kernel arguments.
struct SolverCache {
  static constexpr int NDEV = 29, NPIN = 4;
  DevBuf dev[NDEV];
  PinnedBuf pinned[NPIN];
  struct Graph {
    std::array<uint64_t, 12> key;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  SolverCache() = default;
  SolverCache(const SolverCache &) = delete;
  SolverCache &operator=(const SolverCache &) = delete;
  ~SolverCache()
  {
    for (auto &g : graphs) destroy_graph_exec(g.exec);
  }
};

// ---------------------------------------------------------------------------
// diagonal (Tv^1/2) descriptor shared by the full-space kernels
// ---------------------------------------------------------------------------
// D(s) = tab[i1(s)] * tab[64 + i2(s)] * tab[128 + i0(s)], with the integer
// indices computed from popcounts of the state (no per-state table in HBM).
enum DiagKind : int {
  DIAG_X2D_PBC = 0,   // XAT periodic      (tmat.hpp:218-231)
  DIAG_X2D_OBC = 1,   // XAT open          (tmat.hpp:233-247)
  DIAG_X2D_C5 = 5,    // XAT confined      (tmat.hpp:249-262)
  DIAG_X2D_C9 = 9,    //                   (tmat.hpp:264-277)
  DIAG_X2D_C13 = 13,  //                   (tmat.hpp:279-292)
  DIAG_X2D_C17 = 17,  //                   (tmat.hpp:294-307)
  DIAG_X3D = 30,      // 3D layer energies (tmat3d.cpp:188-199)
  DIAG_XDT_PBC = 40,  // DNNI periodic     (tmat.hpp:445-458)
  DIAG_XDT_C5 = 45,   //                   (tmat.hpp:475-487)
  DIAG_XDT_C9 = 49,   //                   (tmat.hpp:489-501)
  DIAG_ZBT = 60,      // BNNNI/3NN         (tmat.hpp:537-556)
  // two-layer states (l | r << L, 2L bits; L is the layer size)
  DIAG_Y2TV = 70,     // Y2AT pair energy  (tmat.hpp:396-416): i1 = NN bonds in
                      // both layers, i2 = NN between them, i0 = down spins
  DIAG_Y2D = 71       // NN between the layers of a pair only (i1)
};

struct DiagSpec {
  int kind;
  int L;       // bits per state
  int s1, s2;  // 3D cross-section
  // 3D column masks (bits of columns [0,c) in every row) for row rotations
  uint64_t cm1, cmr1, cm2, cmr2;  // cols [0,1), [0,s2-1), [0,2), [0,s2-2)
};

// Largest dynamic LDS allocation per workgroup on the current device (bytes);
// set when the first operator is created.  Tile sizes are chosen from it.
extern size_t g_lds_per_block;
extern int g_num_cu;  // compute units of the device (set at first operator creation)

// ---------------------------------------------------------------------------
// launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
// Kronecker operator y = Dout o (x)_{i<L} K o Din o x, K = [[ka, kb],[kb, ka]].
// in/out are FP64 or FP32 (io), compute precision comp.  scale_in points to a
// device double (input multiplied by it, output divided by it) or is null.
// Kronecker butterflies on bits [b0, b0 + nbits) of an Ltot-bit index (the
// other bits ride along): K = [[ka, kb], [kb, ka]] on every transformed bit,
// optional diagonals on the first load (pro) and the last store (epi).
// Tables are 192-entry device arrays of the compute type.  mid is a
// compute-type buffer of 2^Ltot elements for the passes between the first
// and the last (it may be out, or in when in is compute-type scratch).
// scale (device double or null): input times *scale if sin, output divided
// by it if sout.
struct KronRange {
  int Ltot, b0, nbits;
  bool pro, epi;
  DiagSpec dpro, depi;
  const void *tpro, *tepi;
  double ka, kb;
};
void launch_kron_range(const KronRange &kr, const void *in, void *out, void *mid,
                       tm_dtype io_in, tm_dtype io_out, tm_dtype comp, const double *scale,
                       bool sin, bool sout, bool tune, hipStream_t s);

// ANNNI axial transfer matrices over two-layer states (tm_axial.hip).
//   YAT:  y = nf D(out) Swap (x)_{high bits} K2 x      (TMat/tmat.cpp:207-242)
//   YATT: y = nf (x)_{high bits} K2 Swap D(in) x       (TMat/tmat.cpp:126-183)
//   Y2AT: y = Tv (x)_{low} K2 D1 (x)_{high} K2 Tv x    (TMat/tmat.cpp:265-302)
// D(a, b) = Tv(layer) K(a, b) from the 192-entry table tab (i1 = NN bonds of
// the layer, i2 = down spins of the layer, i0 = popcount(a ^ b)); Y2AT uses
// tab (DIAG_Y2TV, sqrt(nf) folded) and tabb (DIAG_Y2D).
struct AxialPlan {
  int kind;  // TM_KIND_YAT or TM_KIND_Y2AT
  int L;     // layer size; states are 2L bits
  double ka, kb;  // K2 = [[ka, kb], [kb, ka]]
  const double *tab64, *tabb64;
  const float *tab32, *tabb32;
  bool tune = false;
};
void launch_axial(const AxialPlan &plan, const void *in, void *out, tm_dtype io, tm_dtype comp,
                  int transpose, void *work, const double *scale, hipStream_t s);
size_t axial_work_bytes(int L, tm_dtype comp);
// out[(a << L) | b] = f in[(b << L) | a] with f = W(b, a) (wsel 0), W(a, b)
// (wsel 1) or 1 (wsel 2), W = dev::w_layers over tab (compute type; 192
// entries, 320 with ybm); sflags 1: times *scale, 2: divided by it.
void launch_layer_swap(const void *in, void *out, tm_dtype ti, tm_dtype to, tm_dtype comp, int L,
                       const void *tab, int wsel, int ybm, const double *scale, int sflags,
                       hipStream_t s);

struct KronPlan {
  int L;
  bool prologue, epilogue;  // apply D on the first / last pass
  DiagSpec diag;
  const double *tab64;      // device, 192 doubles
  const float *tab32;       // device, 192 floats
  double ka, kb;
  bool tune = false;        // let the autotuner pick the tile size
};
// Dual-number apply: out (2 doubles per state: T in, dT/dtheta in) from a
// real input; tab_dual holds 192 (value, derivative) pairs and dka/dkb the
// derivatives of ka/kb.
void launch_kron_dual(const KronPlan &plan, const double *in, void *out, const void *tab_dual,
                      double dka, double dkb, hipStream_t s);
void launch_kron(const KronPlan &plan, const void *in, void *out, tm_dtype io_in,
                 tm_dtype io_out, tm_dtype comp, void *work, const double *scale_in,
                 hipStream_t s);
size_t kron_work_bytes(int L, tm_dtype comp);

// Sweeps (XDT, ZBT); transpose only differs for ZBT.
struct SweepPlan {
  int kind;           // TM_KIND_XDT or TM_KIND_ZBT
  int L;
  int bc;             // XDT boundary (0, 5, 9)
  DiagSpec diag;
  const double *tab64;
  const float *tab32;
  double b1[2], b2[2], b3[2];
  double nf;          // XDT output factor (ZBT folds it into the diagonal)
  double db1[2], db2[2], db3[2], dnf;  // derivatives (dual-number apply only)
  bool tune = false;  // let the autotuner pick the tile size
};
void launch_sweep_dual(const SweepPlan &plan, const double *in, void *out, int transpose,
                       void *work, const void *tab_dual, hipStream_t s);
size_t sweep_dual_work_bytes(int kind, int L);
void launch_sweep(const SweepPlan &plan, const void *in, void *out, tm_dtype io,
                  tm_dtype comp, int transpose, void *work, const double *scale_in,
                  hipStream_t s);
size_t sweep_work_bytes(int kind, int L, tm_dtype comp);


// ---- orbit enumeration (tm_reduced.hip) ----
struct GroupSpec {
  int dim;          // 2 or 3
  int L;            // total bits
  int s1, s2;       // 3D rows, cols
  int bc;           // 2D: 0 periodic, 1 open
  int censym;       // include global spin flip
  int rotate;       // 3D: include 90-degree rotation
  int zgroup;       // 2D periodic: translations by two sites and the
                    // reflection i -> -i (genZ2DDuplist, TMat/tmatreduce.cpp:242-334)
                    // instead of all translations and i -> L-1-i
  int ylayer;       // > 0: two-layer states of this layer size (L = 2 ylayer
                    // bits): rotations and reflection of both layers
                    // together (genY2DDuplist, TMat/tmatreduce.cpp:114-239)
};
// Enumerate orbits on the device.  Fills reps (ascending) and dups (orbit
// sizes), returns M.  If map != nullptr it receives map[s] = orbit index of s.
int64_t enumerate_orbits(const GroupSpec &g, DevBuf &reps, DevBuf &dups, DevBuf *map,
                         hipStream_t s);
// xf[s] = xr[orbit(s)] / sqrt(orbit size) over all 2^L states (device).
void expand_orbits(const GroupSpec &g, const uint64_t *reps, const uint16_t *dups, int64_t M,
                   const double *xr, double *xf, hipStream_t s);
// Fused-path tables (2D periodic, L = 2n even).  A state s is split into
// its even bits E and odd bits O (n bits each); a rotation of s by two sites
// rotates E and O together, so the transform over O only has to be done for
// columns E that are necklace representatives (minimal rotations).  The
// site mirror i -> -i is in the group too: it maps E to rotl_1(rev E) and O
// to rev O, so a column whose necklace is the mirror of another's is that
// column with its O words reversed.  Pass 1 therefore stores one column per
// bracelet (a necklace class and its mirror), and pass 2 runs only the rows
// that are bracelet representatives.  At n = 14: 1182 necklaces, 687
// bracelets.
//
// Evaluation rows.  The (E class, O class) pairs of an orbit's states form a
// set closed under the one-site shift (a, b) -> (b, a), the mirror
// (a, b) -> (mirror a, mirror b) and, when folded, the spin flip.  The row R
// of an orbit is a function of that set (one of the stored classes among its
// O classes, picked by a balancing rule), and inside a row the orbits are
// ordered by a key that is also a function of the set (the least E class of
// its states with O class R).  The states (E = NR[c], O = rotl_k NR[q]) for
// all k share the pair set, so the orbits pass 1 of column c gathers for the
// O class q form one run of at most 8n orbits: the column map is stored as a
// 32-bit base per (column, O class) and an 8-bit offset per entry.
struct FusedTables {
  int n;                 // L/2
  int nNk;               // number of n-bit necklaces
  int nCol;              // stored columns (bracelet representatives)
  int nRow;              // rows pass 2 runs (bracelet representatives)
  // pass-1 output A, row-major: A[r][s][k] (row necklace r < nNk, stored
  // column s < nCol, slot k < 16) is A(E = NR[colS[s]], O = rotr_k NR[r])
  DevBuf slotO;          // uint16[nNk*16] LDS index sw(O) of (row necklace r, slot k),
                         //   O = rotr_k NR[r] (mod the period) for k < n; 0xFFFF unused
  DevBuf jtab;           // uint32[2^n]   the entries of a column (O words necklace-major:
                         //               class q, then rotation): sw(O) | q << 16, in
                         //               an interleaved position order (each block of
                         //               256 entries transposed as 64 x 4) so that a
                         //               gather instruction covers 64 consecutive ones
  DevBuf cbase;          // uint32[nCol*nNk] internal index of the first orbit of the run
                         //               that column s gathers for O class q
  DevBuf coff;           // uint8[nCol << n] offset of each entry in that run (same
                         //               position order as jtab)
  DevBuf itab;           // uint64[nNk]   pass-2 items: the nCol stored columns, then the
                         //   mirrored classes by source column; bits 0-15 source column,
                         //   16-31 NR of the item's class, 32-39 its period, 40-47 the
                         //   rotation m (255 for a stored class)
  DevBuf rinfo;          // uint64[nRow]  pass-2 rows in launch order (most orbits
                         //   first): bits 0-15 row necklace r, 16-31 its mirror class,
                         //   32-39 t with rev(NR[r]) = rotl_t NR[mirror], 40-55 the
                         //   row's index q into ostart
  DevBuf ostart;         // int64[nRow+1] internal-index range of the orbits whose
                         //               evaluation point has O = NR[rowS[q]]
  DevBuf olist;          // uint32[M]     internal index -> orbit (ascending-rep order)
  DevBuf Elist;          // uint16[M]     LDS index sw(E) of the E part of each orbit's
                         //               evaluation point
  std::vector<uint32_t> perm;  // host copy of olist
};
// The fused operator numbers orbits internally by row (olist order), so
// that pass 2 writes contiguously; vectors cross the API in the reference's
// ascending-representative order: xi[i] = x[perm[i]], y[perm[i]] = yi[i].
void permute_to_internal(const void *x, void *xi, const uint32_t *perm, int64_t M, tm_dtype io,
                         hipStream_t s);
void permute_to_external(const void *yi, void *y, const uint32_t *perm, int64_t M, tm_dtype io,
                         hipStream_t s);
void build_fused_tables(const GroupSpec &g, const DevBuf &reps, int64_t M, FusedTables &ft,
                        hipStream_t s);

// reduced-operator application
struct ReducedPlan {
  int64_t M;
  int L;
  const double *tvec;     // device [M] Tvec (sqrt(nf) exp(e/2) / sqrt(dup))
  const double *dt64;     // device [M] dup*Tvec
  const float *tvec32, *dt32;
  const uint64_t *reps;   // device [M]
  const uint32_t *map;    // device [2^L] (generic path) or null
  const FusedTables *ft;  // fused path tables or null
  double ka, kb;
  // full-space operator between expansion and gather (XDM, ZBM); null = the
  // Kronecker product of K = [[ka, kb], [kb, ka]] (XAM)
  const SweepPlan *sweep;
  int transpose;
  // > 0: symmetrised axial transfer over two-layer states of this layer
  // size (YAM, YBM): K2 = [[ka, kb], [kb, ka]] on the previous layer, the
  // layer swap, W from wtab (dev::w_layers)
  int ylayer;
  int ybm;
  const double *wtab64;
  const float *wtab32;
};
void launch_reduced(const ReducedPlan &plan, const void *in, void *out, tm_dtype io,
                    tm_dtype comp, void *work, const double *scale_in, hipStream_t s);
size_t reduced_work_bytes(const ReducedPlan &plan, tm_dtype comp);

// ---- Krylov vector kernels (tm_krylov.hip) ----
// Vector arguments of type T are basis vectors stored as double or float
// (instantiated for both); every product and sum is accumulated in FP64 and
// the scalar results (h, out, nrm) are double.
namespace kry {
constexpr int NBLK = 2048;  // partial-sum blocks of the reductions
// out[k] = <V_k, w> for k < j, out[j] = <w,w>; partial buffer of NBLK*(j+1).
// The sweep functions below return the number of partial blocks they wrote;
// with reduce = false they leave the final sum to one of the reduce_*
// functions (which also close the Krylov step in the same launch).
template <typename T>
unsigned multidot(const T *V, int64_t ldv, int j, const T *w, int64_t n, double *partial,
                  double *out, hipStream_t s, bool reduce = true);
// w -= sum_k h[k] V_k (k<j); then out[k] = <V_k, w>, out[j] = <w,w>
template <typename T>
unsigned update_dots(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                     double *partial, double *out, hipStream_t s, bool reduce = true);
// update_dots for h zero outside its last two slots (a Lanczos step): w is
// updated from V_{j-2}, V_{j-1} only, then all dots as update_dots
template <typename T>
unsigned update_dots_lanczos(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                             double *partial, double *out, hipStream_t s, bool reduce = true);
// w -= sum_k h[k] V_k; out[0] = <w,w>
template <typename T>
void update_norm(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                 double *partial, double *out, hipStream_t s);
// Sum the nb partial blocks of cnt slots into out, then close Krylov step j
// from the two projections hb1 (first) and out (second): projected-matrix
// column j (+ subdiagonal) into Hd (ld ldh), nrm[0] = beta^2 (Pythagoras after
// a correction), flags[0] += correction, flags[1] = first step whose beta fell
// below 1e-13 |h| or whose correction removed more than half of the norm
// again (must start < 0)
void reduce_step_close(const double *partial, unsigned nb, int cnt, double *out, int j,
                       const double *hb1, double *Hd, int ldh, double *nrm, double *flags,
                       hipStream_t s);
// Sum into out (out[0] = alpha = <v_j, T v_j>), then the Lanczos coefficients
// of step j: hL[0..j] = (row j of Hd left of the diagonal, alpha); the row
// holds beta_{j-1} inside a cycle and the restart's coupling at its first step
void reduce_lanczos_coeffs(const double *partial, unsigned nb, int cnt, double *out, int j,
                           double *Hd, int ldh, double *hL, hipStream_t s);
// Sum into out (the projections h2 of the three-term residual and its
// norm^2), then close the Lanczos step: column j = hL + h2, beta^2 =
// |w1|^2 - |h2|^2; flags as reduce_step_close
void reduce_step_close_lanczos(const double *partial, unsigned nb, int cnt, double *out, int j,
                               const double *hL, double *Hd, int ldh, double *nrm, double *flags,
                               hipStream_t s);
// Lagged Lanczos steps (symmetric T): a step leaves its three-term residual
// r in slot j+1 uncorrected and unnormalised; the next step applies T to r
// and needs one sweep for both the correction of r and its own residual.
// lag_close: from the partial sums of <r, z> (multidot of slot j against
// slot j+1 = T r, 2 slots), the previous correction g (j values), beta^2 =
// nrm[0] and the projected matrix Hd, the coefficients hL[0..j] and cf
// (2 * 72 + 2 doubles) of lanczos_lag.  lanczos_lag: v_j into slot j, the
// next residual into slot j+1, partial sums of the j+2 slots <V_k, w> (k <=
// j), <w, w> for reduce_step_close_lanczos.  j <= LAG_MAXJ.
constexpr int LAG_MAXJ = 24;
void lag_close(const double *partial, unsigned nb, int j, const double *g, const double *Hd,
               int ldh, const double *nrm, double *hL, double *cf, hipStream_t s);
template <typename T>
unsigned lanczos_lag(T *V, int64_t ldv, int j, int64_t n, const double *cf, double *partial,
                     hipStream_t s);
// General T: from the partial sums of the multidot of V_0..V_j (slot j the
// raw residual r) against T r (j+2 slots), the previous correction g, beta^2
// = nrm[0] and Hd, the first projection of T v_j into hL[0..j] and the
// coefficients cf of lanczos_lag.  j <= LAG_MAXJ.
void arnoldi_lag_close(const double *partial, unsigned nb, int j, const double *g,
                       const double *Hd, int ldh, const double *nrm, double *hL, double *cf,
                       hipStream_t s);
// w <- (w - c V h2) / sqrt(nrm2[0]), c = 1 iff force or !(w1 > w0 / 2)
template <typename T>
void update_normalize(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h2,
                      const double *w0, const double *w1, const double *nrm2, int force,
                      hipStream_t s);
// v = w * (1/sqrt(nrm2[0]))
template <typename T>
void normalize(T *w, int64_t n, const double *nrm2, hipStream_t s);
// Vout[:, c] = sum_r Vin[:, r] * Y[r + c*ldy], r < m, c < k  (Vout may alias
// Vin when both have one element type).  Instantiated for (double, double),
// (float, float) and (float, double).
template <typename Ti, typename To>
void gemm_vy(const Ti *Vin, int64_t ldv, int m, const double *Y, int ldy, int k, To *Vout,
             int64_t ldo, int64_t n, hipStream_t s);
template <typename T>
void random_uniform(T *v, int64_t n, uint64_t seed, hipStream_t s);
// b = a with element conversion (all four double/float pairs)
template <typename Ta, typename Tb>
void copy(const Ta *a, Tb *b, int64_t n, hipStream_t s);
void to_f32(const double *a, float *b, int64_t n, hipStream_t s);
void fill(double *a, double val, int64_t n, hipStream_t s);
// max|x| -> 2^-ceil(log2 max) into scale (device double)
void pow2_scale(const double *x, int64_t n, double *partial, double *scale, hipStream_t s);
// G[i*k + j] = <A_i, B_j>  (A, B are k column vectors)
void gram(const double *A, const double *B, int64_t ld, int k, int64_t n, double *partial,
          double *G, hipStream_t s);
// G[i*nv + j] = <U_i, U_j> for the nv column vectors of U (one sweep for nv <= 10)
void gram_sym(const double *U, int64_t ld, int nv, int64_t n, double *partial, double *G,
              hipStream_t s);
// out[c] = || sum_{r < nu} Cm[r + c*ldc] U_r ||^2 for c < nc, without forming
// the combinations (one sweep per 8 columns)
void combine_norms(const double *U, int64_t ld, int nu, const double *Cm, int ldc, int nc,
                   int64_t n, double *partial, double *out, hipStream_t s);
// ---- block Krylov (blocks of b <= 8 vectors) ----
// G[r + c * nx] = <X_r, Y_c> for r < nx <= 64, c < b <= 16, in one sweep on
// FP64 MFMA (X, Y columns of leading dimension ld); partial holds
// gram_xy_partial_doubles() doubles.
template <typename T>
void gram_xy(const T *X, int64_t ld, int nx, const T *Y, int b, int64_t n, double *partial,
             double *G, hipStream_t s);
size_t gram_xy_partial_doubles();
// W <- (W - X H) Rinv: H nx x b (leading dimension ldh, null: none), Rinv b x b
// upper triangular (null: identity); X and W distinct columns of one basis
template <typename T>
void block_update(const T *X, int64_t ld, int nx, T *W, int b, const double *H, int ldh,
                  const double *Rinv, int64_t n, hipStream_t s);
// close a block step on the device (mode 0: projected-matrix column block
// and R from the two projections; mode 1: CholQR's second pass); see
// k_block_close
void block_close(const double *G1, const double *G2, int nc, int b, double *Hd, int ldh, int col0,
                 double *Rinv, double *flags, int step, int mode, hipStream_t s);
}  // namespace kry

// ---------------------------------------------------------------------------
// small dense eigensolvers (tm_dense.cpp)
// ---------------------------------------------------------------------------
// Symmetric: A (n x n, column-major) -> eigenvalues w, eigenvectors Z (columns).
void sym_eig(int n, const double *A, double *w, double *Z);
// General real: eigenvalues (wr, wi) and eigenvectors as complex columns
// Zr/Zi (column j belongs to eigenvalue j; unit 2-norm).
void gen_eig(int n, const double *A, double *wr, double *wi, double *Zr, double *Zi);

}  // namespace ltm

namespace ltm {
// Caller-provided device memory for an apply that must not allocate.
struct Workspace {
  void *work = nullptr;  // the operator workspace
  size_t work_bytes = 0;
  void *px = nullptr, *py = nullptr;  // permutation buffers (perm_bytes each)
  size_t perm_bytes = 0;
};
}  // namespace ltm

// ---------------------------------------------------------------------------
// the operator object behind tm_op*
// ---------------------------------------------------------------------------
struct tm_op {
  tm_params p{};
  tm_kind kind{};
  tm_kind fullkind{};   // XDM, ZBM: the full-space circuit (XDT, ZBT) they reduce
  int L = 0, s1 = 0, s2 = 0;
  int bc = 0;           // boundary code (XAT/XDT) or extend flag (XAM)
  int censym = 0, rotate = 0, nleadbit = 0;
  int64_t nstate = 0, ncol = 0;
  double nf = 1.0;
  bool symmetric = true;
  int64_t matvecs = 0;
  hipStream_t stream = nullptr;
  double bzwj1[2]{}, bzwj2[2]{}, bzwj3[2]{};

  // diagonal tables (full-space operators)
  ltm::DiagSpec diag{};
  ltm::DevBuf tab64, tab32;
  ltm::DevBuf tabb64, tabb32;  // second table (Y2AT: the NN weight between layers)
  // reduced operators
  ltm::DevBuf reps, dups, map, tvec, dt64, tvec32, dt32;
  std::unique_ptr<ltm::FusedTables> ft;
  // internal orbit numbering (fused reduced operator); empty = identity.
  // perm[i] = reference (ascending-representative) index of internal i.
  std::vector<uint32_t> perm;
  // workspaces
  ltm::DevBuf work, one, hx, hy, px, py;
  // eigen-driver buffers and graphs kept between solves (created on demand)
  std::unique_ptr<ltm::SolverCache> scache;
  ltm::SolverCache &solver_cache()
  {
    if (!scache) scache.reset(new ltm::SolverCache);
    return *scache;
  }

  virtual ~tm_op();
  // gate-circuit plan of XDT/ZBT (or of the circuit a reduced XDM/ZBM wraps)
  ltm::SweepPlan sweep_plan() const;
  // Device workspace bytes of one apply with compute type comp (the
  // operator's `work`), and of the two permutation buffers of an apply in
  // the reference order (0 when the operator has no internal order).
  size_t work_bytes(tm_dtype comp) const;
  size_t perm_bytes(tm_dtype io) const;
  // T.x on vectors in the internal order (the eigensolver's basis).  ws
  // null: the operator's own workspace (grown on demand) and the tile
  // autotuner; ws given: that memory (its size checked), no allocation, no
  // tuning.
  void apply_internal(const void *x, void *y, int64_t ld, int nvec, tm_dtype io, tm_dtype comp,
                      int transpose, const double *scale_in, hipStream_t s,
                      const ltm::Workspace *ws = nullptr);
  // T.x on vectors in the reference order (the API); ws as above, its px /
  // py used for the permutation
  void apply_dev(const void *x, void *y, int64_t ld, int nvec, tm_dtype io, tm_dtype comp,
                 int transpose, const double *scale_in, hipStream_t s,
                 const ltm::Workspace *ws = nullptr);
};

namespace ltm {
// Operators over orbit representatives (the reduced matrices).
inline bool is_reduced(tm_kind k)
{
  return k == TM_KIND_XAM2D || k == TM_KIND_XAM3D || k == TM_KIND_XDM || k == TM_KIND_ZBM ||
         k == TM_KIND_YAM || k == TM_KIND_YBM;
}
// tm_op_apply_dev on the no-allocation entry point's scratch array
// (tm_noalloc.hip); throws Error.
void apply_dev_noalloc(tm_op *op, const void *x, void *y, int64_t ld, int nvec, tm_dtype dtype,
                       tm_dtype compute, int transpose, hipStream_t s);
// Construct an operator from parameters (pfunc dispatch); throws Error.
tm_op *make_op(const tm_params &p);
// Hellmann-Feynman numerators: num[i] = <l, dT/dtheta_i r> for theta =
// (betaJ, kappa, kappa2, h), num[4] = <l, r>.  l, r are device vectors.
void grad_numerators(tm_op *op, const double *l_dev, const double *r_dev, double num[5]);
// d ln(normfactor)/d theta for the 4 params.
void dlnnf(const tm_op *op, double out[4]);
// Observable kinds of an eigenvector pair (which weight, which bits carry
// the magnetisation, which extra observables).
enum PhysKind { PHYS_SYM2D, PHYS_SYM3D, PHYS_DNNI, PHYS_GEN, PHYS_GEN_LOW, PHYS_GEN_HIGH,
                PHYS_ZERO };
// sums[0] = sum w, sums[1..4] = sum w m^k, sums[5..8] = the kind's extra
// sums, over n device entries of ql (and qr), states reps[i] or i.
void phys_sums(int pk, const double *ql, const double *qr, const uint64_t *reps, int64_t n, int L,
               int bc, void *work, double sums[9], hipStream_t s);
size_t phys_work_bytes();
// Markov chains of configurations (XAT, YAT and the reduced XAM2D, XDM, ZBM,
// YAM, YBM): vec = |right eigenvector|, vecl = |left eigenvector| (null for
// the symmetric kinds; the start weight is then vec^2) and u = the uniforms,
// all device; states (nchain * length, device) receives the chains; *err
// (device) is set if an orbit's member count disagrees with its size.
void sample_chains(tm_op *op, const double *vec, const double *vecl, const double *u, int nchain,
                   int length, uint64_t *states, int *err);
}  // namespace ltm
