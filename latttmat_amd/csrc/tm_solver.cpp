// tm_solver.cpp -- the eigen driver (pfunc_inner / getlefteigvec of the
// reference, TMat/pfunc.hpp:39-114) with the Krylov restart loop on the GPU.
//
// Method: thick-restart Krylov-Schur (Stewart) for both symmetric and
// general operators.  The reference calls Spectra's implicitly restarted
// Lanczos / Arnoldi (SymEigsSolver / GenEigsSolver, pfunc.hpp:84-98) with
// nev = neigs, ncv = suggested_nev (tmat.hpp:88), LARGEST_MAGN and Spectra's
// defaults tol = 1e-10, maxit = 1000; a thick restart that keeps the wanted
// Ritz vectors spans the same restarted Krylov space as an implicit restart
// with the unwanted Ritz values as shifts.
//
// Device: the basis V (ncv+1 vectors), T.x, CGS2 orthogonalisation in lagged
// steps (a step's residual is corrected and normalised inside the next
// step's sweep: one sweep over V per step for symmetric T, two otherwise),
// the restart V <- V Y.  Host: the (ncv x ncv) projected matrix, its eigen-decomposition,
// convergence tests.  One small device->host copy and one upload per
// restart cycle.
//
// Mixed precision (TM_PREC_MIXED): phase 1 runs the same loop with the
// operator computing in FP32 and, by default, the basis stored in FP32
// (products and sums still accumulated in FP64) to a residual of mixed_tol
// or the FP32 noise floor; phase 2 refines the kept Ritz block with FP64
// matvecs (Rayleigh-Ritz, repeated with one subspace-iteration step) until
// the FP64 residuals or the eigenvalue error bound meet the target.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <numeric>
#include <vector>

#include "tm_internal.hpp"

namespace ltm {
namespace {

constexpr double EPS = 2.220446049250313e-16;

// Operator-cache slots: KrylovSchur role r uses dev[SLOT_ROLE * r + 0..4]
// and pinned[r] (r = 0 main run, 1 FP64 continuation of a mixed run, 2 left
// vector); the refinement block and its scratch, and the Ritz vectors handed
// to the caller, have slots of their own.
constexpr int SLOT_ROLE = 6;
// device projected-matrix buffer: (m+1) m entries, then flags[0..1] and the
// step norm (8 slots), the last step's correction (72 slots), then the
// restart rotation at offset (m+1) m + HD_Y
constexpr int HD_Y = 8 + 72;
constexpr int SLOT_REFINE_B = 18, SLOT_REFINE_PARTIAL = 19, SLOT_REFINE_G = 20,
              SLOT_REFINE_C = 21, SLOT_XR = 22, SLOT_XI = 23, SLOT_CHECK = 24;
// block Krylov-Schur (block runs use role 0's basis slot)
constexpr int SLOT_BK_PART = 25, SLOT_BK_G = 26, SLOT_BK_HD = 27, SLOT_BK_Y = 28;
static_assert(SLOT_BK_Y < SolverCache::NDEV, "operator cache slots");

// A projected-matrix entry read back from the device.  A non-finite value
// means the device results are invalid (an operator or Krylov kernel
// produced garbage); the solve stops with an error instead of turning it
// into eigenvalues.
double finite_or_throw(double v)
{
  if (!std::isfinite(v))
    throw Error(TM_EHIP, "non-finite value in the Krylov projection: device results are invalid");
  return v;
}

struct Ritz {
  std::vector<double> wr, wi;     // eigenvalues (sorted by magnitude, descending)
  std::vector<double> Zr, Zi;     // eigenvectors of the projected matrix (m x m)
  std::vector<double> res;        // Krylov residual estimates
};

// What solve_impl needs of a Krylov-Schur run, whatever the basis storage.
class KrylovRun {
 public:
  virtual ~KrylovRun() = default;
  virtual void set_graph(bool on) = 0;
  virtual void set_stall_stop(bool v) = 0;
  // run to convergence from a random start (v0 null) or from the FP64
  // device vector v0; returns the number of converged wanted pairs
  virtual int run(Ritz &out, const double *v0) = 0;
  // FP64 Ritz vectors (real parts; imaginary parts into Xi if non-null)
  virtual void ritz_vectors(const Ritz &r, int cnt, double *Xr, double *Xi) = 0;
  virtual int64_t graph_cycles() const = 0;
  virtual int64_t ldv() const = 0;
  virtual int64_t matvecs() const = 0;
  virtual int64_t restarts() const = 0;
  virtual bool stalled() const = 0;
  virtual int m() const = 0;
};

// T = element type of the stored basis (double, or float for the FP32
// phase of a mixed solve; the operator then reads and writes FP32 vectors).
template <typename T>
class KrylovSchur final : public KrylovRun {
 public:
  // role (0, 1, 2) selects the operator's cached buffers, so that runs that
  // are alive at the same time (FP32 phase and FP64 continuation, right and
  // left solves) never share them.
  KrylovSchur(tm_op *op, int nev, int m, double tol, int maxit, tm_dtype comp, int transpose,
              uint64_t seed, int verbose, int role)
      : op_(op), nev_(nev), m_(m), tol_(tol), maxit_(maxit), comp_(comp), tr_(transpose),
        seed_(seed), verbose_(verbose), cache_(op->solver_cache()),
        V_(cache_.dev[SLOT_ROLE * role + 0]), partial_(cache_.dev[SLOT_ROLE * role + 1]),
        hbuf_(cache_.dev[SLOT_ROLE * role + 2]), Ydev_(cache_.dev[SLOT_ROLE * role + 3]),
        Hd_(cache_.dev[SLOT_ROLE * role + 4]), stage_(cache_.pinned[role])
  {
    n_ = op->ncol;
    ldv_ = ((n_ + 63) / 64) * 64;
    sym_ = op->symmetric;
    V_.ensure(sizeof(T) * ldv_ * (m_ + 1));
    partial_.ensure(sizeof(double) * kry::NBLK * 72);
    hbuf_.ensure(sizeof(double) * 6 * 72);
    Ydev_.ensure(sizeof(double) * 72 * 72);
    H_.assign((size_t)(m_ + 1) * m_, 0.0);
    // projected matrix, flags and norm (8), the last step's correction (72),
    // the restart rotation (up to 72 x 72)
    Hd_.ensure(sizeof(double) * ((size_t)(m_ + 1) * m_ + HD_Y + 72 * 72));
    // read-back of the projected matrix, then the restart's upload
    stage_.ensure(sizeof(double) * (72 * 72 + (size_t)(m_ + 1) * m_ + HD_Y + 72 * 72));
  }

  KrylovSchur(const KrylovSchur &) = delete;
  KrylovSchur &operator=(const KrylovSchur &) = delete;
  void set_graph(bool on) override { use_graph_ = on; }
  int64_t graph_cycles() const override { return graph_cycles_; }

  T *V(int k) const { return V_.as<T>() + (size_t)k * ldv_; }
  int64_t ldv() const override { return ldv_; }
  int64_t n() const { return n_; }
  double &H(int i, int j) { return H_[i + (size_t)j * (m_ + 1)]; }

  // Start vector: uniform(0,1) from the seed, normalised (pfunc.hpp:97).
  void start()
  {
    hipStream_t s = op_->stream;
    kry::random_uniform(V(0), n_, seed_, s);
    kry::multidot(V_.as<T>(), ldv_, 0, V(0), n_, partial_.as<double>(), hbuf_.as<double>(), s);
    kry::normalize(V(0), n_, hbuf_.as<double>(), s);
    clear_hd();
    k_ = 0;
  }

  // Start from a given FP64 device vector (normalised here).
  void start_from(const double *v0)
  {
    hipStream_t s = op_->stream;
    kry::copy(v0, V(0), n_, s);
    kry::multidot(V_.as<T>(), ldv_, 0, V(0), n_, partial_.as<double>(), hbuf_.as<double>(), s);
    kry::normalize(V(0), n_, hbuf_.as<double>(), s);
    clear_hd();
    k_ = 0;
  }

  // The lagged Lanczos steps read whole leading blocks of the device
  // projected matrix (entries below the subdiagonal included), so a solve
  // starts from zeros there and every restart uploads the restarted matrix.
  void clear_hd()
  {
    double *Hd = Hd_.as<double>();
    kry::fill(Hd, 0.0, (int64_t)(m_ + 1) * m_ + 2, op_->stream);
    kry::fill(Hd + (size_t)(m_ + 1) * m_ + 1, -1.0, 1, op_->stream);
    raw_m_ = false;
  }

  // Expand the Krylov-Schur decomposition from column k_ to m_.
  void expand()
  {
    if (device_steps_) expand_device();
    else expand_host(k_);
  }

  // The whole cycle without host synchronisation: each step's correction
  // decision, projected-matrix column and norm stay on the device
  // (kry::reduce_step_close*); the host reads the m x m block once at the end.  An
  // invariant subspace found at step jb is redone from jb by the host-driven
  // loop below (V(0..jb) are untouched by the steps after it).
  void expand_device()
  {
    hipStream_t s = op_->stream;
    const int ldh = m_ + 1;
    double *Hd = Hd_.as<double>();
    const int k0 = k_;
    // Cycles after the first (k0 > 0, buffers already allocated) are
    // captured once per k0 into a graph and replayed.
    // The graph is kept on the operator and reused by later solves whose
    // launches would carry the same arguments.
    hipGraphExec_t ex = nullptr;
    if (use_graph_ && k0 > 0) {
      const std::array<uint64_t, 12> key{
          (uint64_t)sizeof(T), (uint64_t)comp_, (uint64_t)tr_, (uint64_t)m_, (uint64_t)k0,
          (uint64_t)n_, (uint64_t)ldv_, (uint64_t)(uintptr_t)V_.get(),
          (uint64_t)(uintptr_t)partial_.get(), (uint64_t)(uintptr_t)hbuf_.get(),
          (uint64_t)(uintptr_t)Hd_.get(), (uint64_t)(uintptr_t)op_->work.get()};
      for (const auto &g : cache_.graphs)
        if (g.key == key) ex = g.exec;
      if (!ex) {
        ex = capture_cycle(k0);
        if (ex) cache_.graphs.push_back({key, ex});
        else use_graph_ = false;
      }
    }
    if (ex) {
      TM_HIP(hipGraphLaunch(ex, s));
      op_->matvecs += m_ - k0;
      ++graph_cycles_;
    } else {
      issue_cycle(k0);
    }
    matvecs_ += m_ - k0;
    double *hd = stage_.as<double>();  // (m+1) m + 8 + m <= 72 * 72 doubles
    TM_HIP(hipMemcpyAsync(hd, Hd, sizeof(double) * ((size_t)ldh * m_ + 8 + m_),
                          hipMemcpyDeviceToHost, s));
    stream_sync_poll(s);
    const double *fl = hd + (size_t)ldh * m_;
    reorths_ += (int64_t)fl[0];
    const int jb = fl[1] < 0 ? m_ : (int)fl[1];
    for (int j = k0; j < jb; ++j)
      for (int i = 0; i <= j + 1; ++i) H(i, j) = finite_or_throw(hd[i + (size_t)j * ldh]);
    // a device cycle ends with slot m holding the last residual before its
    // correction and normalisation; the restart applies them
    raw_m_ = jb == m_;
    if (raw_m_)
      for (int i = 0; i < m_; ++i) gm_[i] = finite_or_throw(fl[8 + i]);
    if (jb < m_) {
      matvecs_ -= m_ - jb;
      expand_host(jb);
    }
    k_ = m_;
  }

  // The kernels of one cycle from column k0 (no host synchronisation, no
  // allocation once the operator's workspace exists).
  void issue_cycle(int k0)
  {
    hipStream_t s = op_->stream;
    double *hb = hbuf_.as<double>();
    const int ldh = m_ + 1;
    double *Hd = Hd_.as<double>();
    double *flags = Hd + (size_t)ldh * m_;
    double *nrm = flags + 2;
    // flags[0] = 0 and flags[1] = -1 on entry: set by the solve's start and
    // by every restart's upload
    // lag: slot j holds the previous step's raw three-term residual r (its
    // second projection at hb + 72, beta^2 in nrm), not yet v_j
    bool lag = false;
    double *cf = hb + 216;
    for (int j = k0; j < m_; ++j) {
      T *w = V(j + 1);
      const int nk = j + 1;
      double *part = partial_.as<double>();
      if (lag && j > kry::LAG_MAXJ) {
        // beyond the lagged kernel's register budget: finish v_j first
        kry::update_normalize(V_.as<T>(), ldv_, j, V(j), n_, hb + 72, hb, hb, nrm, 1, s);
        lag = false;
      }
      op_->apply_internal(V(j), w, ldv_, 1, io_, comp_, tr_, nullptr, s);
      if (lag) {
        // lagged step: T r is in slot j+1; one sweep finishes v_j and forms
        // the next residual with its reorthogonalisation dots.  Its first
        // projection: for symmetric T from <r, T r> alone, otherwise from a
        // sweep of the basis against T r
        unsigned nb;
        if (sym_) {
          nb = kry::multidot(V(j), ldv_, 1, w, n_, part, hb, s, false);
          kry::lag_close(part, nb, j, hb + 72, Hd, ldh, nrm, hb + 144, cf, s);
        } else {
          nb = kry::multidot(V_.as<T>(), ldv_, nk, w, n_, part, hb, s, false);
          kry::arnoldi_lag_close(part, nb, j, hb + 72, Hd, ldh, nrm, hb + 144, cf, s);
        }
        nb = kry::lanczos_lag(V_.as<T>(), ldv_, j, n_, cf, part, s);
        kry::reduce_step_close_lanczos(part, nb, nk + 1, hb + 72, j, hb + 144, Hd, ldh, nrm, flags,
                                       s);
        continue;
      }
      if (sym_) {
        // Lanczos step: the first projection is known but for alpha -- T v_j
        // couples to v_{j-1} inside a cycle and, at the first step of a
        // cycle, to every kept Ritz vector (the restart's coupling row) --
        // and is subtracted inside the reorthogonalisation sweep; the
        // correction and normalisation of the residual are left to the next
        // step's sweep (lagged) or, after the last step, to the restart
        // (each reduction also closes its part of the step in the same launch)
        unsigned nb = kry::multidot(V(j), ldv_, 1, w, n_, part, hb, s, false);
        kry::reduce_lanczos_coeffs(part, nb, 2, hb, j, Hd, ldh, hb + 144, s);
        if (j > k0)
          nb = kry::update_dots_lanczos(V_.as<T>(), ldv_, nk, w, n_, hb + 144, part, hb + 72, s,
                                        false);
        else
          nb = kry::update_dots(V_.as<T>(), ldv_, nk, w, n_, hb + 144, part, hb + 72, s, false);
        kry::reduce_step_close_lanczos(part, nb, nk + 1, hb + 72, j, hb + 144, Hd, ldh, nrm, flags,
                                       s);
        lag = true;
        continue;
      }
      // general step: classical Gram-Schmidt, the second projection
      // measured in the same sweep; the correction and normalisation are left
      // to the next (lagged) step's sweep or the restart (an unsafe
      // correction sends the step to the host loop)
      kry::multidot(V_.as<T>(), ldv_, nk, w, n_, part, hb, s);
      const unsigned nb = kry::update_dots(V_.as<T>(), ldv_, nk, w, n_, hb, part, hb + 72, s,
                                           false);
      kry::reduce_step_close_lanczos(part, nb, nk + 1, hb + 72, j, hb, Hd, ldh, nrm, flags, s);
      lag = true;
    }
    // the last step's correction next to the flags, for the restart
    if (lag) kry::copy(hb + 72, flags + 8, m_, s);
  }

  // Record issue_cycle(k0) into an executable graph; nullptr if the runtime
  // refuses (the caller then stays eager).
  hipGraphExec_t capture_cycle(int k0)
  {
    CaptureGuard guard;  // no allocation or free on any thread meanwhile
    hipStream_t s = op_->stream;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    const int64_t mv = op_->matvecs;
    bool ok = true;
    try {
      issue_cycle(k0);
    } catch (...) {
      ok = false;
    }
    op_->matvecs = mv;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    hipGraphExec_t ex = nullptr;
    if (ok && e == hipSuccess && g) {
      if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) ex = nullptr;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return ex;
  }

  // Host-driven expansion from column j0 (one device->host copy per step).
  void expand_host(int j0)
  {
    hipStream_t s = op_->stream;
    double *hb = hbuf_.as<double>();
    std::vector<double> hh(3 * 72);
    for (int j = j0; j < m_; ++j) {
      T *w = V(j + 1);
      op_->apply_internal(V(j), w, ldv_, 1, io_, comp_, tr_, nullptr, s);
      ++matvecs_;
      const int nk = j + 1;
      // classical Gram-Schmidt with one DGKS correction (only when the first
      // projection removed more than half of the norm^2)
      kry::multidot(V_.as<T>(), ldv_, nk, w, n_, partial_.as<double>(), hb, s);
      kry::update_dots(V_.as<T>(), ldv_, nk, w, n_, hb, partial_.as<double>(), hb + 72, s);
      TM_HIP(hipMemcpyAsync(hh.data(), hb, sizeof(double) * 2 * 72, hipMemcpyDeviceToHost, s));
      stream_sync_poll(s);
      const double w0 = hh[nk], w1 = hh[72 + nk];
      const double *nrm_dev = hb + 72 + nk;
      double beta2 = w1;
      const bool reorth = !(w1 > 0.5 * w0);
      for (int kk = 0; kk < nk; ++kk) H(kk, j) = hh[kk];
      if (reorth) {
        kry::update_norm(V_.as<T>(), ldv_, nk, w, n_, hb + 72, partial_.as<double>(), hb + 144, s);
        TM_HIP(hipMemcpyAsync(hh.data() + 144, hb + 144, sizeof(double), hipMemcpyDeviceToHost, s));
        stream_sync_poll(s);
        for (int kk = 0; kk < nk; ++kk) H(kk, j) += hh[72 + kk];
        beta2 = hh[144];
        nrm_dev = hb + 144;
        ++reorths_;
      }
      double hn2 = 0;
      for (int kk = 0; kk < nk; ++kk) hn2 += finite_or_throw(H(kk, j)) * H(kk, j);
      double beta = std::sqrt(std::max(finite_or_throw(beta2), 0.0));
      if (beta <= 1e-13 * std::sqrt(hn2) || !(beta > 0)) {
        // invariant subspace: continue with a fresh random direction
        // orthogonal to the basis; the coupling is zero.
        kry::random_uniform(w, n_, seed_ + 7919 * (uint64_t)(j + 1 + 31 * restarts_), s);
        for (int pass = 0; pass < 2; ++pass) {
          kry::multidot(V_.as<T>(), ldv_, nk, w, n_, partial_.as<double>(), hb, s);
          kry::update_norm(V_.as<T>(), ldv_, nk, w, n_, hb, partial_.as<double>(), hb + 144, s);
        }
        kry::normalize(w, n_, hb + 144, s);
        beta = 0.0;
      } else {
        kry::normalize(w, n_, nrm_dev, s);
      }
      H(j + 1, j) = beta;
    }
    k_ = m_;
  }

  // Eigen-decomposition of the m x m projected matrix, sorted by |lambda|.
  Ritz ritz()
  {
    const int m = m_;
    std::vector<double> A((size_t)m * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) A[i + (size_t)j * m] = H(i, j);
    Ritz r;
    std::vector<double> wr(m, 0.0), wi(m, 0.0), Zr((size_t)m * m), Zi((size_t)m * m, 0.0);
    if (sym_) {
      sym_eig(m, A.data(), wr.data(), Zr.data());
    } else {
      gen_eig(m, A.data(), wr.data(), wi.data(), Zr.data(), Zi.data());
    }
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      const double ma = std::hypot(wr[a], wi[a]), mb = std::hypot(wr[b], wi[b]);
      if (ma != mb) return ma > mb;
      return wi[a] > wi[b];  // conjugate pairs: positive imaginary part first
    });
    r.wr.resize(m);
    r.wi.resize(m);
    r.Zr.resize((size_t)m * m);
    r.Zi.resize((size_t)m * m);
    r.res.resize(m);
    // gen_eig returns each conjugate pair with exactly opposite imaginary
    // parts; order every pair (+, -) next to each other
    if (!sym_) {
      for (int c = 0; c + 1 < m; ++c) {
        const int o = idx[c];
        if (wi[o] == 0) continue;
        int best = -1;
        for (int d = c + 1; d < m; ++d) {
          const int od = idx[d];
          if (wi[od] != 0 && std::fabs(wr[od] - wr[o]) <= 1e-12 * std::fabs(wr[o]) + 1e-300 &&
              std::fabs(wi[od] + wi[o]) <= 1e-12 * std::fabs(wi[o])) { best = d; break; }
        }
        if (best < 0) continue;
        std::swap(idx[c + 1], idx[best]);
        if (wi[idx[c]] < 0) std::swap(idx[c], idx[c + 1]);
        // exact conjugate pairing
        wr[idx[c + 1]] = wr[idx[c]];
        wi[idx[c + 1]] = -wi[idx[c]];
        ++c;
      }
    }
    const double beta = std::fabs(H(m, m - 1));
    for (int c = 0; c < m; ++c) {
      const int o = idx[c];
      r.wr[c] = wr[o];
      r.wi[c] = wi[o];
      for (int i = 0; i < m; ++i) {
        r.Zr[i + (size_t)c * m] = Zr[i + (size_t)o * m];
        r.Zi[i + (size_t)c * m] = Zi[i + (size_t)o * m];
      }
      r.res[c] = beta * std::hypot(r.Zr[(m - 1) + (size_t)c * m], r.Zi[(m - 1) + (size_t)c * m]);
    }
    return r;
  }

  int nconv(const Ritz &r, double tol) const
  {
    int nc = 0;
    const double e23 = std::pow(EPS, 2.0 / 3.0);
    for (int i = 0; i < nev_; ++i) {
      const double mag = std::hypot(r.wr[i], r.wi[i]);
      if (r.res[i] <= tol * std::max(e23, mag)) ++nc;
    }
    return nc;
  }

  // Real orthonormal basis (m x k) of the first k Ritz vectors; conjugate
  // pairs contribute their real and imaginary parts.  Returns k used.
  int ritz_basis(const Ritz &r, int k, std::vector<double> &Y) const
  {
    const int m = m_;
    // a conjugate pair is (i, i+1) with wi[i] != 0 and wi[i+1] == -wi[i]
    auto pair_at = [&](int i) {
      return !sym_ && i + 1 < m && r.wi[i] != 0 && r.wi[i + 1] == -r.wi[i] &&
             r.wr[i + 1] == r.wr[i];
    };
    if (!sym_) {
      // keep conjugate pairs together: if column k-1 opens a pair, widen
      // (or shrink) the kept block by one
      int i = 0;
      while (i < k) {
        if (pair_at(i)) {
          if (i + 1 == k) { k = (k + 1 <= m - 1) ? k + 1 : k - 1; break; }
          i += 2;
        } else {
          ++i;
        }
      }
    }
    Y.assign((size_t)m * k, 0.0);
    int c = 0;
    for (int i = 0; i < k && c < k;) {
      for (int t = 0; t < m; ++t) Y[t + (size_t)c * m] = r.Zr[t + (size_t)i * m];
      ++c;
      if (pair_at(i) && c < k) {
        for (int t = 0; t < m; ++t) Y[t + (size_t)c * m] = r.Zi[t + (size_t)i * m];
        ++c;
        i += 2;
      } else {
        ++i;
      }
    }
    k = c;
    // modified Gram-Schmidt, twice
    for (int pass = 0; pass < 2; ++pass)
      for (int a = 0; a < k; ++a) {
        for (int b = 0; b < a; ++b) {
          double d = 0;
          for (int t = 0; t < m; ++t) d += Y[t + (size_t)a * m] * Y[t + (size_t)b * m];
          for (int t = 0; t < m; ++t) Y[t + (size_t)a * m] -= d * Y[t + (size_t)b * m];
        }
        double nn = 0;
        for (int t = 0; t < m; ++t) nn += Y[t + (size_t)a * m] * Y[t + (size_t)a * m];
        nn = std::sqrt(nn);
        for (int t = 0; t < m; ++t) Y[t + (size_t)a * m] /= nn;
      }
    return k;
  }

  // Thick restart keeping k Ritz vectors.
  void restart(const Ritz &r, int k)
  {
    const int m = m_;
    std::vector<double> Y;
    k = ritz_basis(r, k, Y);
    // S = Y^T H Y,  b = beta * Y(m-1, :)
    std::vector<double> HY((size_t)m * k, 0.0), S((size_t)k * k, 0.0);
    for (int c = 0; c < k; ++c)
      for (int i = 0; i < m; ++i) {
        double s = 0;
        for (int t = 0; t < m; ++t) s += H(i, t) * Y[t + (size_t)c * m];
        HY[i + (size_t)c * m] = s;
      }
    for (int c = 0; c < k; ++c)
      for (int a = 0; a < k; ++a) {
        double s = 0;
        for (int t = 0; t < m; ++t) s += Y[t + (size_t)a * m] * HY[t + (size_t)c * m];
        S[a + (size_t)c * k] = s;
      }
    const double beta = H(m, m - 1);
    if (verbose_ > 1) {
      double e = 0, hn = 0;
      for (int c = 0; c < k; ++c)
        for (int i = 0; i < m; ++i) {
          double t = HY[i + (size_t)c * m];
          for (int a = 0; a < k; ++a) t -= Y[i + (size_t)a * m] * S[a + (size_t)c * k];
          e += t * t;
          hn += HY[i + (size_t)c * m] * HY[i + (size_t)c * m];
        }
      std::fprintf(stderr, "[tm_solve] restart k=%d invariance error %.3e\n", k, std::sqrt(e / hn));
      if (verbose_ > 2) {
        std::fprintf(stderr, "HMAT");
        for (int j = 0; j < m; ++j)
          for (int i = 0; i < m; ++i) std::fprintf(stderr, " %.17g", H(i, j));
        std::fprintf(stderr, "\nWR");
        for (int i = 0; i < m; ++i) std::fprintf(stderr, " %.17g %.17g", r.wr[i], r.wi[i]);
        std::fprintf(stderr, "\nZR");
        for (int i = 0; i < m * m; ++i) std::fprintf(stderr, " %.17g %.17g", r.Zr[i], r.Zi[i]);
        std::fprintf(stderr, "\n");
      }
    }
    hipStream_t s = op_->stream;
    std::fill(H_.begin(), H_.end(), 0.0);
    for (int c = 0; c < k; ++c) {
      for (int a = 0; a < k; ++a) H(a, c) = S[a + (size_t)c * k];
      H(k, c) = beta * Y[(m - 1) + (size_t)c * m];
    }
    // one upload: the restarted projected matrix (zeros beyond column k),
    // the step flags reset (0, -1), and the rotation
    double *up = stage_.as<double>() + 72 * 72;
    const size_t nh = H_.size();
    std::copy(H_.begin(), H_.end(), up);
    std::fill(up + nh, up + nh + HD_Y, 0.0);
    up[nh + 1] = -1.0;
    double *yu = up + nh + HD_Y;
    const double *yd = Hd_.as<double>() + nh + HD_Y;
    int mr = m, kr = k;
    if (raw_m_) {
      // slot m holds r with v_m = (r - V g) / beta: one rotation of the m+1
      // columns by [[Y, -g / beta], [0, 1 / beta]] gives the kept Ritz
      // vectors and v_m in slot k
      TM_CHECK(beta > 0, TM_EHIP, "restart of a symmetric cycle with a zero residual");
      mr = m + 1;
      kr = k + 1;
      for (int c = 0; c < k; ++c) {
        for (int t = 0; t < m; ++t) yu[t + (size_t)c * mr] = Y[t + (size_t)c * m];
        yu[m + (size_t)c * mr] = 0.0;
      }
      for (int t = 0; t < m; ++t) yu[t + (size_t)k * mr] = -gm_[t] / beta;
      yu[m + (size_t)k * mr] = 1.0 / beta;
    } else {
      std::copy(Y.begin(), Y.begin() + (size_t)m * k, yu);
    }
    TM_HIP(hipMemcpyAsync(Hd_.get(), up, sizeof(double) * (nh + HD_Y + (size_t)mr * kr),
                          hipMemcpyHostToDevice, s));
    kry::gemm_vy(V(0), ldv_, mr, yd, mr, kr, V(0), ldv_, n_, s);
    if (!raw_m_) kry::copy(V(m), V(k), n_, s);
    raw_m_ = false;
    k_ = k;
    ++restarts_;
    // no synchronisation: the next cycle is ordered after V Y on the stream,
    // and the staging buffer is next written by a stream-ordered read-back
  }

  // Run to convergence; returns the number of converged wanted pairs.
  int run(Ritz &out, const double *v0) override
  {
    if (v0) start_from(v0); else start();
    int nc = 0;
    double best = 1e300;
    for (int it = 0; it < maxit_; ++it) {
      expand();
      out = ritz();
      nc = nconv(out, tol_);
      double worst = 0;
      for (int i = 0; i < nev_; ++i)
        worst = std::max(worst, out.res[i] / std::max(std::hypot(out.wr[i], out.wi[i]), 1e-300));
      if (verbose_)
        std::fprintf(stderr, "[tm_solve] cycle %d matvecs %lld lambda0 %.15g res0 %.3e worst %.3e nconv %d\n",
                     it, (long long)matvecs_, out.wr[0], out.res[0], worst, nc);
      if (nc >= nev_) break;
      // reduced-precision phase: stop once the residual has reached the
      // noise floor of the FP32 operator (no 30% improvement in a cycle)
      if (stall_stop_ && worst < 1e-4 && worst > 0.7 * best) {
        stalled_ = true;
        break;
      }
      best = std::min(best, worst);
      int k = nev_ + std::min(nc, (m_ - nev_) / 2);
      k = std::max(k, (nev_ + m_) / 2);
      k = std::min(k, m_ - 1);
      restart(out, k);
    }
    return nc;
  }

  // Ritz vectors X[:, c] = V(0..m-1) * Z[:, c] for c < cnt (real part, and
  // imaginary part into Xi if non-null) into device buffers of ld ldv().
  void ritz_vectors(const Ritz &r, int cnt, double *Xr, double *Xi) override
  {
    const int m = m_;
    hipStream_t s = op_->stream;
    std::vector<double> Z((size_t)m * cnt);
    for (int c = 0; c < cnt; ++c)
      for (int t = 0; t < m; ++t) Z[t + (size_t)c * m] = r.Zr[t + (size_t)c * m];
    TM_HIP(hipMemcpyAsync(Ydev_.get(), Z.data(), sizeof(double) * m * cnt, hipMemcpyHostToDevice, s));
    kry::gemm_vy(V(0), ldv_, m, Ydev_.as<double>(), m, cnt, Xr, ldv_, n_, s);
    stream_sync_poll(s);
    if (Xi) {
      for (int c = 0; c < cnt; ++c)
        for (int t = 0; t < m; ++t) Z[t + (size_t)c * m] = r.Zi[t + (size_t)c * m];
      TM_HIP(hipMemcpyAsync(Ydev_.get(), Z.data(), sizeof(double) * m * cnt, hipMemcpyHostToDevice, s));
      kry::gemm_vy(V(0), ldv_, m, Ydev_.as<double>(), m, cnt, Xi, ldv_, n_, s);
      stream_sync_poll(s);
    }
  }

  int64_t matvecs() const override { return matvecs_; }
  int64_t restarts() const override { return restarts_; }
  void set_stall_stop(bool v) override { stall_stop_ = v; }
  int64_t reorths() const { return reorths_; }
  bool stalled() const override { return stalled_; }
  bool sym() const { return sym_; }
  int m() const override { return m_; }

 private:
  tm_op *op_;
  int nev_, m_;
  double tol_;
  int maxit_;
  tm_dtype comp_;
  // vectors handed to the operator are the stored basis columns
  const tm_dtype io_ = sizeof(T) == 4 ? TM_F32 : TM_F64;
  int tr_;
  uint64_t seed_;
  int verbose_;
  int64_t n_ = 0, ldv_ = 0;
  bool device_steps_ = true;
  bool use_graph_ = true;
  int64_t graph_cycles_ = 0;
  SolverCache &cache_;
  DevBuf &V_, &partial_, &hbuf_, &Ydev_, &Hd_;
  PinnedBuf &stage_;
  // slot m after a device cycle of a symmetric run: the last residual
  // before its correction gm_ (m values) and normalisation by H(m, m-1)
  bool raw_m_ = false;
  double gm_[72] = {};
  bool sym_ = true;
  int k_ = 0;
  int64_t matvecs_ = 0, restarts_ = 0;
  bool stall_stop_ = false, stalled_ = false;
  int64_t reorths_ = 0;
  std::vector<double> H_;
};

// Block Krylov-Schur for a symmetric operator: b vectors per step.  The
// basis V (columns of leading dimension ldv; element type T as above) grows
// by blocks: the active block V_A = V[nc-b, nc) is applied, W = T V_A lands
// in V[nc, nc+b), and W is orthogonalised against V[0, nc) by block
// classical Gram-Schmidt twice -- the Gram [V | W]^T W of each pass in one
// sweep on FP64 MFMA (kry::gram_xy), the update W -= V h in one sweep -- then
// orthonormalised by CholQR twice, with the small factorisations on the
// device (kry::block_close), so that a cycle needs no host synchronisation
// until its end.  The projected matrix H = V^T T V is block tridiagonal plus
// the arrow of the kept Ritz vectors; after a cycle of projected size m the
// decomposition is T V_m = V_m H_m + V_res R E^T, and the thick restart keeps
// k Ritz vectors plus the residual block V_res.  Each pass reads the basis
// once for b vectors, where the single-vector run reads it twice per vector.
template <typename T>
class BlockKrylovSchur final : public KrylovRun {
 public:
  BlockKrylovSchur(tm_op *op, int nev, int ncv, int b, double tol, int maxit, tm_dtype comp,
                   uint64_t seed, int verbose)
      : op_(op), nev_(nev), ncv_(ncv), b_(b), tol_(tol), maxit_(maxit), comp_(comp), seed_(seed),
        verbose_(verbose), cache_(op->solver_cache()), V_(cache_.dev[0]),
        part_(cache_.dev[SLOT_BK_PART]), G_(cache_.dev[SLOT_BK_G]), Hd_(cache_.dev[SLOT_BK_HD]),
        Yd_(cache_.dev[SLOT_BK_Y]), stage_(cache_.pinned[0])
  {
    TM_CHECK(b >= 2 && b <= 8, TM_EINVAL, "block size must be in [2, 8]");
    TM_CHECK(ncv >= nev + b, TM_EINVAL, "block Krylov needs ncv >= nev + block size");
    n_ = op->ncol;
    ldv_ = ((n_ + 63) / 64) * 64;
    mmax_ = ncv + b;          // largest projected size
    ncap_ = mmax_ + b;        // basis columns
    TM_CHECK(ncap_ <= 64, TM_EINVAL, "block Krylov basis above 64 vectors");
    ldh_ = ncap_;
    V_.ensure(sizeof(T) * ldv_ * ncap_);
    part_.ensure(sizeof(double) * kry::gram_xy_partial_doubles());
    // G1, G2, G3 (64 x 16 each), Rinv (8 x 8)
    G_.ensure(sizeof(double) * (3 * 64 * 16 + 64));
    Hd_.ensure(sizeof(double) * ((size_t)ldh_ * mmax_ + 8));
    Yd_.ensure(sizeof(double) * 64 * 64);
    stage_.ensure(sizeof(double) * 72 * 72);
    H_.assign((size_t)ldh_ * mmax_, 0.0);
  }
  void set_graph(bool) override {}
  void set_stall_stop(bool v) override { stall_stop_ = v; }
  int64_t graph_cycles() const override { return 0; }
  int64_t ldv() const override { return ldv_; }
  int64_t matvecs() const override { return matvecs_; }
  int64_t restarts() const override { return restarts_; }
  bool stalled() const override { return stalled_; }
  int m() const override { return m_; }

  T *V(int k) const { return V_.as<T>() + (size_t)k * ldv_; }
  double &H(int i, int j) { return H_[i + (size_t)j * ldh_]; }
  double *G1() const { return G_.as<double>(); }
  double *G2() const { return G_.as<double>() + 64 * 16; }
  double *G3() const { return G_.as<double>() + 2 * 64 * 16; }
  double *Rinv() const { return G_.as<double>() + 3 * 64 * 16; }

  // CholQR of the b columns at V(c0) on the host (two passes): used for the
  // start block and for replacement blocks; false if the block is rank
  // deficient
  bool cholqr_host(int c0)
  {
    hipStream_t s = op_->stream;
    const int b = b_;
    for (int pass = 0; pass < 2; ++pass) {
      kry::gram_xy(V(c0), ldv_, b, V(c0), b, n_, part_.as<double>(), G3(), s);
      std::vector<double> G(b * b), R(b * b, 0.0), Ri(b * b, 0.0);
      copy_sync(G.data(), G3(), sizeof(double) * b * b, hipMemcpyDeviceToHost, s);
      double dmax = 0;
      for (int j = 0; j < b; ++j) dmax = std::max(dmax, G[j + j * b]);
      for (int j = 0; j < b; ++j) {
        double d = G[j + j * b];
        for (int k = 0; k < j; ++k) d -= R[k + j * b] * R[k + j * b];
        if (!(d > 1e-24 * dmax) || !(dmax > 0)) {
          if (verbose_) {
            std::fprintf(stderr, "[tm_solve] block CholQR pass %d pivot %d = %.6g; Gram:", pass, j, d);
            for (int t = 0; t < b * b; ++t) std::fprintf(stderr, " %.6g", G[t]);
            std::fprintf(stderr, "\n");
          }
          return false;
        }
        d = std::sqrt(d);
        R[j + j * b] = d;
        for (int i = j + 1; i < b; ++i) {
          double t = G[j + i * b];
          for (int k = 0; k < j; ++k) t -= R[k + j * b] * R[k + i * b];
          R[j + i * b] = t / d;
        }
      }
      for (int j = 0; j < b; ++j) {
        Ri[j + j * b] = 1.0 / R[j + j * b];
        for (int i = j - 1; i >= 0; --i) {
          double t = 0;
          for (int k = i + 1; k <= j; ++k) t += R[i + k * b] * Ri[k + j * b];
          Ri[i + j * b] = -t / R[i + i * b];
        }
      }
      copy_sync(Rinv(), Ri.data(), sizeof(double) * b * b, hipMemcpyHostToDevice, s);
      kry::block_update<T>(nullptr, ldv_, 0, V(c0), b, nullptr, 0, Rinv(), n_, s);
    }
    return true;
  }

  // random block at V(c0) orthogonal to V[0, c0) and orthonormal
  void random_block(int c0, uint64_t salt)
  {
    hipStream_t s = op_->stream;
    for (int c = 0; c < b_; ++c)
      kry::random_uniform(V(c0 + c), n_, seed_ + 7919 * (uint64_t)(c + 1) + salt, s);
    for (int pass = 0; pass < 2 && c0 > 0; ++pass) {
      kry::gram_xy(V(0), ldv_, c0, V(c0), b_, n_, part_.as<double>(), G1(), s);
      kry::block_update<T>(V(0), ldv_, c0, V(c0), b_, G1(), c0, nullptr, n_, s);
    }
    TM_CHECK(cholqr_host(c0), TM_ENOCONV, "block Krylov: could not build an orthonormal block");
  }

  void start(const double *v0)
  {
    random_block(0, 0);
    if (v0) {
      kry::copy(v0, V(0), n_, op_->stream);
      TM_CHECK(cholqr_host(0), TM_ENOCONV, "block Krylov: start block is rank deficient");
    }
    nc_ = b_;
    std::fill(H_.begin(), H_.end(), 0.0);
  }

  // the kernels of one block step at the current nc_ (no host synchronisation)
  void issue_step(int step)
  {
    hipStream_t s = op_->stream;
    const int b = b_, nc = nc_, col0 = nc - b, nx = nc + b;
    double *Hd = Hd_.as<double>(), *flags = Hd + (size_t)ldh_ * mmax_;
    for (int c = 0; c < b; ++c)
      op_->apply_internal(V(col0 + c), V(nc + c), ldv_, 1, io_, comp_, 0, nullptr, s);
    matvecs_ += b;
    kry::gram_xy(V(0), ldv_, nx, V(nc), b, n_, part_.as<double>(), G1(), s);
    kry::block_update<T>(V(0), ldv_, nc, V(nc), b, G1(), nx, nullptr, n_, s);
    kry::gram_xy(V(0), ldv_, nx, V(nc), b, n_, part_.as<double>(), G2(), s);
    kry::block_close(G1(), G2(), nc, b, Hd, ldh_, col0, Rinv(), flags, step, 0, s);
    kry::block_update<T>(V(0), ldv_, nc, V(nc), b, G2(), nx, Rinv(), n_, s);
    kry::gram_xy(V(nc), ldv_, b, V(nc), b, n_, part_.as<double>(), G3(), s);
    kry::block_close(G3(), nullptr, nc, b, Hd, ldh_, col0, Rinv(), flags, step, 1, s);
    kry::block_update<T>(nullptr, ldv_, 0, V(nc), b, nullptr, 0, Rinv(), n_, s);
  }

  // grow the projected size to at least ncv
  void expand()
  {
    hipStream_t s = op_->stream;
    const int b = b_;
    double *Hd = Hd_.as<double>(), *flags = Hd + (size_t)ldh_ * mmax_;
    const int nc0 = nc_;
    kry::fill(flags, 0.0, 1, s);
    kry::fill(flags + 1, -1.0, 1, s);
    int steps = 0;
    std::vector<int> ncs;
    while (nc_ - b < ncv_) {
      ncs.push_back(nc_);
      issue_step(steps++);
      nc_ += b;
    }
    // read the new column blocks and the flags
    const int c_lo = nc0 - b, c_hi = nc_ - b;  // column range [c_lo, c_hi) of H
    double *hd = stage_.as<double>();
    TM_HIP(hipMemcpyAsync(hd, Hd + (size_t)c_lo * ldh_,
                          sizeof(double) * (size_t)ldh_ * (c_hi - c_lo), hipMemcpyDeviceToHost,
                          s));
    double fl[2];
    TM_HIP(hipMemcpyAsync(fl, flags, sizeof fl, hipMemcpyDeviceToHost, s));
    stream_sync_poll(s);
    const int bad = fl[1] < 0 ? steps : (int)fl[1];
    for (int q = 0; q < std::min(bad + 1, steps); ++q) {
      const int nc = ncs[q], col0 = nc - b;
      for (int c = 0; c < b; ++c)
        for (int i = 0; i < nc + b; ++i)
          H(i, col0 + c) = finite_or_throw(hd[i + (size_t)(col0 + c - c_lo) * ldh_]);
    }
    if (bad < steps) {
      // (numerical) rank deficiency at step `bad`: repair its block on the
      // host and continue from the next step
      if (verbose_)
        std::fprintf(stderr, "[tm_solve] block step %d rank deficient; repaired\n", bad);
      nc_ = ncs[bad];
      repair_block(nc_ - b);
      nc_ += b;
      matvecs_ -= (int64_t)b * (steps - bad - 1);
      expand();
    }
  }

  // W = V[nc, nc+b) of the block step with active block col0 = nc - b, and
  // the stored coupling C (rows [nc, nc+b) of column block col0) with
  // T V_A = V H_col + W C.  Rank-revealing re-orthonormalisation: with
  // W^T W = Q diag(l) Q^T the directions l_i > 1e-16 (|H_col|^2 + max l)
  // are kept (W Q_i / sqrt(l_i), coupling sqrt(l_i) Q_i^T), the others are
  // replaced by random directions with zero coupling; the block is then
  // projected once more against V and orthonormalised (CholQR twice), the
  // projection folded into H_col and the factors into the coupling.
  void repair_block(int col0)
  {
    hipStream_t s = op_->stream;
    const int b = b_, nc = col0 + b;
    kry::gram_xy(V(nc), ldv_, b, V(nc), b, n_, part_.as<double>(), G3(), s);
    std::vector<double> WW(b * b);
    copy_sync(WW.data(), G3(), sizeof(double) * b * b, hipMemcpyDeviceToHost, s);
    std::vector<double> Hc((size_t)(nc + b) * b);
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < nc + b; ++i) Hc[i + (size_t)c * (nc + b)] = H(i, col0 + c);
    std::vector<double> lam(b), Q(b * b);
    for (int c1 = 0; c1 < b; ++c1)
      for (int c2 = 0; c2 < b; ++c2)
        WW[c1 + c2 * b] = 0.5 * (WW[c1 + c2 * b] + WW[c2 + c1 * b]);
    sym_eig(b, WW.data(), lam.data(), Q.data());
    double hn = 0, lmax = 0;
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < nc; ++i) hn += Hc[i + (size_t)c * (nc + b)] * Hc[i + (size_t)c * (nc + b)];
    for (double l : lam) lmax = std::max(lmax, l);
    std::vector<int> keep;
    for (int i = 0; i < b; ++i)
      if (lam[i] > 1e-16 * (hn + lmax)) keep.push_back(i);
    const int r = (int)keep.size();
    // Cm: W <- W Cm over the kept directions; Rp: their coupling (r x b)
    std::vector<double> Cm((size_t)b * b, 0.0), Rp((size_t)b * b, 0.0);
    for (int q = 0; q < r; ++q) {
      const int i = keep[q];
      const double sl = std::sqrt(lam[i]);
      for (int t = 0; t < b; ++t) {
        Cm[t + (size_t)q * b] = Q[t + (size_t)i * b] / sl;
        Rp[q + (size_t)t * b] = sl * Q[t + (size_t)i * b];
      }
    }
    if (r > 0) {
      copy_sync(Yd_.get(), Cm.data(), sizeof(double) * b * b, hipMemcpyHostToDevice, s);
      kry::gemm_vy(V(nc), ldv_, b, Yd_.as<double>(), b, r, V(nc), ldv_, n_, s);
    }
    for (int c = r; c < b; ++c)
      kry::random_uniform(V(nc + c), n_, seed_ + 104729 * (uint64_t)(c + 1 + 17 * restarts_), s);
    // project the whole block against V (twice) and orthonormalise it
    std::vector<double> h3((size_t)nc * b, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      kry::gram_xy(V(0), ldv_, nc, V(nc), b, n_, part_.as<double>(), G1(), s);
      std::vector<double> g((size_t)nc * b);
      copy_sync(g.data(), G1(), sizeof(double) * nc * b, hipMemcpyDeviceToHost, s);
      for (size_t t = 0; t < g.size(); ++t) h3[t] += g[t];
      kry::block_update<T>(V(0), ldv_, nc, V(nc), b, G1(), nc, nullptr, n_, s);
    }
    // W_kept = W_new R3 + V h3 with R3 from CholQR (two passes, host)
    std::vector<double> R3((size_t)b * b, 0.0);
    for (int i = 0; i < b; ++i) R3[i + i * b] = 1.0;
    for (int pass = 0; pass < 2; ++pass) {
      kry::gram_xy(V(nc), ldv_, b, V(nc), b, n_, part_.as<double>(), G3(), s);
      std::vector<double> G(b * b), R(b * b, 0.0), Ri(b * b, 0.0);
      copy_sync(G.data(), G3(), sizeof(double) * b * b, hipMemcpyDeviceToHost, s);
      for (int j = 0; j < b; ++j) {
        double d = G[j + j * b];
        for (int k = 0; k < j; ++k) d -= R[k + j * b] * R[k + j * b];
        TM_CHECK(d > 0, TM_ENOCONV, "block Krylov: repaired block is rank deficient");
        d = std::sqrt(d);
        R[j + j * b] = d;
        for (int i = j + 1; i < b; ++i) {
          double t = G[j + i * b];
          for (int k = 0; k < j; ++k) t -= R[k + j * b] * R[k + i * b];
          R[j + i * b] = t / d;
        }
      }
      for (int j = 0; j < b; ++j) {
        Ri[j + j * b] = 1.0 / R[j + j * b];
        for (int i = j - 1; i >= 0; --i) {
          double t = 0;
          for (int k = i + 1; k <= j; ++k) t += R[i + k * b] * Ri[k + j * b];
          Ri[i + j * b] = -t / R[i + i * b];
        }
      }
      copy_sync(Rinv(), Ri.data(), sizeof(double) * b * b, hipMemcpyHostToDevice, s);
      kry::block_update<T>(nullptr, ldv_, 0, V(nc), b, nullptr, 0, Rinv(), n_, s);
      std::vector<double> t3((size_t)b * b, 0.0);
      for (int i = 0; i < b; ++i)
        for (int j = 0; j < b; ++j)
          for (int k = 0; k < b; ++k) t3[i + j * b] += R[i + k * b] * R3[k + j * b];
      R3 = t3;
    }
    // T V_A = V Hc + W_old C_old, W_old = W_kept Rp (kept directions only),
    // W_kept = W_new R3 + V h3 (h3 over the kept and the random columns;
    // the random ones carry no coupling):
    //   H_col += h3 Rp C_old, coupling = R3 Rp C_old
    std::vector<double> Cold((size_t)b * b);
    for (int c = 0; c < b; ++c)
      for (int i = 0; i < b; ++i) Cold[i + (size_t)c * b] = H(nc + i, col0 + c);
    std::vector<double> RC((size_t)b * b, 0.0);
    for (int i = 0; i < b; ++i)
      for (int c = 0; c < b; ++c)
        for (int k = 0; k < b; ++k) RC[i + (size_t)c * b] += Rp[i + (size_t)k * b] * Cold[k + (size_t)c * b];
    for (int c = 0; c < b; ++c) {
      for (int i = 0; i < nc; ++i) {
        double t = 0;
        for (int k = 0; k < b; ++k) t += h3[i + (size_t)k * nc] * RC[k + (size_t)c * b];
        H(i, col0 + c) += t;
      }
      for (int i = 0; i < b; ++i) {
        double t = 0;
        for (int k = 0; k < b; ++k) t += R3[i + (size_t)k * b] * RC[k + (size_t)c * b];
        H(nc + i, col0 + c) = t;
      }
    }
  }

  Ritz ritz()
  {
    const int m = nc_ - b_;
    m_ = m;
    std::vector<double> A((size_t)m * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) A[i + (size_t)j * m] = 0.5 * (H(i, j) + H(j, i));
    std::vector<double> w(m), Z((size_t)m * m);
    sym_eig(m, A.data(), w.data(), Z.data());
    std::vector<int> idx(m);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int c) { return std::fabs(w[a]) > std::fabs(w[c]); });
    Ritz r;
    r.wr.resize(m);
    r.wi.assign(m, 0.0);
    r.Zr.resize((size_t)m * m);
    r.Zi.assign((size_t)m * m, 0.0);
    r.res.resize(m);
    for (int c = 0; c < m; ++c) {
      r.wr[c] = w[idx[c]];
      for (int i = 0; i < m; ++i) r.Zr[i + (size_t)c * m] = Z[i + (size_t)idx[c] * m];
      // residual ||R z_last||, R = H[m .. m+b, m-b .. m)
      double rr = 0;
      for (int i = 0; i < b_; ++i) {
        double t = 0;
        for (int q = 0; q < b_; ++q) t += H(m + i, m - b_ + q) * r.Zr[(m - b_ + q) + (size_t)c * m];
        rr += t * t;
      }
      r.res[c] = std::sqrt(rr);
    }
    return r;
  }

  int nconv(const Ritz &r, double tol) const
  {
    int nc = 0;
    const double e23 = std::pow(EPS, 2.0 / 3.0);
    for (int i = 0; i < nev_; ++i)
      if (r.res[i] <= tol * std::max(e23, std::fabs(r.wr[i]))) ++nc;
    return nc;
  }

  void restart(const Ritz &r, int k)
  {
    const int m = m_, b = b_;
    hipStream_t s = op_->stream;
    double *hy = stage_.as<double>();
    for (int c = 0; c < k; ++c)
      for (int i = 0; i < m; ++i) hy[i + (size_t)c * m] = r.Zr[i + (size_t)c * m];
    TM_HIP(hipMemcpyAsync(Yd_.get(), hy, sizeof(double) * m * k, hipMemcpyHostToDevice, s));
    kry::gemm_vy(V(0), ldv_, m, Yd_.as<double>(), m, k, V(0), ldv_, n_, s);
    for (int c = 0; c < b; ++c) kry::copy(V(m + c), V(k + c), n_, s);
    std::vector<double> Rm((size_t)b * b);
    for (int i = 0; i < b; ++i)
      for (int q = 0; q < b; ++q) Rm[i + (size_t)q * b] = H(m + i, m - b + q);
    std::fill(H_.begin(), H_.end(), 0.0);
    for (int c = 0; c < k; ++c) {
      H(c, c) = r.wr[c];
      for (int i = 0; i < b; ++i) {
        double t = 0;
        for (int q = 0; q < b; ++q) t += Rm[i + (size_t)q * b] * r.Zr[(m - b + q) + (size_t)c * m];
        H(k + i, c) = t;
      }
    }
    nc_ = k + b;
    ++restarts_;
    // the staging buffer is next written by a stream-ordered read-back
  }

  int run(Ritz &out, const double *v0) override
  {
    start(v0);
    int nc = 0;
    double best = 1e300;
    for (int it = 0; it < maxit_; ++it) {
      expand();
      out = ritz();
      nc = nconv(out, tol_);
      double worst = 0;
      for (int i = 0; i < nev_; ++i)
        worst = std::max(worst, out.res[i] / std::max(std::fabs(out.wr[i]), 1e-300));
      if (verbose_)
        std::fprintf(stderr,
                     "[tm_solve] block cycle %d matvecs %lld m %d lambda0 %.15g worst %.3e "
                     "nconv %d\n",
                     it, (long long)matvecs_, m_, out.wr[0], worst, nc);
      if (nc >= nev_) break;
      if (stall_stop_ && worst < 1e-4 && worst > 0.7 * best) {
        stalled_ = true;
        break;
      }
      best = std::min(best, worst);
      const int m = m_;
      int k = nev_ + std::min(nc, (m - nev_) / 2);
      k = std::max(k, (nev_ + m) / 2);
      k = std::min(k, m - b_);
      k = std::max(k, nev_);
      restart(out, k);
    }
    return nc;
  }

  void ritz_vectors(const Ritz &r, int cnt, double *Xr, double *Xi) override
  {
    const int m = m_;
    hipStream_t s = op_->stream;
    std::vector<double> Z((size_t)m * cnt);
    for (int c = 0; c < cnt; ++c)
      for (int t = 0; t < m; ++t) Z[t + (size_t)c * m] = r.Zr[t + (size_t)c * m];
    TM_HIP(hipMemcpyAsync(Yd_.get(), Z.data(), sizeof(double) * m * cnt, hipMemcpyHostToDevice, s));
    kry::gemm_vy(V(0), ldv_, m, Yd_.as<double>(), m, cnt, Xr, ldv_, n_, s);
    stream_sync_poll(s);
    if (Xi) kry::fill(Xi, 0.0, ldv_ * cnt, s);
  }

 private:
  tm_op *op_;
  int nev_, ncv_, b_;
  double tol_;
  int maxit_;
  tm_dtype comp_;
  const tm_dtype io_ = sizeof(T) == 4 ? TM_F32 : TM_F64;
  uint64_t seed_;
  int verbose_;
  SolverCache &cache_;
  DevBuf &V_, &part_, &G_, &Hd_, &Yd_;
  PinnedBuf &stage_;
  int64_t n_ = 0, ldv_ = 0;
  int mmax_ = 0, ncap_ = 0, ldh_ = 0;
  int nc_ = 0, m_ = 0;
  int64_t matvecs_ = 0, restarts_ = 0;
  bool stall_stop_ = false, stalled_ = false;
  std::vector<double> H_;
};

// Cholesky of a small SPD matrix (k x k, column-major), lower factor in place.
bool cholesky(int k, std::vector<double> &G)
{
  for (int j = 0; j < k; ++j) {
    double d = G[j + (size_t)j * k];
    for (int t = 0; t < j; ++t) d -= G[j + (size_t)t * k] * G[j + (size_t)t * k];
    if (!(d > 0)) return false;
    d = std::sqrt(d);
    G[j + (size_t)j * k] = d;
    for (int i = j + 1; i < k; ++i) {
      double s = G[i + (size_t)j * k];
      for (int t = 0; t < j; ++t) s -= G[i + (size_t)t * k] * G[j + (size_t)t * k];
      G[i + (size_t)j * k] = s / d;
    }
    for (int i = 0; i < j; ++i) G[i + (size_t)j * k] = 0;
  }
  return true;
}

struct RefineOut {
  std::vector<double> theta, res;
  double bound = 0;
  int64_t matvecs = 0;
  bool failed = false;  // Y^T Y not positive definite
  bool met = false;     // the last iteration met the residual or the error bound
};

// FP64 Rayleigh-Ritz refinement of the block Y = B[0..k) (leading dimension
// ld; any basis of the subspace, e.g. Ritz vectors of an FP32-stored Krylov
// basis) for a symmetric operator.  B holds 2k vectors [Y | W].  Each
// iteration: W = A Y (k FP64 matvecs); the Gram matrix of [Y | W] in one
// sweep; on the host, with Y^T Y = L L^T (Cholesky), the Ritz pairs
// (theta, q) of L^-1 sym(Y^T W) L^-T and Z = L^-T Q, so that X = Y Z has
// orthonormal columns; then the residual norms ||W z - theta Y z|| in one
// sweep that forms the sums explicitly (no cancellation) without storing
// them.  Accepted when every wanted residual meets tol or, when the caller
// takes eigenvalues only (vectors = false), the Rayleigh-Ritz error bound
// r^2 / (gap |theta|) (gap to the Krylov Ritz values outside the block) is
// below 1e-13 -- returned eigenvectors always meet the residual tolerance;
// otherwise Y <- W Z diag(1/theta) (= A X scaled, one
// subspace-iteration step) and repeat.  On return B[0..nev) hold X.
// failed = true if Y^T Y is not positive definite (a rank-deficient block); met =
// false if maxiter iterations end without meeting either test (the caller
// then does not use the block).
RefineOut refine_sym(tm_op *op, double *B, int64_t ld, int64_t n, int k, int nev, double tol,
                     int maxiter, int verbose, const std::vector<double> &outside, bool vectors)
{
  hipStream_t s = op->stream;
  RefineOut ro;
  const int k2 = 2 * k;
  SolverCache &cache = op->solver_cache();
  DevBuf &partial = cache.dev[SLOT_REFINE_PARTIAL], &G = cache.dev[SLOT_REFINE_G],
         &C = cache.dev[SLOT_REFINE_C];
  PinnedBuf &stage = cache.pinned[3];
  partial.ensure(sizeof(double) * kry::NBLK * 72);
  G.ensure(sizeof(double) * ((size_t)k2 * k2 + k));
  C.ensure(sizeof(double) * ((size_t)k2 * k + (size_t)k * k));
  stage.ensure(sizeof(double) * ((size_t)k2 * k2 + (size_t)k2 * k + (size_t)k * k + k));
  double *hG = stage.as<double>();               // Gram matrix read back
  double *hCr = hG + (size_t)k2 * k2;            // residual combinations (k2 x k)
  double *hZ = hCr + (size_t)k2 * k;             // next-block / output combinations
  double *hR = hZ + (size_t)k * k;               // residual norms^2 read back
  double *dCr = C.as<double>(), *dZ = dCr + (size_t)k2 * k;
  double *dR = G.as<double>() + (size_t)k2 * k2;
  auto col = [&](int c) { return B + (size_t)c * ld; };
  for (int it = 0; it < maxiter; ++it) {
    for (int c = 0; c < k; ++c) {
      op->apply_internal(col(c), col(k + c), ld, 1, TM_F64, TM_F64, 0, nullptr, s);
      ++ro.matvecs;
    }
    kry::gram_sym(col(0), ld, k2, n, partial.as<double>(), G.as<double>(), s);
    TM_HIP(hipMemcpyAsync(hG, G.get(), sizeof(double) * k2 * k2, hipMemcpyDeviceToHost, s));
    stream_sync_poll(s);
    // L = chol(Y^T Y), Li = L^-1, M = Li sym(Y^T W) Li^T
    std::vector<double> Lf((size_t)k * k), Li((size_t)k * k, 0.0), M((size_t)k * k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) Lf[i + (size_t)j * k] = hG[i * k2 + j];
    if (!cholesky(k, Lf)) {
      ro.failed = true;
      return ro;
    }
    for (int c = 0; c < k; ++c)  // forward substitution, column c of Li
      for (int i = c; i < k; ++i) {
        double t = i == c ? 1.0 : 0.0;
        for (int q = c; q < i; ++q) t -= Lf[i + (size_t)q * k] * Li[q + (size_t)c * k];
        Li[i + (size_t)c * k] = t / Lf[i + (size_t)i * k];
      }
    std::vector<double> S((size_t)k * k), LS((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j)
        S[i + (size_t)j * k] = 0.5 * (hG[i * k2 + k + j] + hG[j * k2 + k + i]);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        double t = 0;
        for (int q = 0; q <= i; ++q) t += Li[i + (size_t)q * k] * S[q + (size_t)j * k];
        LS[i + (size_t)j * k] = t;
      }
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        double t = 0;
        for (int q = 0; q <= j; ++q) t += LS[i + (size_t)q * k] * Li[j + (size_t)q * k];
        M[i + (size_t)j * k] = t;
      }
    std::vector<double> th(k), Q((size_t)k * k);
    sym_eig(k, M.data(), th.data(), Q.data());
    std::vector<int> idx(k);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return std::fabs(th[a]) > std::fabs(th[b]); });
    // Z = Li^T Q (sorted)
    std::vector<double> ths(k), Z((size_t)k * k);
    for (int c = 0; c < k; ++c) {
      ths[c] = th[idx[c]];
      for (int t = 0; t < k; ++t) {
        double z = 0;
        for (int q = t; q < k; ++q) z += Li[q + (size_t)t * k] * Q[q + (size_t)idx[c] * k];
        Z[t + (size_t)c * k] = z;
      }
    }
    // residual r_c = [Y | W] [-theta_c z_c; z_c]: norms of the wanted ones
    for (int c = 0; c < k; ++c)
      for (int t = 0; t < k; ++t) {
        hCr[t + (size_t)c * k2] = -ths[c] * Z[t + (size_t)c * k];
        hCr[k + t + (size_t)c * k2] = Z[t + (size_t)c * k];
      }
    TM_HIP(hipMemcpyAsync(dCr, hCr, sizeof(double) * k2 * nev, hipMemcpyHostToDevice, s));
    kry::combine_norms(col(0), ld, k2, dCr, k2, nev, n, partial.as<double>(), dR, s);
    TM_HIP(hipMemcpyAsync(hR, dR, sizeof(double) * nev, hipMemcpyDeviceToHost, s));
    stream_sync_poll(s);
    ro.theta = ths;
    ro.res.assign(k, 0.0);
    double worst = 0, bound = 0;
    std::vector<double> bnd(nev);
    for (int c = 0; c < nev; ++c) {
      const double r = std::sqrt(std::max(hR[c], 0.0));
      ro.res[c] = r / std::fabs(ths[c]);
      worst = std::max(worst, ro.res[c]);
      // error bound r^2 / (gap |theta|), gap to the spectrum outside the block
      double gap = 1e300;
      for (double t : outside) gap = std::min(gap, std::fabs(std::fabs(ths[c]) - std::fabs(t)));
      if (outside.empty())
        for (int d = 0; d < k; ++d)
          if (d != c) gap = std::min(gap, std::fabs(ths[c] - ths[d]));
      bnd[c] = r * r / (gap * std::fabs(ths[c]));
      bound = std::max(bound, bnd[c]);
    }
    if (verbose) {
      std::fprintf(stderr, "[tm_refine] it %d theta0 %.16g res %.3e bound", it, ths[0], worst);
      for (int c = 0; c < nev; ++c) std::fprintf(stderr, " %.2e", bnd[c]);
      std::fprintf(stderr, "\n");
    }
    ro.bound = bound;
    ro.met = worst <= tol || (!vectors && bound <= 1e-13);
    const bool done = ro.met || it == maxiter - 1;
    if (done) {
      // X = Y Z[:, :nev] in place (every row is read before it is written)
      for (int c = 0; c < nev; ++c)
        for (int t = 0; t < k; ++t) hZ[t + (size_t)c * k] = Z[t + (size_t)c * k];
      TM_HIP(hipMemcpyAsync(dZ, hZ, sizeof(double) * k * nev, hipMemcpyHostToDevice, s));
      kry::gemm_vy(col(0), ld, k, dZ, k, nev, col(0), ld, n, s);
      stream_sync_poll(s);
      break;
    }
    // next block Y = W Z diag(1/theta) (close to orthonormal)
    for (int c = 0; c < k; ++c)
      for (int t = 0; t < k; ++t)
        hZ[t + (size_t)c * k] = Z[t + (size_t)c * k] / (ths[c] != 0 ? ths[c] : 1.0);
    TM_HIP(hipMemcpyAsync(dZ, hZ, sizeof(double) * k * k, hipMemcpyHostToDevice, s));
    kry::gemm_vy(col(k), ld, k, dZ, k, k, col(0), ld, n, s);
    stream_sync_poll(s);
  }
  return ro;
}

int suggested_ncv(int neigs) { return neigs > 4 ? 2 * neigs + 1 : 2 * neigs + 3; }  // tmat.hpp:88

// Explicit FP64 residual of the leading Ritz pair (lr + i li, xr + i xi),
// vectors of leading dimension ld in the operator's internal order: one
// FP64 T.x per real vector, then ||T x - lambda x|| / (|lambda| ||x||) with
// every sum formed explicitly (no cancellation).  Independent of the Krylov
// recurrence, so a corrupted basis or operator result cannot pass it.
double leading_residual(tm_op *op, const double *xr, const double *xi, int64_t ld, int64_t n,
                        double lr, double li)
{
  hipStream_t s = op->stream;
  SolverCache &cache = op->solver_cache();
  const bool cplx = li != 0 && xi != nullptr;
  const int h = cplx ? 2 : 1, nv = 2 * h;
  DevBuf &C = cache.dev[SLOT_CHECK], &partial = cache.dev[SLOT_REFINE_PARTIAL];
  partial.ensure(sizeof(double) * kry::NBLK * 72);
  C.ensure(sizeof(double) * ((size_t)ld * nv + 32));
  double *U = C.as<double>();
  double *cm = U + (size_t)ld * nv, *res = cm + 16;
  kry::copy(xr, U, n, s);
  if (cplx) kry::copy(xi, U + ld, n, s);
  for (int c = 0; c < h; ++c)
    op->apply_internal(U + (size_t)c * ld, U + (size_t)(h + c) * ld, ld, 1, TM_F64, TM_F64, 0,
                       nullptr, s);
  // columns: the residual(s) T x - lambda x, then the pieces of ||x||
  std::vector<double> co((size_t)nv * nv, 0.0);
  if (!cplx) {
    co[0] = -lr;  // T x - lr x
    co[1] = 1.0;
    co[2] = 1.0;  // x
  } else {
    const double a = lr, b = li;
    double *c0 = &co[0], *c1 = &co[4], *c2 = &co[8], *c3 = &co[12];
    c0[0] = -a; c0[1] = b; c0[2] = 1.0;   // T xr - (a xr - b xi)
    c1[0] = -b; c1[1] = -a; c1[3] = 1.0;  // T xi - (b xr + a xi)
    c2[0] = 1.0;
    c3[1] = 1.0;
  }
  copy_sync(cm, co.data(), sizeof(double) * co.size(), hipMemcpyHostToDevice, s);
  kry::combine_norms(U, ld, nv, cm, nv, nv, n, partial.as<double>(), res, s);
  double hr[4] = {0, 0, 0, 0};
  copy_sync(hr, res, sizeof(double) * nv, hipMemcpyDeviceToHost, s);
  const double rn = cplx ? hr[0] + hr[1] : hr[0], xn = cplx ? hr[2] + hr[3] : hr[1];
  const double lam = std::hypot(lr, li);
  return std::sqrt(std::max(rn, 0.0)) / (lam * std::sqrt(xn));
}

// Largest explicit relative residual of the leading pair accepted as a
// converged result (the Krylov test asks for 1e-10; an invalid result is
// O(1)).
constexpr double VERIFY_MAX = 1e-6;

}  // namespace

// ---------------------------------------------------------------------------
// public solve
// ---------------------------------------------------------------------------
tm_status solve_impl(tm_op *op, const tm_solve_opts &o, tm_solve_out *out)
{
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = op->ncol;
  const int nev = o.nev;
  TM_CHECK(nev >= 1 && nev < n, TM_EINVAL, "nev must be in [1, cols())");
  int ncv = o.ncv > 0 ? o.ncv : suggested_ncv(nev);
  if (ncv >= n) ncv = (int)n - 1;  // pfunc.hpp:90
  TM_CHECK(ncv > nev, TM_EINVAL, "ncv must exceed nev");
  TM_CHECK(ncv <= 60, TM_EINVAL, "ncv above 60 is not supported by the device kernels");
  const bool mixed = o.precision == TM_PREC_MIXED;
  const int64_t mv0 = op->matvecs;
  int64_t mv32 = 0, mv64 = 0, restarts = 0, gcycles = 0;
  int trial = 0, nc = 0;
  Ritz rz;
  std::unique_ptr<KrylovRun> ks;
  // pfunc.hpp:87-109: up to maxtrial attempts, ncv += 2 on failure
  for (;;) {
    const double tol1 = mixed ? std::max(o.mixed_tol, o.tol) : o.tol;
    // the block form for symmetric operators whose dimension is well above
    // the basis (smaller ones run the single-vector form)
    const bool blk = o.block > 1 && op->symmetric && n >= 8 * (int64_t)(ncv + 2 * o.block);
    if (blk && mixed && o.basis_f32)
      ks.reset(new BlockKrylovSchur<float>(op, nev, ncv, o.block, tol1, o.maxit, TM_F32, o.seed,
                                           o.verbose));
    else if (blk)
      ks.reset(new BlockKrylovSchur<double>(op, nev, ncv, o.block, tol1, o.maxit,
                                            mixed ? TM_F32 : TM_F64, o.seed, o.verbose));
    else if (mixed && o.basis_f32)
      ks.reset(new KrylovSchur<float>(op, nev, ncv, tol1, o.maxit, TM_F32, 0, o.seed, o.verbose,
                                      0));
    else
      ks.reset(new KrylovSchur<double>(op, nev, ncv, tol1, o.maxit, mixed ? TM_F32 : TM_F64, 0,
                                       o.seed, o.verbose, 0));
    if (mixed) ks->set_stall_stop(true);
    ks->set_graph(o.use_graph != 0);
    nc = ks->run(rz, nullptr);
    if (mixed && ks->stalled()) nc = nev;  // FP32 phase ends at its noise floor
    restarts += ks->restarts();
    gcycles += ks->graph_cycles();
    if (mixed) mv32 += ks->matvecs(); else mv64 += ks->matvecs();
    ++trial;
    if (nc >= nev) break;
    if (o.verbose) std::fprintf(stderr, "eigensolver not converged (%d/%d)\n", trial, o.maxtrial);
    if (trial >= o.maxtrial) break;
    ncv += 2;
    if (ncv >= n) ncv = (int)n - 1;
  }
  if (nc == 0 && !mixed) {
    out->nconv = 0;
    out->trials = trial;
    throw Error(TM_ENOCONV, "Number of converged eigenvalue is zero");
  }
  const int64_t ldv = ks->ldv();
  std::vector<double> evr(nev), evi(nev), res(nev);
  SolverCache &cache = op->solver_cache();
  DevBuf &Xr = cache.dev[SLOT_XR], &Xi = cache.dev[SLOT_XI];
  const bool cplx = !op->symmetric;
  const int refine = o.refine != 0 ? o.refine : (op->symmetric ? 1 : 2);
  const double *xr = nullptr, *xi = nullptr;  // device Ritz vectors (ld ldv)
  bool refined = false;
  if (mixed && op->symmetric && refine == 1) {
    // FP64 block Rayleigh-Ritz refinement of the kept FP32 Ritz vectors
    const int k = std::min(ks->m() - 1, nev + std::max(0, o.refine_extra));
    DevBuf &B = cache.dev[SLOT_REFINE_B];
    B.ensure(sizeof(double) * ldv * 2 * k);
    ks->ritz_vectors(rz, k, B.as<double>(), nullptr);
    std::vector<double> outside;
    for (int d = k; d < ks->m(); ++d) outside.push_back(rz.wr[d]);
    // with eigenvectors wanted, one iteration: subspace iteration brings the
    // FP32 Ritz vectors' residuals (about 1e-7) to tol only slowly, at k FP64
    // T.x per iteration, so a block that misses tol after one goes straight
    // to the FP64 Krylov-Schur below
    RefineOut ro = refine_sym(op, B.as<double>(), ldv, n, k, nev, o.tol, o.want_vectors ? 1 : 4,
                              o.verbose, outside, o.want_vectors != 0);
    mv64 += ro.matvecs;
    if (!ro.failed && ro.met) {
      for (int i = 0; i < nev; ++i) {
        evr[i] = ro.theta[i];
        evi[i] = 0;
        res[i] = ro.res[i];
      }
      nc = 0;
      for (int i = 0; i < nev; ++i) nc += std::isfinite(ro.theta[i]) ? 1 : 0;
      xr = B.as<double>();
      refined = true;
    } else if (o.verbose) {
      std::fprintf(stderr, "[tm_refine] %s; FP64 Krylov-Schur instead\n",
                   ro.failed ? "rank-deficient block" : "residual and error bound not met");
    }
  }
  if (!refined) {
    std::unique_ptr<KrylovRun> k2;
    KrylovRun *fin = ks.get();
    if (mixed) {
      // FP64 Krylov-Schur warm-started from the sum of the wanted FP32 Ritz
      // vectors (real parts): the FP64 run converges to tol in a cycle or two
      DevBuf &X0 = cache.dev[SLOT_REFINE_B], &v0 = cache.dev[SLOT_REFINE_C];
      X0.ensure(sizeof(double) * ldv * nev);
      v0.ensure(sizeof(double) * ldv);
      ks->ritz_vectors(rz, nev, X0.as<double>(), nullptr);
      std::vector<double> ones((size_t)nev * 1, 1.0);
      DevBuf &Y1 = cache.dev[SLOT_REFINE_G];
      Y1.ensure(sizeof(double) * nev);
      copy_sync(Y1.get(), ones.data(), sizeof(double) * nev, hipMemcpyHostToDevice, op->stream);
      kry::gemm_vy(X0.as<double>(), ldv, nev, Y1.as<double>(), nev, 1, v0.as<double>(), ldv, n,
                   op->stream);
      k2.reset(new KrylovSchur<double>(op, nev, ncv, o.tol, o.maxit, TM_F64, 0, o.seed, o.verbose,
                                       1));
      k2->set_graph(o.use_graph != 0);
      nc = k2->run(rz, v0.as<double>());
      mv64 += k2->matvecs();
      restarts += k2->restarts();
      gcycles += k2->graph_cycles();
      fin = k2.get();
    }
    Xr.ensure(sizeof(double) * fin->ldv() * nev);
    if (cplx) Xi.ensure(sizeof(double) * fin->ldv() * nev);
    fin->ritz_vectors(rz, nev, Xr.as<double>(), cplx ? Xi.as<double>() : nullptr);
    xr = Xr.as<double>();
    xi = cplx ? Xi.as<double>() : nullptr;
    for (int i = 0; i < nev; ++i) {
      evr[i] = rz.wr[i];
      evi[i] = rz.wi[i];
      const double mag = std::hypot(rz.wr[i], rz.wi[i]);
      res[i] = rz.res[i] / (mag > 0 ? mag : 1.0);
    }
  }
  // explicit residual of the leading pair: the refinement's residuals are
  // already explicit FP64 ones; otherwise one FP64 T.x of the Ritz vector
  out->resid_check = std::nan("");
  if (o.verify && nc >= 1) {
    double rc = refined ? res[0]
                        : leading_residual(op, xr, xi, ldv, n, evr[0], cplx ? evi[0] : 0.0);
    if (!refined) ++mv64;
    out->resid_check = rc;
    if (!(rc <= VERIFY_MAX)) {
      char msg[160];
      std::snprintf(msg, sizeof msg,
                    "explicit residual check failed: ||Tx - lambda x|| / |lambda| = %.3e for "
                    "lambda = %.17g",
                    rc, evr[0]);
      throw Error(TM_EHIP, msg);
    }
  }
  out->nconv = std::min(nc, nev);
  for (int i = 0; i < nev; ++i) {
    if (out->evals_re) out->evals_re[i] = evr[i];
    if (out->evals_im) out->evals_im[i] = evi[i];
    if (out->resid) out->resid[i] = res[i];
  }
  if (o.want_vectors && out->vecs) {
    if (!cplx) {
      std::vector<double> tmp(op->perm.empty() ? 0 : n);
      for (int i = 0; i < nev; ++i) {
        double *dst = out->vecs + (size_t)i * n;
        if (op->perm.empty()) {
          copy_large(dst, xr + (size_t)i * ldv, sizeof(double) * n, hipMemcpyDeviceToHost,
                     op->stream);
        } else {
          // internal orbit order -> ascending representatives
          copy_large(tmp.data(), xr + (size_t)i * ldv, sizeof(double) * n,
                     hipMemcpyDeviceToHost, op->stream);
          for (int64_t t = 0; t < n; ++t) dst[op->perm[t]] = tmp[t];
        }
      }
    } else {
      // cols() x nev complex128 interleaved, column-major (pfunc.hpp:139)
      std::vector<double> re(n), im(n);
      for (int i = 0; i < nev; ++i) {
        copy_large(re.data(), xr + (size_t)i * ldv, sizeof(double) * n, hipMemcpyDeviceToHost,
                   op->stream);
        copy_large(im.data(), xi + (size_t)i * ldv, sizeof(double) * n, hipMemcpyDeviceToHost,
                   op->stream);
        for (int64_t t = 0; t < n; ++t) {
          out->vecs[2 * ((size_t)i * n + t)] = re[t];
          out->vecs[2 * ((size_t)i * n + t) + 1] = im[t];
        }
      }
    }
  }
  // left eigenvector (getlefteigvec, pfunc.hpp:47-67): T^T with nev=1, ncv=7
  out->eval_left = std::nan("");
  if (o.want_left && !op->symmetric) {
    int ncvl = 7;
    if (ncvl >= n) ncvl = (int)n - 1;
    KrylovSchur<double> kl(op, 1, ncvl, o.tol, o.maxit, TM_F64, 1, o.seed + 1, o.verbose, 2);
    kl.set_graph(o.use_graph != 0);
    Ritz rl;
    const int ncl = kl.run(rl, nullptr);
    mv64 += kl.matvecs();
    gcycles += kl.graph_cycles();
    if (ncl >= 1) {
      out->eval_left = rl.wr[0];
      if (out->vecl) {
        DevBuf &XL = cache.dev[SLOT_XI];  // the right vectors are on the host by now
        XL.ensure(sizeof(double) * kl.ldv());
        kl.ritz_vectors(rl, 1, XL.as<double>(), nullptr);
        copy_large(out->vecl, XL.get(), sizeof(double) * n, hipMemcpyDeviceToHost, op->stream);
      }
    } else if (o.verbose) {
      std::fprintf(stderr, "Warning: left eigenvalue calculation failed.\n");
    }
  }
  out->matvecs_f32 = mv32;
  out->matvecs_f64 = mv64;
  (void)mv0;
  out->restarts = restarts;
  out->graph_cycles = gcycles;
  out->trials = trial;
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return nc >= nev ? TM_OK : TM_ENOCONV;
}

}  // namespace ltm
