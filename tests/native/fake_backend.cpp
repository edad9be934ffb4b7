// CPU implementations of the Krylov vector kernels, a dense-matrix operator
// and stubs for the device paths the C ABI reaches, for unit tests of the
// eigen driver and the C ABI without a GPU.  Every "kernel" is queued on its
// stream through the deferred-execution runtime stand-in (fake_runtime.cpp)
// and runs only when the host waits, in stream order, like a device launch.
// The reductions keep the device data flow: a sweep writes its sums into the
// partial buffer (one block), the reduce step reads them from there.  Test
// infrastructure only.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../latttmat_amd/csrc/tm_internal.hpp"

namespace ltm {
namespace kry {
namespace {
template <typename T>
void dots_now(const T *V, int64_t ldv, int j, const T *w, int64_t n, double *out)
{
  for (int k = 0; k <= j; ++k) out[k] = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < j; ++k) out[k] += (double)V[k * ldv + i] * (double)w[i];
    out[j] += (double)w[i] * (double)w[i];
  }
}
template <typename T>
void subtract_now(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h)
{
  std::vector<double> hh(h, h + j);
  for (int64_t i = 0; i < n; ++i) {
    double wi = w[i];
    for (int k = 0; k < j; ++k) wi -= hh[k] * (double)V[k * ldv + i];
    w[i] = (T)wi;
  }
}
}  // namespace

template <typename T>
unsigned multidot(const T *V, int64_t ldv, int j, const T *w, int64_t n, double *partial,
                  double *out, hipStream_t s, bool reduce)
{
  fake_enqueue(s, [=] {
    dots_now(V, ldv, j, w, n, partial);
    if (reduce) std::memcpy(out, partial, sizeof(double) * (j + 1));
  });
  return 1;
}
template <typename T>
unsigned update_dots(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                     double *partial, double *out, hipStream_t s, bool reduce)
{
  fake_enqueue(s, [=] {
    subtract_now(V, ldv, j, w, n, h);
    dots_now(V, ldv, j, w, n, partial);
    if (reduce) std::memcpy(out, partial, sizeof(double) * (j + 1));
  });
  return 1;
}
template <typename T>
unsigned update_dots_lanczos(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                             double *p, double *out, hipStream_t st, bool r)
{
  return update_dots(V, ldv, j, w, n, h, p, out, st, r);
}
template <typename T>
void update_norm(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                 double *partial, double *out, hipStream_t s)
{
  fake_enqueue(s, [=] {
    subtract_now(V, ldv, j, w, n, h);
    double t = 0;
    for (int64_t i = 0; i < n; ++i) t += (double)w[i] * (double)w[i];
    partial[0] = t;
    out[0] = partial[0];
  });
}
// The step-closing arithmetic of the device reductions (k_reduce_epi):
// sums from the partial buffer into out, then the close.
void reduce_step_close(const double *partial, unsigned, int cnt, double *out, int j,
                       const double *hb1, double *Hd, int ldh, double *nrm, double *flags,
                       hipStream_t s)
{
  fake_enqueue(s, [=] {
    std::memcpy(out, partial, sizeof(double) * cnt);
    const double *hb2 = out;
    const int nk = j + 1;
    const double w0 = hb1[nk], w1 = hb2[nk];
    const bool re = !(w1 > 0.5 * w0);
    double hn2 = 0, h22 = 0;
    for (int k = 0; k < nk; ++k) {
      const double hk = hb1[k] + (re ? hb2[k] : 0.0);
      Hd[k + (int64_t)j * ldh] = hk;
      hn2 += hk * hk;
      h22 += hb2[k] * hb2[k];
    }
    const double beta2 = re ? w1 - h22 : w1;
    const double beta = std::sqrt(std::fmax(beta2, 0.0));
    Hd[nk + (int64_t)j * ldh] = beta;
    nrm[0] = beta2;
    if (re) flags[0] += 1.0;
    const bool unsafe = re && !(h22 <= 0.5 * w1);
    if ((unsafe || beta <= 1e-13 * std::sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
  });
}
void reduce_lanczos_coeffs(const double *partial, unsigned, int cnt, double *out, int j,
                           double *Hd, int ldh, double *hL, hipStream_t s)
{
  fake_enqueue(s, [=] {
    std::memcpy(out, partial, sizeof(double) * cnt);
    for (int k = 0; k <= j; ++k) hL[k] = k == j ? out[0] : Hd[j + (int64_t)k * ldh];
  });
}
void reduce_step_close_lanczos(const double *partial, unsigned, int cnt, double *out, int j,
                               const double *hL, double *Hd, int ldh, double *nrm, double *flags,
                               hipStream_t s)
{
  fake_enqueue(s, [=] {
    std::memcpy(out, partial, sizeof(double) * cnt);
    const double *hb2 = out;
    const int nk = j + 1;
    const double w1 = hb2[nk];
    double hn2 = 0, h22 = 0;
    for (int k = 0; k < nk; ++k) {
      const double hk = hL[k] + hb2[k];
      Hd[k + (int64_t)j * ldh] = hk;
      hn2 += hk * hk;
      h22 += hb2[k] * hb2[k];
    }
    const double beta2 = w1 - h22;
    const double beta = std::sqrt(std::fmax(beta2, 0.0));
    Hd[nk + (int64_t)j * ldh] = beta;
    nrm[0] = beta2;
    flags[0] += 1.0;
    const bool unsafe = !(h22 <= 0.5 * w1);
    if ((unsafe || beta <= 1e-13 * std::sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
  });
}
// lagged Lanczos step (k_lag_close, k_lanczos_lag_pk)
void lag_close(const double *partial, unsigned, int j, const double *g, const double *Hd,
               int ldh, const double *nrm, double *hL, double *cf, hipStream_t s)
{
  fake_enqueue(s, [=] {
    const double rz = partial[0];
    const double b2 = nrm[0], b = std::sqrt(std::fmax(b2, 0.0)), invb = 1.0 / b;
    double gc = 0;
    for (int i = 0; i < j; ++i) {
      double c = 0;
      for (int t = 0; t < j; ++t) c += Hd[i + (int64_t)t * ldh] * g[t];
      gc += g[i] * c;
      cf[i] = g[i] * invb;
      cf[72 + i] = c * invb + (i == j - 1 ? b : 0.0);
    }
    const double alpha = (rz - gc) / b2 - 2.0 * g[j - 1];
    for (int i = 0; i <= j; ++i) hL[i] = i == j ? alpha : (i == j - 1 ? b : 0.0);
    cf[144] = alpha + g[j - 1];
    cf[145] = invb;
  });
}
void arnoldi_lag_close(const double *partial, unsigned, int j, const double *g,
                       const double *Hd, int ldh, const double *nrm, double *hL, double *cf,
                       hipStream_t s)
{
  fake_enqueue(s, [=] {
    const double b2 = nrm[0], b = std::sqrt(std::fmax(b2, 0.0)), invb = 1.0 / b;
    double gpz = 0;
    for (int i = 0; i < j; ++i) {
      double c = 0;
      for (int t = 0; t < j; ++t) c += Hd[i + (int64_t)t * ldh] * g[t];
      const double hp = (partial[i] - c) * invb;
      gpz += g[i] * partial[i];
      hL[i] = hp;
      cf[i] = g[i] * invb;
      cf[72 + i] = c * invb + hp;
    }
    const double hj = (partial[j] - gpz) / b2 - g[j - 1];
    hL[j] = hj;
    cf[144] = g[j - 1] + hj;
    cf[145] = invb;
  });
}
template <typename T>
unsigned lanczos_lag(T *V, int64_t ldv, int j, int64_t n, const double *cf, double *partial,
                     hipStream_t s)
{
  fake_enqueue(s, [=] {
    std::vector<double> acc(j + 2, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      double vd = (double)V[j * ldv + i] * cf[145], wd = (double)V[(j + 1) * ldv + i] * cf[145];
      for (int k = 0; k < j; ++k) {
        vd -= cf[k] * (double)V[k * ldv + i];
        wd -= cf[72 + k] * (double)V[k * ldv + i];
      }
      const T vs = (T)vd;
      const T ws = (T)(wd - cf[144] * (double)vs);
      V[j * ldv + i] = vs;
      V[(j + 1) * ldv + i] = ws;
      for (int k = 0; k < j; ++k) acc[k] += (double)V[k * ldv + i] * (double)ws;
      acc[j] += (double)vs * (double)ws;
      acc[j + 1] += (double)ws * (double)ws;
    }
    for (int k = 0; k < j + 2; ++k) partial[k] = acc[k];
  });
  return 1;
}
template <typename T>
void update_normalize(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h2,
                      const double *w0, const double *w1, const double *nrm2, int force,
                      hipStream_t s)
{
  fake_enqueue(s, [=] {
    const bool re = force || !(w1[0] > 0.5 * w0[0]);
    std::vector<double> hh(h2, h2 + j);
    const double inv = 1.0 / std::sqrt(nrm2[0]);
    for (int64_t i = 0; i < n; ++i) {
      double wi = w[i];
      if (re)
        for (int k = 0; k < j; ++k) wi -= hh[k] * (double)V[k * ldv + i];
      w[i] = (T)(wi * inv);
    }
  });
}
template <typename T>
void normalize(T *w, int64_t n, const double *nrm2, hipStream_t s)
{
  fake_enqueue(s, [=] {
    const double inv = 1.0 / std::sqrt(nrm2[0]);
    for (int64_t i = 0; i < n; ++i) w[i] = (T)((double)w[i] * inv);
  });
}
template <typename Ti, typename To>
void gemm_vy(const Ti *Vin, int64_t ldv, int m, const double *Y, int ldy, int k, To *Vout,
             int64_t ldo, int64_t n, hipStream_t s)
{
  TM_CHECK(m <= 64 && k <= m, TM_EINVAL, "restart block too large");
  fake_enqueue(s, [=] {
    std::vector<double> row(m);
    for (int64_t i = 0; i < n; ++i) {
      for (int r = 0; r < m; ++r) row[r] = Vin[r * ldv + i];
      for (int c = 0; c < k; ++c) {
        double t = 0;
        for (int r = 0; r < m; ++r) t += row[r] * Y[r + (int64_t)c * ldy];
        Vout[c * ldo + i] = (To)t;
      }
    }
  });
}
template <typename T>
void random_uniform(T *v, int64_t n, uint64_t seed, hipStream_t s)
{
  fake_enqueue(s, [=] {
    uint64_t x = seed * 0x9e3779b97f4a7c15ull + 1;
    for (int64_t i = 0; i < n; ++i) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      v[i] = (T)(((double)(x >> 11) + 0.5) / 9007199254740992.0);
    }
  });
}
template <typename Ta, typename Tb>
void copy(const Ta *a, Tb *b, int64_t n, hipStream_t s)
{
  fake_enqueue(s, [=] {
    for (int64_t i = 0; i < n; ++i) b[i] = (Tb)a[i];
  });
}
void to_f32(const double *a, float *b, int64_t n, hipStream_t s) { copy(a, b, n, s); }
void fill(double *a, double v, int64_t n, hipStream_t s)
{
  fake_enqueue(s, [=] {
    for (int64_t i = 0; i < n; ++i) a[i] = v;
  });
}
void pow2_scale(const double *x, int64_t n, double *, double *scale, hipStream_t s)
{
  fake_enqueue(s, [=] {
    double m = 0;
    for (int64_t i = 0; i < n; ++i) m = std::fmax(m, std::fabs(x[i]));
    int e = 0;
    if (m > 0) std::frexp(m, &e);
    scale[0] = m > 0 ? std::ldexp(1.0, -e) : 1.0;
  });
}
void gram(const double *A, const double *B, int64_t ld, int k, int64_t n, double *,
          double *G, hipStream_t s)
{
  fake_enqueue(s, [=] {
    for (int r = 0; r < k; ++r)
      for (int c = 0; c < k; ++c) {
        double t = 0;
        for (int64_t i = 0; i < n; ++i) t += A[r * ld + i] * B[c * ld + i];
        G[r * k + c] = t;
      }
  });
}
void gram_sym(const double *U, int64_t ld, int nv, int64_t n, double *p, double *G,
              hipStream_t st)
{
  gram(U, U, ld, nv, n, p, G, st);
}
void combine_norms(const double *U, int64_t ld, int nu, const double *Cm, int ldc, int nc,
                   int64_t n, double *, double *out, hipStream_t s)
{
  fake_enqueue(s, [=] {
    for (int c = 0; c < nc; ++c) {
      double t = 0;
      for (int64_t i = 0; i < n; ++i) {
        double x = 0;
        for (int r = 0; r < nu; ++r) x += Cm[r + (int64_t)c * ldc] * U[r * ld + i];
        t += x * x;
      }
      out[c] = t;
    }
  });
}
template <typename T>
void gram_xy(const T *X, int64_t ld, int nx, const T *Y, int b, int64_t n, double *partial,
             double *G, hipStream_t s)
{
  TM_CHECK(nx >= 1 && nx <= 64 && b >= 1 && b <= 16, TM_EINVAL, "block Gram too large");
  fake_enqueue(s, [=] {
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < nx; ++r) {
        double t = 0;
        for (int64_t i = 0; i < n; ++i) t += (double)X[r * ld + i] * (double)Y[c * ld + i];
        partial[r + c * nx] = t;
      }
  });
  fake_enqueue(s, [=] {
    for (int t = 0; t < nx * b; ++t) G[t] = partial[t];
  });
}
size_t gram_xy_partial_doubles() { return 64 * 16; }
template <typename T>
void block_update(const T *X, int64_t ld, int nx, T *W, int b, const double *H, int ldh,
                  const double *Rinv, int64_t n, hipStream_t s)
{
  fake_enqueue(s, [=] {
    std::vector<double> w(b), o(b);
    for (int64_t i = 0; i < n; ++i) {
      for (int c = 0; c < b; ++c) w[c] = W[c * ld + i];
      if (H && nx > 0)
        for (int r = 0; r < nx; ++r)
          for (int c = 0; c < b; ++c) w[c] -= (double)X[r * ld + i] * H[r + (int64_t)c * ldh];
      for (int c = 0; c < b; ++c) {
        double t = 0;
        for (int q = 0; q <= c; ++q) t += w[q] * (Rinv ? Rinv[q + c * b] : (q == c ? 1.0 : 0.0));
        o[c] = t;
      }
      for (int c = 0; c < b; ++c) W[c * ld + i] = (T)o[c];
    }
  });
}
void block_close(const double *G1, const double *G2, int nc, int b, double *Hd, int ldh, int col0,
                 double *Rinv, double *flags, int step, int mode, hipStream_t s)
{
  fake_enqueue(s, [=] {
    const int nx = nc + b;
    std::vector<double> S(b * b), R(b * b, 0.0), Ri(b * b, 0.0);
    if (mode == 0) {
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < nc; ++r)
          Hd[r + (int64_t)(col0 + c) * ldh] = G1[r + c * nx] + G2[r + c * nx];
      for (int c1 = 0; c1 < b; ++c1)
        for (int c2 = 0; c2 < b; ++c2) {
          double t = G2[nc + c1 + c2 * nx];
          for (int r = 0; r < nc; ++r) t -= G2[r + c1 * nx] * G2[r + c2 * nx];
          S[c1 + c2 * b] = t;
        }
    } else {
      for (int t = 0; t < b * b; ++t) S[t] = G1[t];
    }
    double dmax = 0;
    for (int i = 0; i < b; ++i) dmax = std::fmax(dmax, S[i + i * b]);
    bool ok = dmax > 0;
    for (int j = 0; ok && j < b; ++j) {
      double d = S[j + j * b];
      for (int k = 0; k < j; ++k) d -= R[k + j * b] * R[k + j * b];
      if (!(d > 1e-24 * dmax)) {
        ok = false;
        break;
      }
      d = std::sqrt(d);
      R[j + j * b] = d;
      for (int i = j + 1; i < b; ++i) {
        double t = 0.5 * (S[j + i * b] + S[i + j * b]);
        for (int k = 0; k < j; ++k) t -= R[k + j * b] * R[k + i * b];
        R[j + i * b] = t / d;
      }
    }
    if (!ok) {
      if (flags[1] < 0) flags[1] = step;
      for (int t = 0; t < b * b; ++t) Rinv[t] = (t % (b + 1)) == 0 ? 1.0 : 0.0;
      if (mode == 0)
        for (int c = 0; c < b; ++c)
          for (int r = 0; r < b; ++r) Hd[nc + r + (int64_t)(col0 + c) * ldh] = r == c ? 1.0 : 0.0;
      return;
    }
    for (int j = 0; j < b; ++j) {
      Ri[j + j * b] = 1.0 / R[j + j * b];
      for (int i = j - 1; i >= 0; --i) {
        double t = 0;
        for (int k = i + 1; k <= j; ++k) t += R[i + k * b] * Ri[k + j * b];
        Ri[i + j * b] = -t / R[i + i * b];
      }
    }
    for (int t = 0; t < b * b; ++t) Rinv[t] = Ri[t];
    if (mode == 0) {
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < b; ++r) Hd[nc + r + (int64_t)(col0 + c) * ldh] = R[r + c * b];
    } else {
      std::vector<double> O(b * b);
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < b; ++r) O[r + c * b] = Hd[nc + r + (int64_t)(col0 + c) * ldh];
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < b; ++r) {
          double t = 0;
          for (int k = 0; k < b; ++k) t += R[r + k * b] * O[k + c * b];
          Hd[nc + r + (int64_t)(col0 + c) * ldh] = t;
        }
    }
  });
}
template void gram_xy<double>(const double *, int64_t, int, const double *, int, int64_t, double *,
                              double *, hipStream_t);
template void gram_xy<float>(const float *, int64_t, int, const float *, int, int64_t, double *,
                             double *, hipStream_t);
template void block_update<double>(const double *, int64_t, int, double *, int, const double *,
                                   int, const double *, int64_t, hipStream_t);
template void block_update<float>(const float *, int64_t, int, float *, int, const double *, int,
                                  const double *, int64_t, hipStream_t);
#define FAKE_KRY_INST(T)                                                                        \
  template unsigned multidot<T>(const T *, int64_t, int, const T *, int64_t, double *,         \
                                double *, hipStream_t, bool);                                  \
  template unsigned update_dots<T>(const T *, int64_t, int, T *, int64_t, const double *,       \
                                   double *, double *, hipStream_t, bool);                     \
  template void update_norm<T>(const T *, int64_t, int, T *, int64_t, const double *, double *, \
                               double *, hipStream_t);                                         \
  template unsigned lanczos_lag<T>(T *, int64_t, int, int64_t, const double *, double *,        \
                                   hipStream_t);                                                \
  template unsigned update_dots_lanczos<T>(const T *, int64_t, int, T *, int64_t,               \
                                           const double *, double *, double *, hipStream_t,    \
                                           bool);                                              \
  template void update_normalize<T>(const T *, int64_t, int, T *, int64_t, const double *,      \
                                    const double *, const double *, const double *, int,       \
                                    hipStream_t);                                              \
  template void normalize<T>(T *, int64_t, const double *, hipStream_t);                        \
  template void random_uniform<T>(T *, int64_t, uint64_t, hipStream_t);
FAKE_KRY_INST(double)
FAKE_KRY_INST(float)
template void gemm_vy<double, double>(const double *, int64_t, int, const double *, int, int,
                                      double *, int64_t, int64_t, hipStream_t);
template void gemm_vy<float, float>(const float *, int64_t, int, const double *, int, int,
                                    float *, int64_t, int64_t, hipStream_t);
template void gemm_vy<float, double>(const float *, int64_t, int, const double *, int, int,
                                     double *, int64_t, int64_t, hipStream_t);
template void copy<double, double>(const double *, double *, int64_t, hipStream_t);
template void copy<float, float>(const float *, float *, int64_t, hipStream_t);
template void copy<double, float>(const double *, float *, int64_t, hipStream_t);
template void copy<float, double>(const float *, double *, int64_t, hipStream_t);
}  // namespace kry

// The C ABI's device-only paths have no CPU model.
tm_op *make_op(const tm_params &) { throw Error(TM_EUNSUPPORTED, "no operators in the CPU model"); }
void grad_numerators(tm_op *, const double *, const double *, double *)
{
  throw Error(TM_EUNSUPPORTED, "no gradients in the CPU model");
}
void dlnnf(const tm_op *, double *) {}
void phys_sums(int, const double *, const double *, const uint64_t *, int64_t, int, int, void *,
               double *, hipStream_t)
{
  throw Error(TM_EUNSUPPORTED, "no observables in the CPU model");
}
size_t phys_work_bytes() { return 0; }
void apply_dev_noalloc(tm_op *, const void *, void *, int64_t, int, tm_dtype, tm_dtype, int,
                       hipStream_t)
{
  throw Error(TM_EUNSUPPORTED, "no scratch array in the CPU model");
}
void sample_chains(tm_op *, const double *, const double *, const double *, int, int, uint64_t *,
                   int *)
{
  throw Error(TM_EUNSUPPORTED, "no sampler in the CPU model");
}
}  // namespace ltm

// Dense operator: op->tab64 holds the n x n column-major matrix.  Like the
// device operators, an apply stages its input through the operator's
// workspace op->work (so two applies of one operator on different streams
// share it, as the device operators share theirs).
// Fault injection for the CPU tests: from this many matvecs of an operator
// on, its applies return a perturbed matrix's product (-1: never).
extern "C" {
int g_fake_perturb_after = -1;
}

tm_op::~tm_op()
{
  if (stream) (void)hipStreamDestroy(stream);
}
void tm_op::apply_dev(const void *x, void *y, int64_t ld, int nvec, tm_dtype io, tm_dtype comp,
                      int transpose, const double *sc, hipStream_t s, const ltm::Workspace *ws)
{
  apply_internal(x, y, ld, nvec, io, comp, transpose, sc, s, ws);
}
size_t tm_op::work_bytes(tm_dtype) const { return sizeof(double) * ncol; }
size_t tm_op::perm_bytes(tm_dtype) const { return 0; }
void tm_op::apply_internal(const void *x, void *y, int64_t ld, int nvec, tm_dtype io,
                           tm_dtype comp, int transpose, const double *sc, hipStream_t s,
                           const ltm::Workspace *ws)
{
  const double *A = tab64.as<double>();
  const int64_t n = ncol;
  double *xv;
  if (ws) {
    TM_CHECK(ws->work_bytes >= work_bytes(comp), TM_EINVAL, "workspace too small");
    xv = static_cast<double *>(ws->work);
  } else {
    work.ensure(work_bytes(comp));
    xv = work.as<double>();
  }
  for (int v = 0; v < nvec; ++v) {
    const bool perturb = g_fake_perturb_after >= 0 && matvecs >= g_fake_perturb_after;
    fake_enqueue(s, [=] {
      // stage x (times the input scale) in the workspace
      const double si = sc ? sc[0] : 1.0;
      for (int64_t j = 0; j < n; ++j)
        xv[j] = si * (io == TM_F32 ? (double)static_cast<const float *>(x)[v * ld + j]
                                   : static_cast<const double *>(x)[v * ld + j]);
    });
    fake_enqueue(s, [=] {
      const double so = sc ? 1.0 / sc[0] : 1.0;
      for (int64_t i = 0; i < n; ++i) {
        double t = 0;
        for (int64_t j = 0; j < n; ++j) t += (transpose ? A[j + i * n] : A[i + j * n]) * xv[j];
        if (perturb) t += 0.25 * xv[(i + 1) % n];
        // FP32 compute: round the result to single precision
        if (comp == TM_F32) t = (double)(float)t;
        t *= so;
        if (io == TM_F32) static_cast<float *>(y)[v * ld + i] = (float)t;
        else static_cast<double *>(y)[v * ld + i] = t;
      }
    });
    ++matvecs;
  }
}
