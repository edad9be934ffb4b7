// tm_krylov.hip -- Krylov-basis kernels for the device eigensolver (basis
// stored in FP64, or in FP32 for the FP32 phase of a mixed solve; FP64
// arithmetic throughout).
//
// The basis V is column-major (vector k at V + k*ldv).  Every reduction is
// two-level and deterministic: a fixed grid of NBLK workgroups writes
// partial sums, a single workgroup adds them in a fixed order.  Dot products
// use wavefront (64-lane) shuffles, then LDS across the 4 wavefronts.
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2): the
// update kernel subtracts V h and, in the same sweep over the data, already
// accumulates the dot products needed by the next correction.
#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <type_traits>

#include "tm_device.hpp"

namespace ltm {
namespace kry {
namespace {

constexpr int THREADS = 256;
constexpr int MAXJ = 64;  // largest basis handled by the fused kernels

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum of p[b * stride] over b = lane, lane + 64, ... < nblk into lane 0 of
// the wavefront: four independent accumulators per lane (loads in flight),
// combined in a fixed order, then a fixed shuffle tree -- deterministic.
__device__ __forceinline__ double lane_sum(const double *__restrict__ p, int nblk, int64_t stride,
                                           int lane)
{
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int b = lane;
  for (; b + 192 < nblk; b += 256) {
    s0 += p[(int64_t)b * stride];
    s1 += p[(int64_t)(b + 64) * stride];
    s2 += p[(int64_t)(b + 128) * stride];
    s3 += p[(int64_t)(b + 192) * stride];
  }
  for (; b < nblk; b += 64) s0 += p[(int64_t)b * stride];
  return wave_sum((s0 + s1) + (s2 + s3));
}

// Block-reduce the first `cnt` of the J values acc[] (compile-time indexed,
// so the array stays in registers) and write slot k of this workgroup to
// partial[k * NBLK + blockIdx.x]: each slot's partial sums are contiguous,
// so the reduction kernels read them with coalesced loads (grids of at most
// NBLK workgroups).
template <int J>
__device__ __forceinline__ void block_reduce_write(const double (&acc)[J], int cnt, double *partial)
{
  double *out = partial + blockIdx.x;
  __shared__ double red[THREADS / 64][J];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < J; ++k) {
    if (k < cnt) {
      const double v = wave_sum(acc[k]);
      if (lane == 0) red[wid][k] = v;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
    double s = 0;
    for (int w = 0; w < THREADS / 64; ++w) s += red[w][k];
    out[(int64_t)k * NBLK] = s;
  }
}

// Basis vectors are stored as T = double (FP64 phase) or float (the FP32
// phase of the mixed-precision solve); every product and sum is formed in
// FP64 either way.  The packed kernels move 16 bytes per lane and element
// group: Pk<T>::P elements (2 doubles or 4 floats); they need V, w 16-byte
// aligned and ldv a multiple of P, and block 0 thread 0 takes the n % P tail.
template <typename T> struct Pk;
template <> struct Pk<double> {
  typedef double V __attribute__((ext_vector_type(2)));
  static constexpr int P = 2;
};
template <> struct Pk<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static constexpr int P = 4;
};

// Waves per SIMD the dot sweeps are compiled for: the compiler keeps every
// basis load of an element group in flight, so registers grow with J; cap
// them where the loads still fit without spilling.
constexpr int sweep_waves(int J) { return J <= 8 ? 8 : (J <= 12 ? 6 : 1); }

// sum_e a[e] b[e] in FP64
template <typename PV, int P>
__device__ __forceinline__ double pdot(const PV &a, const PV &b)
{
  double d = (double)a[0] * (double)b[0];
#pragma unroll
  for (int e = 1; e < P; ++e) d += (double)a[e] * (double)b[e];
  return d;
}

// out slots: <V_k, w> for k < j, then <w, w> in slot j
template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_multidot(const T *__restrict__ V, int64_t ldv,
                                                      int j, const T *__restrict__ w,
                                                      int64_t n, double *__restrict__ partial)
{
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    const double wi = w[i];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) acc[k] += (double)V[k * ldv + i] * wi;
    acc[J] += wi * wi;
  }
  // move <w,w> into slot j
#pragma unroll
  for (int k = 0; k < J; ++k)
    if (k == j) acc[k] = acc[J];
  block_reduce_write<J + 1>(acc, j + 1, partial);
}

// w -= sum_k h[k] V_k; then (dots=1) slots <V_k, w>, <w,w> or (dots=0) slot 0 = <w,w>
// (the products use w as stored)
template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_update_dots(const T *__restrict__ V,
                                                         int64_t ldv, int j,
                                                         T *__restrict__ w, int64_t n,
                                                         const double *__restrict__ h,
                                                         double *__restrict__ partial, int dots)
{
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = k < j ? h[k] : 0.0;
  __syncthreads();
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double vk[J];
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        vk[k] = V[k * ldv + i];
        wi -= hk[k] * vk[k];
      }
    const T ws = (T)wi;
    w[i] = ws;
    wi = ws;
    if (dots) {
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += vk[k] * wi;
    }
    acc[J] += wi * wi;
  }
  if (dots) {
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k == j) acc[k] = acc[J];
    block_reduce_write<J + 1>(acc, j + 1, partial);
  } else {
    acc[0] = acc[J];
    block_reduce_write<J + 1>(acc, 1, partial);
  }
}

// One wavefront per output slot: lane l adds partials l, l+64, ... in a fixed
// order, then a fixed shuffle tree -> deterministic.
__global__ __launch_bounds__(64) void k_reduce_partials(const double *__restrict__ partial,
                                                        int nblk, int cnt,
                                                        double *__restrict__ out)
{
  const int k = blockIdx.x;
  const double s = lane_sum(partial + (int64_t)k * NBLK, nblk, 1, threadIdx.x);
  if (threadIdx.x == 0) out[k] = s;
}

// ---- packed variants of the same kernels (same two-level reductions)
template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_multidot_pk(const T *__restrict__ V, int64_t ldv,
                                                         int j, const T *__restrict__ w,
                                                         int64_t n, double *__restrict__ partial)
{
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  const PV *__restrict__ Vp = reinterpret_cast<const PV *>(V);
  const PV *__restrict__ wp = reinterpret_cast<const PV *>(w);
  const int64_t ldp = ldv / P, np = n / P;
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    const PV wi = wp[p];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) acc[k] += pdot<PV, P>(Vp[k * ldp + p], wi);
    acc[J] += pdot<PV, P>(wi, wi);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      const double wi = w[i];
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += (double)V[k * ldv + i] * wi;
      acc[J] += wi * wi;
    }
#pragma unroll
  for (int k = 0; k < J; ++k)
    if (k == j) acc[k] = acc[J];
  block_reduce_write<J + 1>(acc, j + 1, partial);
}

template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_update_dots_pk(const T *__restrict__ V,
                                                            int64_t ldv, int j,
                                                            T *__restrict__ w, int64_t n,
                                                            const double *__restrict__ h,
                                                            double *__restrict__ partial, int dots,
                                                            const double *w0p, const double *w1p)
{
  // conditional form (w0p != null): only when the first projection removed
  // more than half of the norm^2; the whole grid takes the same branch
  if (w0p && (w1p[0] > 0.5 * w0p[0])) return;
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = k < j ? h[k] : 0.0;
  __syncthreads();
  const PV *__restrict__ Vp = reinterpret_cast<const PV *>(V);
  PV *__restrict__ wp = reinterpret_cast<PV *>(w);
  const int64_t ldp = ldv / P, np = n / P;
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    PV vk[J];
    const PV wi = wp[p];
    double wd[P];
#pragma unroll
    for (int e = 0; e < P; ++e) wd[e] = wi[e];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        vk[k] = Vp[k * ldp + p];
#pragma unroll
        for (int e = 0; e < P; ++e) wd[e] -= hk[k] * (double)vk[k][e];
      }
    PV wo;
#pragma unroll
    for (int e = 0; e < P; ++e) wo[e] = (T)wd[e];
    wp[p] = wo;
    if (dots) {
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += pdot<PV, P>(vk[k], wo);
    }
    acc[J] += pdot<PV, P>(wo, wo);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      double vk[J];
      double wi = w[i];
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) {
          vk[k] = V[k * ldv + i];
          wi -= hk[k] * vk[k];
        }
      const T ws = (T)wi;
      w[i] = ws;
      wi = ws;
      if (dots) {
#pragma unroll
        for (int k = 0; k < J; ++k)
          if (k < j) acc[k] += vk[k] * wi;
      }
      acc[J] += wi * wi;
    }
  if (dots) {
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k == j) acc[k] = acc[J];
    block_reduce_write<J + 1>(acc, j + 1, partial);
  } else {
    acc[0] = acc[J];
    block_reduce_write<J + 1>(acc, 1, partial);
  }
}

// Lanczos form of k_update_dots_pk (dots = 1): h is zero except in its last
// two slots (beta_{j-1}, alpha), so w is updated from the last two basis
// vectors first and the other vectors are streamed straight into the dot
// products (no register array of J vectors).  Same results as the general
// kernel: the skipped terms subtract exact zeros.
template <int J, typename T>
__global__ __launch_bounds__(THREADS, sweep_waves(J)) void k_lanczos_dots_pk(const T *__restrict__ V,
                                                             int64_t ldv, int j,
                                                             T *__restrict__ w, int64_t n,
                                                             const double *__restrict__ h,
                                                             double *__restrict__ partial)
{
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  const double ha = j >= 2 ? h[j - 2] : 0.0, hb = h[j - 1];
  const PV *__restrict__ Vp = reinterpret_cast<const PV *>(V);
  PV *__restrict__ wp = reinterpret_cast<PV *>(w);
  const int64_t ldp = ldv / P, np = n / P;
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    const PV wi = wp[p];
    const PV va = j >= 2 ? Vp[(j - 2) * ldp + p] : PV{};
    const PV vb = Vp[(j - 1) * ldp + p];
    PV wo;
#pragma unroll
    for (int e = 0; e < P; ++e) {
      double x = wi[e];
      if (j >= 2) x -= ha * (double)va[e];
      x -= hb * (double)vb[e];
      wo[e] = (T)x;
    }
    wp[p] = wo;
#pragma unroll
    for (int k = 0; k < J; ++k) {
      if (k < j - 2) acc[k] += pdot<PV, P>(Vp[k * ldp + p], wo);
      else if (k == j - 2) acc[k] += pdot<PV, P>(va, wo);
      else if (k == j - 1) acc[k] += pdot<PV, P>(vb, wo);
    }
    acc[J] += pdot<PV, P>(wo, wo);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      double x = w[i];
      if (j >= 2) x -= ha * (double)V[(j - 2) * ldv + i];
      x -= hb * (double)V[(j - 1) * ldv + i];
      const T ws = (T)x;
      w[i] = ws;
      x = ws;
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += (double)V[k * ldv + i] * x;
      acc[J] += x * x;
    }
#pragma unroll
  for (int k = 0; k < J; ++k)
    if (k == j) acc[k] = acc[J];
  block_reduce_write<J + 1>(acc, j + 1, partial);
}

template <typename T>
__global__ void k_normalize_pk(T *__restrict__ w, int64_t n, const double *__restrict__ nrm2)
{
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  const double inv = 1.0 / sqrt(nrm2[0]);
  PV *__restrict__ wp = reinterpret_cast<PV *>(w);
  const int64_t np = n / P;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * blockDim.x) {
    PV v = wp[p];
#pragma unroll
    for (int e = 0; e < P; ++e) v[e] = (T)((double)v[e] * inv);
    wp[p] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) w[i] = (T)((double)w[i] * inv);
}

// Close a Krylov step in one sweep: w <- (w - c V h2) / beta, with c = 1
// when the DGKS correction is due (w1 <= w0 / 2) and beta^2 = nrm2[0] as set
// by the step close (w1 - |h2|^2 after a correction, w1 otherwise).
template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_update_normalize(const T *__restrict__ V,
                                                              int64_t ldv, int j,
                                                              T *__restrict__ w, int64_t n,
                                                              const double *__restrict__ h2,
                                                              const double *w0p, const double *w1p,
                                                              const double *__restrict__ nrm2,
                                                              int force)
{
  const bool re = force || !(w1p[0] > 0.5 * w0p[0]);
  const double inv = 1.0 / sqrt(nrm2[0]);
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = (re && k < j) ? h2[k] : 0.0;
  __syncthreads();
  const int jj = re ? j : 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < jj) wi -= hk[k] * (double)V[k * ldv + i];
    w[i] = (T)(wi * inv);
  }
}

template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_update_normalize_pk(const T *__restrict__ V,
                                                                 int64_t ldv, int j,
                                                                 T *__restrict__ w, int64_t n,
                                                                 const double *__restrict__ h2,
                                                                 const double *w0p,
                                                                 const double *w1p,
                                                                 const double *__restrict__ nrm2,
                                                                 int force)
{
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  const bool re = force || !(w1p[0] > 0.5 * w0p[0]);
  const double inv = 1.0 / sqrt(nrm2[0]);
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = (re && k < j) ? h2[k] : 0.0;
  __syncthreads();
  const int jj = re ? j : 0;
  const PV *__restrict__ Vp = reinterpret_cast<const PV *>(V);
  PV *__restrict__ wp = reinterpret_cast<PV *>(w);
  const int64_t ldp = ldv / P, np = n / P;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    const PV wi = wp[p];
    double wd[P];
#pragma unroll
    for (int e = 0; e < P; ++e) wd[e] = wi[e];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < jj) {
        const PV v = Vp[k * ldp + p];
#pragma unroll
        for (int e = 0; e < P; ++e) wd[e] -= hk[k] * (double)v[e];
      }
    PV wo;
#pragma unroll
    for (int e = 0; e < P; ++e) wo[e] = (T)(wd[e] * inv);
    wp[p] = wo;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      double wi = w[i];
      for (int k = 0; k < jj; ++k) wi -= hk[k] * (double)V[k * ldv + i];
      w[i] = (T)(wi * inv);
    }
}

// One thread closes step j from the two projections alone: column j of the
// projected matrix = h1 + c h2 (c as above), beta^2 = w1 - |h2|^2 after a
// correction (Pythagoras; V orthonormal) or w1; nrm[0] = beta^2.  flags[0]
// counts corrections; flags[1] records the first step j whose beta is below
// 1e-13 |h| or whose correction removed more than half of w1 again (a
// numerically invariant subspace: the host redoes the cycle from j with
// explicit norms).
// Partial sums of slots 0..cnt-1 (one wavefront per slot, the same order as
// k_reduce_partials) into out[], then the first wavefront closes the step
// from them, one lane per projected-matrix row (values through LDS, the
// norms by a wavefront sum):
//   EPI 1, Lanczos coefficients: a[0..j] = hL = (row j of Hd left of the
//     diagonal, alpha) with alpha = <v_j, T v_j> = out[0]: T v_j couples to
//     v_{j+1}, v_j and the vectors whose columns reach row j -- v_{j-1}
//     (beta_{j-1} = Hd[j, j-1]) inside a cycle, every kept Ritz vector (the
//     restart's coupling row) at its first step; the rest of the row is 0;
//   EPI 2, general step (a = hb1, the first projection, out = hb2, the
//     second): the DGKS correction is due when |w1|^2 <= |w0|^2 / 2; column j
//     = hb1 + (correction ? hb2 : 0), beta^2 = |w1|^2 - |hb2|^2 after a
//     correction, |w1|^2 otherwise; unsafe when the correction removed more
//     than half of |w1|^2 again;
//   EPI 3, Lanczos step (a = hL, out = hb2; the full reorthogonalisation
//     always runs): column j = hL + hb2, beta^2 = |w1|^2 - |hb2|^2, unsafe as
//     for EPI 2.
// flags[0] counts corrections; flags[1] takes the first step whose beta
// is tiny (<= 1e-13 |column|), not positive, or unsafe.  nrm[0] = beta^2.
template <int EPI>
__global__ __launch_bounds__(1024) void k_reduce_epi(const double *__restrict__ partial, int nblk,
                                                     int cnt, double *__restrict__ out, int j,
                                                     double *__restrict__ a,
                                                     double *__restrict__ Hd, int ldh,
                                                     double *__restrict__ nrm,
                                                     double *__restrict__ flags)
{
  __shared__ double sv[MAXJ + 2], sa[MAXJ + 2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nk = j + 1;
  if (EPI != 1)
    for (int k = threadIdx.x; k < (EPI == 2 ? nk + 1 : nk); k += blockDim.x) sa[k] = a[k];
  for (int k = wid; k < cnt; k += nw) {
    const double s = lane_sum(partial + (int64_t)k * NBLK, nblk, 1, lane);
    if (lane == 0) {
      out[k] = s;
      sv[k] = s;
    }
  }
  __syncthreads();
  if (wid != 0) return;
  if (EPI == 1) {
    for (int k = lane; k <= j; k += 64) a[k] = k == j ? sv[0] : Hd[j + (int64_t)k * ldh];
    return;
  }
  const double w1 = sv[nk];
  const bool re = EPI == 3 || !(w1 > 0.5 * sa[nk]);
  double hn2 = 0, h22 = 0;
  for (int k = lane; k < nk; k += 64) {
    const double hk = sa[k] + (re ? sv[k] : 0.0);
    Hd[k + (int64_t)j * ldh] = hk;
    hn2 += hk * hk;
    h22 += sv[k] * sv[k];
  }
  hn2 = wave_sum(hn2);
  h22 = wave_sum(h22);
  if (lane != 0) return;
  const double beta2 = re ? w1 - h22 : w1;
  const double beta = sqrt(fmax(beta2, 0.0));
  Hd[nk + (int64_t)j * ldh] = beta;
  nrm[0] = beta2;
  if (re) flags[0] += 1.0;
  const bool unsafe = re && !(h22 <= 0.5 * w1);
  if ((unsafe || beta <= 1e-13 * sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
}

// Lagged Lanczos step j.  The previous step left its three-term residual r
// in slot j unnormalised and uncorrected, with its second projection g
// (g_k = <P_k, r>, P = V_0 .. V_{j-1}) and beta^2 = nrm[0] = |r - P g|^2;
// T was applied to r itself (z = T r in slot j+1).  Then
//   v_j = (r - P g) / beta,
// and the Krylov relation T P = P H + beta v_j e_{j-1}^T (H = the projected
// matrix so far, Hd) gives
//   T v_j = (z - P c) / beta - g_{j-1} v_j,   c = H_{0..j-1, 0..j-1} g,
//   alpha = <v_j, T v_j> = (<r, z> - <g, c>) / beta^2 - 2 g_{j-1}
// (T symmetric: <P_k, z> = <T P_k, r>).  One wavefront sums <r, z> (slot 0
// of the partial blocks, cnt slots each) and writes hL[0..j] = (0 .., beta,
// alpha) for the step close and the sweep coefficients
//   cf[k] = g_k / beta, cf[LAG_LD + k] = c_k / beta + beta [k = j-1] (k < j),
//   cf[2 LAG_LD] = alpha + g_{j-1}, cf[2 LAG_LD + 1] = 1 / beta.
// An error in alpha or c only moves weight into the step's second
// projection, which the full reorthogonalisation measures and the close adds
// to the projected-matrix column.
constexpr int LAG_LD = 72;
__global__ __launch_bounds__(64) void k_lag_close(const double *__restrict__ partial, int nblk,
                                                  int cnt, int j, const double *__restrict__ g,
                                                  const double *__restrict__ Hd, int ldh,
                                                  const double *__restrict__ nrm,
                                                  double *__restrict__ hL, double *__restrict__ cf)
{
  const int lane = threadIdx.x;
  // H_{0..j-1, 0..j-1} and g staged in LDS by parallel loads (j <= LAG_MAXJ)
  __shared__ double sh[LAG_MAXJ * LAG_MAXJ], sg[LAG_MAXJ];
  for (int e = lane; e < j * j; e += 64) sh[e] = Hd[(e % j) + (int64_t)(e / j) * ldh];
  for (int e = lane; e < j; e += 64) sg[e] = g[e];
  const double rz = __shfl(lane_sum(partial, nblk, 1, lane), 0, 64);
  const double b2 = nrm[0], b = sqrt(fmax(b2, 0.0)), invb = 1.0 / b;
  __syncthreads();
  double gc = 0;
  for (int i = lane; i < j; i += 64) {
    double c = 0;
    for (int t = 0; t < j; ++t) c += sh[i + t * j] * sg[t];
    gc += g[i] * c;
    cf[i] = g[i] * invb;
    cf[LAG_LD + i] = c * invb + (i == j - 1 ? b : 0.0);
  }
  gc = __shfl(wave_sum(gc), 0, 64);
  const double gl = g[j - 1];
  const double alpha = (rz - gc) / b2 - 2.0 * gl;
  for (int i = lane; i <= j; i += 64) hL[i] = i == j ? alpha : (i == j - 1 ? b : 0.0);
  if (lane == 0) {
    cf[2 * LAG_LD] = alpha + gl;
    cf[2 * LAG_LD + 1] = invb;
  }
}

// Lagged Arnoldi step j (general T; the previous step left r, g and beta^2 as
// for k_lag_close).  The first projection of T v_j is not known beforehand:
// a multidot of [P | r] against z = T r gives the slots <P_k, z> (k < j),
// <r, z> (slot j) and <z, z> (slot j+1), and with c = H_{0..j-1, 0..j-1} g
// and V orthonormal
//   <P_k, T v_j> = (<P_k, z> - c_k) / beta,
//   <v_j, T v_j> = (<r, z> - <g, P^T z>) / beta^2 - g_{j-1}.
// Writes that projection to hL[0..j] (the step close adds the second one)
// and the sweep coefficients of k_lanczos_lag_pk: cf[k] = g_k / beta,
// cf[LAG_LD + k] = c_k / beta + hL[k], cf[2 LAG_LD] = g_{j-1} + hL[j],
// cf[2 LAG_LD + 1] = 1 / beta.  Its waves sum the j+2 slots in parallel;
// the first wavefront closes.
__global__ __launch_bounds__(1024) void k_arnoldi_lag_close(const double *__restrict__ partial,
                                                            int nblk, int j,
                                                            const double *__restrict__ g,
                                                            const double *__restrict__ Hd,
                                                            int ldh,
                                                            const double *__restrict__ nrm,
                                                            double *__restrict__ hL,
                                                            double *__restrict__ cf)
{
  __shared__ double sv[LAG_MAXJ + 2], sh[LAG_MAXJ * LAG_MAXJ], sg[LAG_MAXJ];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int e = threadIdx.x; e < j * j; e += blockDim.x)
    sh[e] = Hd[(e % j) + (int64_t)(e / j) * ldh];
  for (int e = threadIdx.x; e < j; e += blockDim.x) sg[e] = g[e];
  for (int k = wid; k < j + 2; k += nw) {
    const double sum = lane_sum(partial + (int64_t)k * NBLK, nblk, 1, lane);
    if (lane == 0) sv[k] = sum;
  }
  __syncthreads();
  if (wid != 0) return;
  const double b2 = nrm[0], b = sqrt(fmax(b2, 0.0)), invb = 1.0 / b;
  double gpz = 0;
  for (int i = lane; i < j; i += 64) {
    double c = 0;
    for (int t = 0; t < j; ++t) c += sh[i + t * j] * sg[t];
    const double hp = (sv[i] - c) * invb;
    gpz += sg[i] * sv[i];
    hL[i] = hp;
    cf[i] = sg[i] * invb;
    cf[LAG_LD + i] = c * invb + hp;
  }
  gpz = __shfl(wave_sum(gpz), 0, 64);
  const double gl = sg[j - 1];
  const double hj = (sv[j] - gpz) / b2 - gl;
  if (lane == 0) {
    hL[j] = hj;
    cf[2 * LAG_LD] = gl + hj;
    cf[2 * LAG_LD + 1] = invb;
  }
}

// The sweep of a lagged Lanczos step j (coefficients from k_lag_close): per
// element, from slot j (r), slot j+1 (z) and P,
//   v_j = r / beta - sum_k (g_k / beta) P_k                 -> slot j
//   w   = z / beta - sum_k d_k P_k - (alpha + g_{j-1}) v_j  -> slot j+1
// (w = T v_j - alpha v_j - beta v_{j-1}, the next three-term residual), then
// the slots <P_k, w> (k < j), <v_j, w>, <w, w> of the full
// reorthogonalisation, as k_lanczos_dots_pk writes them.  One read of P,
// r and z replaces the correction sweep of step j-1 and the dot sweep of
// step j; each P_k of an element group stays in registers between its two
// uses.  The dots use v_j and w as stored.
template <int J, typename T>
__global__ __launch_bounds__(THREADS) void k_lanczos_lag_pk(T *__restrict__ V, int64_t ldv, int j,
                                                            int64_t n,
                                                            const double *__restrict__ cf,
                                                            double *__restrict__ partial)
{
  using PV = typename Pk<T>::V;
  constexpr int P = Pk<T>::P;
  // the coefficients are uniform: scalar loads into scalar registers
  const double *__restrict__ sg = cf;
  const double *__restrict__ sd = cf + LAG_LD;
  const double ca = cf[2 * LAG_LD], invb = cf[2 * LAG_LD + 1];
  PV *__restrict__ Vp = reinterpret_cast<PV *>(V);
  const int64_t ldp = ldv / P, np = n / P;
  double acc[J + 2];
#pragma unroll
  for (int k = 0; k < J + 2; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    PV vk[J];
    const PV r = Vp[j * ldp + p], z = Vp[(j + 1) * ldp + p];
    double vd[P], wd[P];
#pragma unroll
    for (int e = 0; e < P; ++e) {
      vd[e] = (double)r[e] * invb;
      wd[e] = (double)z[e] * invb;
    }
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        vk[k] = Vp[k * ldp + p];
#pragma unroll
        for (int e = 0; e < P; ++e) {
          vd[e] -= sg[k] * (double)vk[k][e];
          wd[e] -= sd[k] * (double)vk[k][e];
        }
      }
    PV vo, wo;
#pragma unroll
    for (int e = 0; e < P; ++e) vo[e] = (T)vd[e];
#pragma unroll
    for (int e = 0; e < P; ++e) wo[e] = (T)(wd[e] - ca * (double)vo[e]);
    Vp[j * ldp + p] = vo;
    Vp[(j + 1) * ldp + p] = wo;
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) acc[k] += pdot<PV, P>(vk[k], wo);
    acc[J] += pdot<PV, P>(vo, wo);
    acc[J + 1] += pdot<PV, P>(wo, wo);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      double vk[J];
      double vd = (double)V[j * ldv + i] * invb, wd = (double)V[(j + 1) * ldv + i] * invb;
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) {
          vk[k] = V[k * ldv + i];
          vd -= sg[k] * vk[k];
          wd -= sd[k] * vk[k];
        }
      const T vs = (T)vd;
      const T ws = (T)(wd - ca * (double)vs);
      V[j * ldv + i] = vs;
      V[(j + 1) * ldv + i] = ws;
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += vk[k] * (double)ws;
      acc[J] += (double)vs * (double)ws;
      acc[J + 1] += (double)ws * (double)ws;
    }
  // slots 0 .. j-1 stay; <v_j, w> into slot j, <w, w> into slot j+1 (j <= J)
  const double a1 = acc[J], a2 = acc[J + 1];
#pragma unroll
  for (int k = 0; k < J + 2; ++k) {
    if (k == j) acc[k] = a1;
    else if (k == j + 1) acc[k] = a2;
  }
  block_reduce_write<J + 2>(acc, j + 2, partial);
}

template <typename T>
__global__ void k_normalize(T *__restrict__ w, int64_t n, const double *__restrict__ nrm2)
{
  const double inv = 1.0 / sqrt(nrm2[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = (T)((double)w[i] * inv);
}

// Vout may be Vin (the restart rotates the basis in place), so Vin is not
// __restrict__: every read of row i happens before any write of row i.
template <int M, typename Ti, typename To>
__global__ __launch_bounds__(THREADS) void k_gemm_vy(const Ti *Vin, int64_t ldv,
                                                     int m, const double *__restrict__ Y,
                                                     int ldy, int k, To *Vout, int64_t ldo,
                                                     int64_t n)
{
  __shared__ double sY[M * M];
  for (int t = threadIdx.x; t < m * k; t += blockDim.x) {
    const int r = t % m, c = t / m;
    sY[r + c * M] = Y[r + (int64_t)c * ldy];
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = r < m ? (double)Vin[r * ldv + i] : 0.0;
    // all reads of row i happen before any write of row i (Vout may alias)
    for (int c = 0; c < k; ++c) {
      double s = 0;
#pragma unroll
      for (int r = 0; r < M; ++r) s += v[r] * sY[r + c * M];
      Vout[c * ldo + i] = (To)s;
    }
  }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// uniform(0,1) from a counter-based hash of (seed, index)
// k_gemm_vy over element groups: Pk<Ti>::P consecutive elements per thread
// with one 16-byte load per input column (Vin, Vout 16-byte aligned, ldv and
// ldo multiples of P); block 0 thread 0 takes the n % P tail.  Vout may
// alias Vin as above.
template <int M, typename Ti, typename To>
__global__ __launch_bounds__(THREADS) void k_gemm_vy_pk(const Ti *Vin, int64_t ldv, int m,
                                                        const double *__restrict__ Y, int ldy,
                                                        int k, To *Vout, int64_t ldo, int64_t n)
{
  using PV = typename Pk<Ti>::V;
  constexpr int P = Pk<Ti>::P;
  __shared__ double sY[M * M];
  for (int t = threadIdx.x; t < m * k; t += blockDim.x) {
    const int r = t % m, c = t / m;
    sY[r + c * M] = Y[r + (int64_t)c * ldy];
  }
  __syncthreads();
  const PV *Vp = reinterpret_cast<const PV *>(Vin);
  const int64_t ldp = ldv / P, np = n / P;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < np;
       p += (int64_t)gridDim.x * THREADS) {
    PV v[M];
#pragma unroll
    for (int r = 0; r < M; ++r) v[r] = r < m ? Vp[r * ldp + p] : PV{};
    for (int c = 0; c < k; ++c) {
      double acc[P];
#pragma unroll
      for (int e = 0; e < P; ++e) acc[e] = 0;
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const double y = sY[r + c * M];
#pragma unroll
        for (int e = 0; e < P; ++e) acc[e] += (double)v[r][e] * y;
      }
      To *o = Vout + c * ldo + p * P;
      if constexpr (sizeof(To) * P == 16) {
        typename Pk<To>::V ov;
#pragma unroll
        for (int e = 0; e < P; ++e) ov[e] = (To)acc[e];
        *reinterpret_cast<typename Pk<To>::V *>(o) = ov;
      } else {
        // float in, double out: two 16-byte stores
        reinterpret_cast<double2 *>(o)[0] = make_double2(acc[0], acc[1]);
        reinterpret_cast<double2 *>(o)[1] = make_double2(acc[2], acc[3]);
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t i = np * P; i < n; ++i) {
      double v[M];
#pragma unroll
      for (int r = 0; r < M; ++r) v[r] = r < m ? (double)Vin[r * ldv + i] : 0.0;
      for (int c = 0; c < k; ++c) {
        double s = 0;
#pragma unroll
        for (int r = 0; r < M; ++r) s += v[r] * sY[r + c * M];
        Vout[c * ldo + i] = (To)s;
      }
    }
}

template <typename T>
__global__ void k_random(T *v, int64_t n, uint64_t seed)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = splitmix64(splitmix64(seed) ^ (uint64_t)i);
    v[i] = (T)(((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0));
  }
}

template <typename Ta, typename Tb>
__global__ void k_copy(const Ta *a, Tb *b, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (Tb)a[i];
}

__global__ void k_to_f32(const double *a, float *b, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_fill(double *a, double v, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ __launch_bounds__(THREADS) void k_maxabs(const double *x, int64_t n, double *partial)
{
  double m = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS)
    m = fmax(m, fabs(x[i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  __shared__ double red[THREADS / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0;
    for (int w = 0; w < THREADS / 64; ++w) r = fmax(r, red[w]);
    partial[blockIdx.x] = r;
  }
}

__global__ void k_pow2_scale(const double *partial, int nblk, double *scale)
{
  if (threadIdx.x == 0) {
    double m = 0;
    for (int b = 0; b < nblk; ++b) m = fmax(m, partial[b]);
    int e = 0;
    if (m > 0) frexp(m, &e);
    scale[0] = m > 0 ? ldexp(1.0, -e) : 1.0;
  }
}

template <int J, typename T = double>
__global__ __launch_bounds__(THREADS) void k_gram_row(const double *__restrict__ A,
                                                      const double *__restrict__ B, int64_t ld,
                                                      int k, int64_t n, int row,
                                                      double *__restrict__ partial)
{
  double acc[J];
#pragma unroll
  for (int c = 0; c < J; ++c) acc[c] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    const double a = A[row * ld + i];
#pragma unroll
    for (int c = 0; c < J; ++c)
      if (c < k) acc[c] += a * B[c * ld + i];
  }
  block_reduce_write<J>(acc, k, partial);
}

// Rows [r0, r0 + RB) of the Gram matrix in one sweep: each vector of A's
// row block and of B is read once.  Partial slot r * J + c.
template <int RB, int J>
__global__ __launch_bounds__(THREADS) void k_gram_block(const double *__restrict__ A,
                                                        const double *__restrict__ B, int64_t ld,
                                                        int k, int64_t n, int r0, int nr,
                                                        double *__restrict__ partial)
{
  double acc[RB * J];
#pragma unroll
  for (int t = 0; t < RB * J; ++t) acc[t] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double a[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) a[r] = r < nr ? A[(r0 + r) * ld + i] : 0.0;
#pragma unroll
    for (int c = 0; c < J; ++c)
      if (c < k) {
        const double b = B[c * ld + i];
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r * J + c] += a[r] * b;
      }
  }
  block_reduce_write<RB * J>(acc, RB * J, partial);
}

// out[r * k + c] = sum over blocks of slot r * J + c, for r < nr, c < k (one
// wavefront per output, fixed order)
__global__ __launch_bounds__(64) void k_reduce_gram(const double *__restrict__ partial, int nblk,
                                                    int stride, int J, int k,
                                                    double *__restrict__ out)
{
  const int r = blockIdx.x / k, c = blockIdx.x % k;
  const double s = lane_sum(partial + (int64_t)(r * J + c) * NBLK, nblk, 1, threadIdx.x);
  if (threadIdx.x == 0) out[r * k + c] = s;
}

// Gram matrix of NV <= 10 vectors U_0..U_{nv-1} (columns at U + i*ld) in one
// sweep: packed lower triangle, slot a(a+1)/2 + b for b <= a.
template <int NV>
__global__ __launch_bounds__(THREADS) void k_gram_tri(const double *__restrict__ U, int64_t ld,
                                                      int nv, int64_t n,
                                                      double *__restrict__ partial)
{
  constexpr int NT = NV * (NV + 1) / 2;
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double u[NV];
#pragma unroll
    for (int a = 0; a < NV; ++a) u[a] = a < nv ? U[a * ld + i] : 0.0;
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) acc[a * (a + 1) / 2 + b] += u[a] * u[b];
  }
  block_reduce_write<NT>(acc, NT, partial);
}

// G[a * nv + b] = G[b * nv + a] = sum over blocks of triangle slot (a, b)
__global__ __launch_bounds__(64) void k_reduce_tri(const double *__restrict__ partial, int nblk,
                                                   int nt, int nv, double *__restrict__ G)
{
  const int t = blockIdx.x;
  int a = 0;
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  const int b = t - a * (a + 1) / 2;
  const double s = lane_sum(partial + (int64_t)t * NBLK, nblk, 1, threadIdx.x);
  if (threadIdx.x == 0 && a < nv) {
    G[a * nv + b] = s;
    G[b * nv + a] = s;
  }
}

// out[c0 + c] = || sum_r C[r + (c0 + c) * ldc] U_r ||^2 for c < nc <= 8, r < nu
// <= NU: combinations of the columns of U and their squared norms in one
// sweep, without writing the combinations.
template <int NU>
__global__ __launch_bounds__(THREADS) void k_combine_norms(const double *__restrict__ U,
                                                           int64_t ld, int nu,
                                                           const double *__restrict__ Cm, int ldc,
                                                           int c0, int nc, int64_t n,
                                                           double *__restrict__ partial)
{
  constexpr int NC = 8;
  __shared__ double sC[NU * NC];
  for (int t = threadIdx.x; t < NU * NC; t += blockDim.x) {
    const int r = t % NU, c = t / NU;
    sC[t] = (r < nu && c < nc) ? Cm[r + (int64_t)(c0 + c) * ldc] : 0.0;
  }
  __syncthreads();
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    // keep the coefficients in LDS: without this barrier the compiler hoists
    // all NU * NC of them into registers
    __asm__ volatile("" ::: "memory");
    double u[NU];
#pragma unroll
    for (int r = 0; r < NU; ++r) u[r] = r < nu ? U[r * ld + i] : 0.0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double x = 0;
#pragma unroll
      for (int r = 0; r < NU; ++r) x += sC[r + c * NU] * u[r];
      acc[c] += x * x;
    }
  }
  block_reduce_write<NC>(acc, NC, partial);
}

unsigned nblocks(int64_t n)
{
  int64_t b = (n + THREADS - 1) / THREADS;
  if (b > NBLK) b = NBLK;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// Basis-size classes of the fused kernels: the register arrays are sized by
// the template argument, so keep it close to the actual count.  T is the
// vector element type.
// Launch KERNEL<J, T> for the smallest compiled J >= jj on nb workgroups,
// nb first capped at what that kernel keeps resident on the device at once
// (occupancy_cap): a grid-stride sweep with more workgroups than fit runs a
// second, partial round with most CUs idle.  nb (an unsigned lvalue) holds
// the grid launched, which the partial-sum reductions read.
#define TM_JDISPATCH_ONE(KERNEL, JJ, T, nb, ...)                                                \
  do {                                                                                          \
    nb = occupancy_cap(reinterpret_cast<const void *>(&KERNEL<JJ, T>), THREADS, nb);           \
    hipLaunchKernelGGL((KERNEL<JJ, T>), dim3(nb), dim3(THREADS), 0, s, __VA_ARGS__);           \
  } while (0)
#define TM_JDISPATCH(KERNEL, T, jj, nb, ...)                                                   \
  do {                                                                                         \
    if ((jj) <= 4) TM_JDISPATCH_ONE(KERNEL, 4, T, nb, __VA_ARGS__);                            \
    else if ((jj) <= 8) TM_JDISPATCH_ONE(KERNEL, 8, T, nb, __VA_ARGS__);                       \
    else if ((jj) <= 12) TM_JDISPATCH_ONE(KERNEL, 12, T, nb, __VA_ARGS__);                     \
    else if ((jj) <= 16) TM_JDISPATCH_ONE(KERNEL, 16, T, nb, __VA_ARGS__);                     \
    else if ((jj) <= 24) TM_JDISPATCH_ONE(KERNEL, 24, T, nb, __VA_ARGS__);                     \
    else if ((jj) <= 32) TM_JDISPATCH_ONE(KERNEL, 32, T, nb, __VA_ARGS__);                     \
    else TM_JDISPATCH_ONE(KERNEL, 64, T, nb, __VA_ARGS__);                                     \
    TM_LAUNCH_CHECK();                                                                         \
  } while (0)

namespace {
// Workgroups of `threads` threads that a kernel keeps resident on the whole
// device at once (occupancy query, cached per kernel); nb capped at that.
// 0 from the query (unknown) leaves nb as it is.
unsigned occupancy_cap(const void *kernel, int threads, unsigned nb)
{
  static std::mutex mu;
  static std::unordered_map<const void *, unsigned> cache;
  unsigned cap;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) {
      cap = it->second;
    } else {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) !=
              hipSuccess ||
          per_cu <= 0) {
        (void)hipGetLastError();
        per_cu = 0;
      }
      cap = (unsigned)per_cu * (unsigned)(g_num_cu > 0 ? g_num_cu : 256);
      cache.emplace(kernel, cap);
    }
  }
  return cap ? std::min(nb, cap) : nb;
}

// the packed kernels apply: 16-byte aligned vectors, ld a multiple of the
// pack width
template <typename T>
bool packs_ok(const void *a, const void *b, int64_t ld)
{
  return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0 &&
         (ld % Pk<T>::P) == 0;
}
// grid of a packed (or scalar) sweep over n elements
template <typename T>
unsigned sweep_blocks(bool packed, int64_t n)
{
  return packed ? nblocks(n / Pk<T>::P) : nblocks(n);
}
}  // namespace

template <typename T>
unsigned multidot(const T *V, int64_t ldv, int j, const T *w, int64_t n, double *partial,
                  double *out, hipStream_t s, bool reduce)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool pk = packs_ok<T>(V, w, ldv);
  unsigned nb = sweep_blocks<T>(pk, n);
  if (pk) TM_JDISPATCH(k_multidot_pk, T, j, nb, V, ldv, j, w, n, partial);
  else TM_JDISPATCH(k_multidot, T, j, nb, V, ldv, j, w, n, partial);
  if (reduce) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(64), 0, s, partial, (int)nb, j + 1,
                       out);
    TM_LAUNCH_CHECK();
  }
  return nb;
}

template <typename T>
unsigned update_dots(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                     double *partial, double *out, hipStream_t s, bool reduce)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool pk = packs_ok<T>(V, w, ldv);
  unsigned nb = sweep_blocks<T>(pk, n);
  if (pk)
    TM_JDISPATCH(k_update_dots_pk, T, j, nb, V, ldv, j, w, n, h, partial, 1, nullptr,
                 nullptr);
  else
    TM_JDISPATCH(k_update_dots, T, j, nb, V, ldv, j, w, n, h, partial, 1);
  if (reduce) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(64), 0, s, partial, (int)nb, j + 1,
                       out);
    TM_LAUNCH_CHECK();
  }
  return nb;
}

template <typename T>
unsigned update_dots_lanczos(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                             double *partial, double *out, hipStream_t s, bool reduce)
{
  TM_CHECK(j >= 1 && j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool pk = packs_ok<T>(V, w, ldv);
  unsigned nb = sweep_blocks<T>(pk, n);
  if (pk)
    TM_JDISPATCH(k_lanczos_dots_pk, T, j, nb, V, ldv, j, w, n, h, partial);
  else
    TM_JDISPATCH(k_update_dots, T, j, nb, V, ldv, j, w, n, h, partial, 1);
  if (reduce) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(64), 0, s, partial, (int)nb, j + 1,
                       out);
    TM_LAUNCH_CHECK();
  }
  return nb;
}

template <typename T>
void update_norm(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h,
                 double *partial, double *out, hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool pk = packs_ok<T>(V, w, ldv);
  unsigned nb = sweep_blocks<T>(pk, n);
  if (pk)
    TM_JDISPATCH(k_update_dots_pk, T, j, nb, V, ldv, j, w, n, h, partial, 0, nullptr,
                 nullptr);
  else
    TM_JDISPATCH(k_update_dots, T, j, nb, V, ldv, j, w, n, h, partial, 0);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, s, partial, (int)nb, 1, out);
  TM_LAUNCH_CHECK();
}

void reduce_step_close(const double *partial, unsigned nb, int cnt, double *out, int j,
                       const double *hb1, double *Hd, int ldh, double *nrm, double *flags,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_epi<2>, dim3(1), dim3(1024), 0, s, partial, (int)nb, cnt, out, j,
                     const_cast<double *>(hb1), Hd, ldh, nrm, flags);
  TM_LAUNCH_CHECK();
}

template <typename T>
void update_normalize(const T *V, int64_t ldv, int j, T *w, int64_t n, const double *h2,
                      const double *w0, const double *w1, const double *nrm2, int force,
                      hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool pk = packs_ok<T>(V, w, ldv);
  unsigned nb = sweep_blocks<T>(pk, n);
  if (pk)
    TM_JDISPATCH(k_update_normalize_pk, T, j, nb, V, ldv, j, w, n, h2, w0, w1, nrm2,
                 force);
  else
    TM_JDISPATCH(k_update_normalize, T, j, nb, V, ldv, j, w, n, h2, w0, w1, nrm2, force);
}

void reduce_lanczos_coeffs(const double *partial, unsigned nb, int cnt, double *out, int j,
                           double *Hd, int ldh, double *hL, hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_epi<1>, dim3(1), dim3(1024), 0, s, partial, (int)nb, cnt, out, j, hL,
                     Hd, ldh, nullptr, nullptr);
  TM_LAUNCH_CHECK();
}

void reduce_step_close_lanczos(const double *partial, unsigned nb, int cnt, double *out, int j,
                               const double *hL, double *Hd, int ldh, double *nrm, double *flags,
                               hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce_epi<3>, dim3(1), dim3(1024), 0, s, partial, (int)nb, cnt, out, j,
                     const_cast<double *>(hL), Hd, ldh, nrm, flags);
  TM_LAUNCH_CHECK();
}

void lag_close(const double *partial, unsigned nb, int j, const double *g, const double *Hd,
               int ldh, const double *nrm, double *hL, double *cf, hipStream_t s)
{
  TM_CHECK(j >= 1 && j <= LAG_MAXJ, TM_EINVAL, "lagged Lanczos step out of range");
  hipLaunchKernelGGL(k_lag_close, dim3(1), dim3(64), 0, s, partial, (int)nb, 2, j, g, Hd, ldh,
                     nrm, hL, cf);
  TM_LAUNCH_CHECK();
}

void arnoldi_lag_close(const double *partial, unsigned nb, int j, const double *g,
                       const double *Hd, int ldh, const double *nrm, double *hL, double *cf,
                       hipStream_t s)
{
  TM_CHECK(j >= 1 && j <= LAG_MAXJ, TM_EINVAL, "lagged Arnoldi step out of range");
  hipLaunchKernelGGL(k_arnoldi_lag_close, dim3(1), dim3(1024), 0, s, partial, (int)nb, j, g, Hd,
                     ldh, nrm, hL, cf);
  TM_LAUNCH_CHECK();
}

template <typename T>
unsigned lanczos_lag(T *V, int64_t ldv, int j, int64_t n, const double *cf, double *partial,
                     hipStream_t s)
{
  TM_CHECK(j >= 1 && j <= LAG_MAXJ, TM_EINVAL, "lagged Lanczos step out of range");
  TM_CHECK(packs_ok<T>(V, V, ldv), TM_EINVAL,
           "lagged Lanczos step: basis columns must be 16-byte aligned");
  unsigned nb = sweep_blocks<T>(true, n);
  if (j <= 4) TM_JDISPATCH_ONE(k_lanczos_lag_pk, 4, T, nb, V, ldv, j, n, cf, partial);
  else if (j <= 8) TM_JDISPATCH_ONE(k_lanczos_lag_pk, 8, T, nb, V, ldv, j, n, cf, partial);
  else if (j <= 12) TM_JDISPATCH_ONE(k_lanczos_lag_pk, 12, T, nb, V, ldv, j, n, cf, partial);
  else if (j <= 16) TM_JDISPATCH_ONE(k_lanczos_lag_pk, 16, T, nb, V, ldv, j, n, cf, partial);
  else TM_JDISPATCH_ONE(k_lanczos_lag_pk, 24, T, nb, V, ldv, j, n, cf, partial);
  TM_LAUNCH_CHECK();
  return nb;
}

template <typename T>
void normalize(T *w, int64_t n, const double *nrm2, hipStream_t s)
{
  if ((reinterpret_cast<uintptr_t>(w) & 15) == 0)
    hipLaunchKernelGGL(k_normalize_pk<T>, dim3(nblocks(n / Pk<T>::P) * 4), dim3(THREADS), 0, s,
                       w, n, nrm2);
  else
    hipLaunchKernelGGL(k_normalize<T>, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, w, n, nrm2);
  TM_LAUNCH_CHECK();
}

template <typename Ti, typename To>
void gemm_vy(const Ti *Vin, int64_t ldv, int m, const double *Y, int ldy, int k, To *Vout,
             int64_t ldo, int64_t n, hipStream_t s)
{
  TM_CHECK(m <= MAXJ && k <= m, TM_EINVAL, "restart block too large");
  TM_CHECK(sizeof(Ti) == sizeof(To) || (const void *)Vin != (const void *)Vout, TM_EINVAL,
           "in-place basis rotation needs one element type");
  // (more workgroups than fit at once: measured faster for this kernel than
  // a grid capped at the resident count)
  // (the packed form holds M x P converted elements: up to 24 columns)
  const bool pk = m <= 24 && packs_ok<Ti>(Vin, Vout, ldv) && ldo % Pk<Ti>::P == 0;
  const unsigned nb = (pk ? nblocks(n / Pk<Ti>::P) : nblocks(n)) * 4;
  auto go = [&](auto m_tag) {
    constexpr int MM = decltype(m_tag)::value;
    if (MM <= 24 && pk)
      hipLaunchKernelGGL((k_gemm_vy_pk<(MM <= 24 ? MM : 24), Ti, To>), dim3(nb), dim3(THREADS), 0,
                         s, Vin, ldv, m, Y, ldy, k, Vout, ldo, n);
    else
      hipLaunchKernelGGL((k_gemm_vy<MM, Ti, To>), dim3(nb), dim3(THREADS), 0, s, Vin, ldv, m, Y,
                         ldy, k, Vout, ldo, n);
  };
  if (m <= 8) go(std::integral_constant<int, 8>());
  else if (m <= 16) go(std::integral_constant<int, 16>());
  else if (m <= 24) go(std::integral_constant<int, 24>());
  else if (m <= 32) go(std::integral_constant<int, 32>());
  else go(std::integral_constant<int, 64>());
  TM_LAUNCH_CHECK();
}

template <typename T>
void random_uniform(T *v, int64_t n, uint64_t seed, hipStream_t s)
{
  hipLaunchKernelGGL(k_random<T>, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, v, n, seed);
  TM_LAUNCH_CHECK();
}

template <typename Ta, typename Tb>
void copy(const Ta *a, Tb *b, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((k_copy<Ta, Tb>), dim3(nblocks(n) * 4), dim3(THREADS), 0, s, a, b, n);
  TM_LAUNCH_CHECK();
}

void to_f32(const double *a, float *b, int64_t n, hipStream_t s) { copy(a, b, n, s); }

void fill(double *a, double val, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, a, val, n);
  TM_LAUNCH_CHECK();
}

void pow2_scale(const double *x, int64_t n, double *partial, double *scale, hipStream_t s)
{
  const unsigned nb = nblocks(n);
  hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(THREADS), 0, s, x, n, partial);
  hipLaunchKernelGGL(k_pow2_scale, dim3(1), dim3(64), 0, s, partial, (int)nb, scale);
  TM_LAUNCH_CHECK();
}

void gram(const double *A, const double *B, int64_t ld, int k, int64_t n, double *partial,
          double *G, hipStream_t s)
{
  TM_CHECK(k <= MAXJ, TM_EINVAL, "gram block too large");
  unsigned nb = nblocks(n);
  // row blocks of 4 (k <= 8) or 2 (k <= 16): each sweep reads RB + k vectors
  auto blocked = [&](auto rb_tag, auto j_tag) {
    constexpr int RB = decltype(rb_tag)::value, J = decltype(j_tag)::value;
    for (int r0 = 0; r0 < k; r0 += RB) {
      const int nr = std::min(RB, k - r0);
      hipLaunchKernelGGL((k_gram_block<RB, J>), dim3(nb), dim3(THREADS), 0, s, A, B, ld, k, n, r0,
                         nr, partial);
      hipLaunchKernelGGL(k_reduce_gram, dim3(nr * k), dim3(64), 0, s, partial, (int)nb, RB * J, J,
                         k, G + r0 * k);
      TM_LAUNCH_CHECK();
    }
  };
  if (k <= 4) {
    blocked(std::integral_constant<int, 4>(), std::integral_constant<int, 4>());
  } else if (k <= 8) {
    blocked(std::integral_constant<int, 4>(), std::integral_constant<int, 8>());
  } else if (k <= 16) {
    blocked(std::integral_constant<int, 2>(), std::integral_constant<int, 16>());
  } else {
    for (int row = 0; row < k; ++row) {
      TM_JDISPATCH(k_gram_row, double, k, nb, A, B, ld, k, n, row, partial);
      // G[row + c*k] = <A_row, B_c>
      hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(64), 0, s, partial, (int)nb, k,
                         G + row * k);
      TM_LAUNCH_CHECK();
    }
  }
}

void gram_sym(const double *U, int64_t ld, int nv, int64_t n, double *partial, double *G,
              hipStream_t s)
{
  TM_CHECK(nv >= 1 && nv <= MAXJ, TM_EINVAL, "gram block too large");
  if (nv > 10) {
    gram(U, U, ld, nv, n, partial, G, s);
    return;
  }
  // (the triangle's accumulators stay in registers: 6, 8 and 10 vectors)
  const int nt = nv <= 6 ? 21 : (nv <= 8 ? 36 : 55);
  const void *kf = nv <= 6   ? reinterpret_cast<const void *>(&k_gram_tri<6>)
                   : nv <= 8 ? reinterpret_cast<const void *>(&k_gram_tri<8>)
                             : reinterpret_cast<const void *>(&k_gram_tri<10>);
  const unsigned nb = occupancy_cap(kf, THREADS, nblocks(n));
  if (nv <= 6)
    hipLaunchKernelGGL(k_gram_tri<6>, dim3(nb), dim3(THREADS), 0, s, U, ld, nv, n, partial);
  else if (nv <= 8)
    hipLaunchKernelGGL(k_gram_tri<8>, dim3(nb), dim3(THREADS), 0, s, U, ld, nv, n, partial);
  else
    hipLaunchKernelGGL(k_gram_tri<10>, dim3(nb), dim3(THREADS), 0, s, U, ld, nv, n, partial);
  hipLaunchKernelGGL(k_reduce_tri, dim3(nt), dim3(64), 0, s, partial, (int)nb, nt, nv, G);
  TM_LAUNCH_CHECK();
}

void combine_norms(const double *U, int64_t ld, int nu, const double *Cm, int ldc, int nc,
                   int64_t n, double *partial, double *out, hipStream_t s)
{
  TM_CHECK(nu >= 1 && nu <= MAXJ && nc >= 1, TM_EINVAL, "combination block too large");
  const void *kf = nu <= 8    ? reinterpret_cast<const void *>(&k_combine_norms<8>)
                   : nu <= 16 ? reinterpret_cast<const void *>(&k_combine_norms<16>)
                   : nu <= 32 ? reinterpret_cast<const void *>(&k_combine_norms<32>)
                              : reinterpret_cast<const void *>(&k_combine_norms<64>);
  const unsigned nb = occupancy_cap(kf, THREADS, nblocks(n));
  for (int c0 = 0; c0 < nc; c0 += 8) {
    const int m = std::min(8, nc - c0);
    if (nu <= 8)
      hipLaunchKernelGGL(k_combine_norms<8>, dim3(nb), dim3(THREADS), 0, s, U, ld, nu, Cm, ldc, c0,
                         m, n, partial);
    else if (nu <= 16)
      hipLaunchKernelGGL(k_combine_norms<16>, dim3(nb), dim3(THREADS), 0, s, U, ld, nu, Cm, ldc,
                         c0, m, n, partial);
    else if (nu <= 32)
      hipLaunchKernelGGL(k_combine_norms<32>, dim3(nb), dim3(THREADS), 0, s, U, ld, nu, Cm, ldc,
                         c0, m, n, partial);
    else
      hipLaunchKernelGGL(k_combine_norms<64>, dim3(nb), dim3(THREADS), 0, s, U, ld, nu, Cm, ldc,
                         c0, m, n, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(m), dim3(64), 0, s, partial, (int)nb, 8,
                       out + c0);
    TM_LAUNCH_CHECK();
  }
}

// ---- block Krylov kernels ----------------------------------------------
namespace {
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int GRAM_THREADS = 256;
constexpr int GRAM_S = 16;          // elements per lane group and sweep step
constexpr int GRAM_MAXBLK = 2048;   // partial-sum blocks of the MFMA Gram

// Partial Gram G = X^T Y (X: nx <= 16 NG columns, Y: b <= 16 columns, both
// of length n, leading dimension ld) on FP64 MFMA (v_mfma_f64_16x16x4f64).
// In one MFMA the 64 lanes of a wave supply a 16 x 4 block of X^T (lane l:
// column l mod 16 of the group, element slot l / 16) and the matching 4 x 16
// block of Y (lane l: column l mod 16, element slot l / 16).  The sum runs
// over every element, so which element fills which slot is free: lane
// (column c, slot q) loads the four consecutive elements 16 s + 4 q .. +3 of
// its columns with one vector load (the four lanes of a column cover 64
// contiguous bytes) and feeds them to four MFMAs, one element each.  The
// 16 x 16 accumulator of group g holds <X_{16g+i}, Y_j> in
// lane (j + 16 (i mod 4)), register i / 4 (the accumulator layout of this
// instruction on gfx950, measured with one-hot operands).  Each workgroup
// adds its waves' accumulators in LDS and writes nx x 16 partial sums.
// v[0..3] = p[e .. e+3] as doubles (zero where on is false or beyond n):
// one 16-byte (float) or two 16-byte (double) loads when all four are in
// range (p + e is then 16 / 32 byte aligned: e is a multiple of 4 and the
// columns start on 64-element boundaries)
__device__ __forceinline__ void gram_load4(const float *p, int64_t e, int64_t n, bool on,
                                           double (&v)[4])
{
  if (on && e + 3 < n) {
    const float4 a = *reinterpret_cast<const float4 *>(p + e);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (on && e + i < n) ? (double)p[e + i] : 0.0;
  }
}
__device__ __forceinline__ void gram_load4(const double *p, int64_t e, int64_t n, bool on,
                                           double (&v)[4])
{
  if (on && e + 3 < n) {
    const double2 a = reinterpret_cast<const double2 *>(p + e)[0];
    const double2 b = reinterpret_cast<const double2 *>(p + e)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (on && e + i < n) ? p[e + i] : 0.0;
  }
}

template <int NG, typename T>
__global__ __launch_bounds__(GRAM_THREADS) void k_gram_xy_mfma(const T *__restrict__ X, int64_t ld,
                                                                int nx, const T *__restrict__ Y,
                                                                int b, int64_t n,
                                                                double *__restrict__ partial)
{
  __shared__ double red[GRAM_THREADS / 64][NG * 256];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  constexpr int NW = GRAM_THREADS / 64;
  d4 acc[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) acc[g] = d4{0.0, 0.0, 0.0, 0.0};
  const int64_t chunk = 4 * GRAM_S;  // elements per wave and step
  const T *yc = Y + (int64_t)col * ld;
  const bool yon = col < b;
  for (int64_t base = ((int64_t)blockIdx.x * NW + wid) * chunk; base < n;
       base += (int64_t)gridDim.x * NW * chunk) {
#pragma unroll
    for (int s = 0; s < GRAM_S / 4; ++s) {
      const int64_t e = base + 16 * s + 4 * kq;
      double yv[4], xv[NG][4];
      gram_load4(yc, e, n, yon, yv);
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        const int r = g * 16 + col;
        gram_load4(X + (int64_t)r * ld, e, n, r < nx, xv[g]);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int g = 0; g < NG; ++g)
          acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[g][v], yv[v], acc[g], 0, 0, 0);
    }
  }
  // accumulator element (i, j) of group g: lane j + 16 (i mod 4), register
  // i / 4, so register v of this lane holds row i = kq + 4 v, column col
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wid][g * 256 + (kq + 4 * v) * 16 + col] = acc[g][v];
  __syncthreads();
  for (int t = threadIdx.x; t < nx * 16; t += blockDim.x) {
    const int r = t >> 4, j = t & 15;
    double sum = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sum += red[w][(r >> 4) * 256 + (r & 15) * 16 + j];
    partial[(int64_t)blockIdx.x * nx * 16 + t] = sum;
  }
}

// G[r + c * nx] = sum over blocks of partial slot (r, c), r < nx, c < b
// (one wavefront per output, fixed order)
__global__ __launch_bounds__(64) void k_reduce_gram_xy(const double *__restrict__ partial,
                                                       int nblk, int nx, int b,
                                                       double *__restrict__ G)
{
  const int r = blockIdx.x / b, c = blockIdx.x % b;
  const double s = lane_sum(partial + r * 16 + c, nblk, (int64_t)nx * 16, threadIdx.x);
  if (threadIdx.x == 0) G[r + c * nx] = s;
}

// W <- (W - X H) Rinv over b <= 8 columns: H (nx x b, leading dimension
// ldh) may be null (no subtraction), Rinv (b x b upper triangular,
// column-major) may be null (no scaling).  X (nx <= NX columns) and W are
// distinct columns of one basis.
template <int NX, typename T>
__global__ __launch_bounds__(THREADS) void k_block_update(const T *X, int64_t ld, int nx, T *W,
                                                          int b, const double *__restrict__ H,
                                                          int ldh,
                                                          const double *__restrict__ Rinv,
                                                          int64_t n)
{
  constexpr int BM = 8;
  __shared__ double sH[NX * BM];
  __shared__ double sR[BM * BM];
  for (int t = threadIdx.x; t < NX * BM; t += blockDim.x) {
    const int r = t % NX, c = t / NX;
    sH[t] = (H && r < nx && c < b) ? H[r + (int64_t)c * ldh] : 0.0;
  }
  for (int t = threadIdx.x; t < BM * BM; t += blockDim.x) {
    const int r = t % BM, c = t / BM;
    sR[t] = Rinv ? ((r < b && c < b && r <= c) ? Rinv[r + c * b] : 0.0) : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double w[BM];
#pragma unroll
    for (int c = 0; c < BM; ++c) w[c] = c < b ? (double)W[(int64_t)c * ld + i] : 0.0;
    if (H) {
      for (int r = 0; r < nx; ++r) {
        const double x = X[(int64_t)r * ld + i];
#pragma unroll
        for (int c = 0; c < BM; ++c) w[c] -= x * sH[r + c * NX];
      }
    }
#pragma unroll
    for (int c = 0; c < BM; ++c) {
      if (c >= b) break;
      double o = 0;
#pragma unroll
      for (int q = 0; q < BM; ++q)
        if (q <= c) o += w[q] * sR[q + c * BM];
      W[(int64_t)c * ld + i] = (T)o;
    }
  }
}

// Cholesky factor R (upper, b x b, column-major) of S = R^T R; false if a
// pivot is not positive (relative to the largest diagonal entry)
__device__ bool chol_upper(const double *S, int b, double *R)
{
  double dmax = 0;
  for (int i = 0; i < b; ++i) dmax = fmax(dmax, S[i + i * b]);
  for (int j = 0; j < b; ++j) {
    for (int i = 0; i < b; ++i) R[i + j * b] = 0.0;
  }
  for (int j = 0; j < b; ++j) {
    double d = S[j + j * b];
    for (int k = 0; k < j; ++k) d -= R[k + j * b] * R[k + j * b];
    if (!(d > 1e-24 * dmax) || !(dmax > 0)) return false;
    d = sqrt(d);
    R[j + j * b] = d;
    for (int i = j + 1; i < b; ++i) {
      double t = S[j + i * b];
      for (int k = 0; k < j; ++k) t -= R[k + j * b] * R[k + i * b];
      R[j + i * b] = t / d;
    }
  }
  return true;
}

// inverse of an upper triangular R (b x b)
__device__ void inv_upper(const double *R, int b, double *Ri)
{
  for (int j = 0; j < b; ++j) {
    for (int i = 0; i < b; ++i) Ri[i + j * b] = 0.0;
    Ri[j + j * b] = 1.0 / R[j + j * b];
    for (int i = j - 1; i >= 0; --i) {
      double t = 0;
      for (int k = i + 1; k <= j; ++k) t += R[i + k * b] * Ri[k + j * b];
      Ri[i + j * b] = -t / R[i + i * b];
    }
  }
}

// Close a block step (one thread).  mode 0, after the second projection:
// G1, G2 = [V | W]^T W of the two passes (nx = nc + b rows, b columns);
// column block col0 of the projected matrix gets h1 + h2 in rows [0, nc) and
// R in rows [nc, nc + b), where R^T R = W2^T W2 = G2[W] - h2^T h2
// (Pythagoras, V orthonormal), and Rinv = R^-1.  mode 1, CholQR's second
// pass: G1 = W^T W (b x b), R2 = chol; the stored R block becomes R2 R and
// Rinv = R2^-1.  A failed factorisation records the step in flags[1] (if
// not yet set) and leaves Rinv = identity.
__global__ void k_block_close(const double *__restrict__ G1, const double *__restrict__ G2, int nc,
                              int b, double *__restrict__ Hd, int ldh, int col0,
                              double *__restrict__ Rinv, double *__restrict__ flags, int step,
                              int mode)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  constexpr int BM = 8;
  double S[BM * BM], R[BM * BM], Ri[BM * BM];
  const int nx = nc + b;
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
  // symmetrise
  for (int c1 = 0; c1 < b; ++c1)
    for (int c2 = c1 + 1; c2 < b; ++c2) {
      const double a = 0.5 * (S[c1 + c2 * b] + S[c2 + c1 * b]);
      S[c1 + c2 * b] = S[c2 + c1 * b] = a;
    }
  const bool ok = chol_upper(S, b, R);
  if (!ok) {
    // W is left as it is (Rinv = identity) with the stored coupling block
    // such that W_before = W R_block (identity after a failed first
    // factorisation); the host repairs the block from there
    if (flags[1] < 0) flags[1] = step;
    for (int t = 0; t < b * b; ++t) Rinv[t] = (t % (b + 1)) == 0 ? 1.0 : 0.0;
    if (mode == 0)
      for (int c = 0; c < b; ++c)
        for (int r = 0; r < b; ++r) Hd[nc + r + (int64_t)(col0 + c) * ldh] = r == c ? 1.0 : 0.0;
    return;
  }
  inv_upper(R, b, Ri);
  for (int t = 0; t < b * b; ++t) Rinv[t] = Ri[t];
  if (mode == 0) {
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) Hd[nc + r + (int64_t)(col0 + c) * ldh] = R[r + c * b];
  } else {
    // the R block sits in rows [nc, nc + b) of column block col0 (nc is the
    // row offset of W in that case)
    double O[BM * BM];
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) O[r + c * b] = Hd[nc + r + (int64_t)(col0 + c) * ldh];
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) {
        double t = 0;
        for (int k = 0; k < b; ++k) t += R[r + k * b] * O[k + c * b];
        Hd[nc + r + (int64_t)(col0 + c) * ldh] = t;
      }
  }
}
}  // namespace

template <typename T>
void gram_xy(const T *X, int64_t ld, int nx, const T *Y, int b, int64_t n, double *partial,
             double *G, hipStream_t s)
{
  TM_CHECK(nx >= 1 && nx <= 64 && b >= 1 && b <= 16, TM_EINVAL, "block Gram too large");
  TM_CHECK(((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(Y)) %
            (4 * sizeof(T))) == 0 && ld % 4 == 0,
           TM_EINVAL, "block Gram needs columns aligned to 4 elements");
  int64_t nb = (n + GRAM_THREADS / 64 * 4 * GRAM_S - 1) / (GRAM_THREADS / 64 * 4 * GRAM_S);
  if (nb > GRAM_MAXBLK) nb = GRAM_MAXBLK;
  if (nb < 1) nb = 1;
  const int ng = (nx + 15) / 16;
  // at most the workgroups that stay resident (one round)
  auto go = [&](auto ng_tag) {
    constexpr int NG = decltype(ng_tag)::value;
    nb = occupancy_cap(reinterpret_cast<const void *>(&k_gram_xy_mfma<NG, T>), GRAM_THREADS,
                       (unsigned)nb);
    hipLaunchKernelGGL((k_gram_xy_mfma<NG, T>), dim3((unsigned)nb), dim3(GRAM_THREADS), 0, s, X,
                       ld, nx, Y, b, n, partial);
  };
  if (ng == 1) go(std::integral_constant<int, 1>());
  else if (ng == 2) go(std::integral_constant<int, 2>());
  else if (ng == 3) go(std::integral_constant<int, 3>());
  else go(std::integral_constant<int, 4>());
  TM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_gram_xy, dim3(nx * b), dim3(64), 0, s, partial, (int)nb, nx, b, G);
  TM_LAUNCH_CHECK();
}

size_t gram_xy_partial_doubles() { return (size_t)GRAM_MAXBLK * 64 * 16; }

template <typename T>
void block_update(const T *X, int64_t ld, int nx, T *W, int b, const double *H, int ldh,
                  const double *Rinv, int64_t n, hipStream_t s)
{
  TM_CHECK(nx >= 0 && nx <= 64 && b >= 1 && b <= 8, TM_EINVAL, "block update too large");
  const unsigned nb = nblocks(n);
  const double *h = nx > 0 ? H : nullptr;
  if (nx <= 16)
    hipLaunchKernelGGL((k_block_update<16, T>), dim3(nb), dim3(THREADS), 0, s, X, ld, nx, W, b, h, ldh, Rinv, n);
  else if (nx <= 32)
    hipLaunchKernelGGL((k_block_update<32, T>), dim3(nb), dim3(THREADS), 0, s, X, ld, nx, W, b, h, ldh, Rinv, n);
  else
    hipLaunchKernelGGL((k_block_update<64, T>), dim3(nb), dim3(THREADS), 0, s, X, ld, nx, W, b, h, ldh, Rinv, n);
  TM_LAUNCH_CHECK();
}

void block_close(const double *G1, const double *G2, int nc, int b, double *Hd, int ldh, int col0,
                 double *Rinv, double *flags, int step, int mode, hipStream_t s)
{
  TM_CHECK(b >= 1 && b <= 8, TM_EINVAL, "block size above 8");
  hipLaunchKernelGGL(k_block_close, dim3(1), dim3(64), 0, s, G1, G2, nc, b, Hd, ldh, col0, Rinv,
                     flags, step, mode);
  TM_LAUNCH_CHECK();
}

#define TM_KRY_INST(T)                                                                          \
  template unsigned multidot<T>(const T *, int64_t, int, const T *, int64_t, double *,         \
                                double *, hipStream_t, bool);                                  \
  template unsigned update_dots<T>(const T *, int64_t, int, T *, int64_t, const double *,       \
                                   double *, double *, hipStream_t, bool);                     \
  template unsigned update_dots_lanczos<T>(const T *, int64_t, int, T *, int64_t,               \
                                           const double *, double *, double *, hipStream_t,    \
                                           bool);                                              \
  template void update_norm<T>(const T *, int64_t, int, T *, int64_t, const double *, double *, \
                               double *, hipStream_t);                                         \
  template void update_normalize<T>(const T *, int64_t, int, T *, int64_t, const double *,      \
                                    const double *, const double *, const double *, int,       \
                                    hipStream_t);                                              \
  template void normalize<T>(T *, int64_t, const double *, hipStream_t);                        \
  template unsigned lanczos_lag<T>(T *, int64_t, int, int64_t, const double *, double *,         \
                                   hipStream_t);                                                \
  template void random_uniform<T>(T *, int64_t, uint64_t, hipStream_t);
TM_KRY_INST(double)
TM_KRY_INST(float)
#undef TM_KRY_INST
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
template void gram_xy<double>(const double *, int64_t, int, const double *, int, int64_t, double *,
                              double *, hipStream_t);
template void gram_xy<float>(const float *, int64_t, int, const float *, int, int64_t, double *,
                             double *, hipStream_t);
template void block_update<double>(const double *, int64_t, int, double *, int, const double *,
                                   int, const double *, int64_t, hipStream_t);
template void block_update<float>(const float *, int64_t, int, float *, int, const double *, int,
                                  const double *, int64_t, hipStream_t);

}  // namespace kry
}  // namespace ltm
