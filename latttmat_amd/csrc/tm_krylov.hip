// tm_krylov.hip -- FP64 Krylov-basis kernels for the device eigensolver.
//
// The basis V is column-major (vector k at V + k*ldv).  Every reduction is
// two-level and deterministic: a fixed grid of NBLK workgroups writes
// partial sums, a single workgroup adds them in a fixed order.  Dot products
// use wavefront (64-lane) shuffles, then LDS across the 4 wavefronts.
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2): the
// update kernel subtracts V h and, in the same sweep over the data, already
// accumulates the dot products needed by the next correction.
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

// Block-reduce the first `cnt` of the J values acc[] (compile-time indexed,
// so the array stays in registers) and write them to out[0..cnt).
template <int J>
__device__ __forceinline__ void block_reduce_write(const double (&acc)[J], int cnt, double *out)
{
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
    out[k] = s;
  }
}

// out slots: <V_k, w> for k < j, then <w, w> in slot j
template <int J>
__global__ __launch_bounds__(THREADS) void k_multidot(const double *__restrict__ V, int64_t ldv,
                                                      int j, const double *__restrict__ w,
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
      if (k < j) acc[k] += V[k * ldv + i] * wi;
    acc[J] += wi * wi;
  }
  // move <w,w> into slot j
#pragma unroll
  for (int k = 0; k < J; ++k)
    if (k == j) acc[k] = acc[J];
  block_reduce_write<J + 1>(acc, j + 1, partial + (int64_t)blockIdx.x * (j + 1));
}

// w -= sum_k h[k] V_k; then (dots=1) slots <V_k, w>, <w,w> or (dots=0) slot 0 = <w,w>
template <int J>
__global__ __launch_bounds__(THREADS) void k_update_dots(const double *__restrict__ V,
                                                         int64_t ldv, int j,
                                                         double *__restrict__ w, int64_t n,
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
    w[i] = wi;
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
    block_reduce_write<J + 1>(acc, j + 1, partial + (int64_t)blockIdx.x * (j + 1));
  } else {
    acc[0] = acc[J];
    block_reduce_write<J + 1>(acc, 1, partial + (int64_t)blockIdx.x);
  }
}

// One wavefront per output slot: lane l adds partials l, l+64, ... in a fixed
// order, then a fixed shuffle tree -> deterministic.
__global__ __launch_bounds__(64) void k_reduce_partials(const double *__restrict__ partial,
                                                        int nblk, int cnt,
                                                        double *__restrict__ out)
{
  const int k = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += 64) s += partial[(int64_t)b * cnt + k];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[k] = s;
}

// ---- 16-byte (two-double) variants: V columns and w 16-byte aligned, ldv
// even.  Lane p handles elements 2p and 2p+1; block 0 thread 0 adds an odd
// tail element.  Same two-level deterministic reductions.
template <int J>
__global__ __launch_bounds__(THREADS) void k_multidot2(const double *__restrict__ V, int64_t ldv,
                                                       int j, const double *__restrict__ w,
                                                       int64_t n, double *__restrict__ partial)
{
  const double2 *__restrict__ V2 = reinterpret_cast<const double2 *>(V);
  const double2 *__restrict__ w2 = reinterpret_cast<const double2 *>(w);
  const int64_t ld2 = ldv >> 1, n2 = n >> 1;
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < n2;
       p += (int64_t)gridDim.x * THREADS) {
    const double2 wi = w2[p];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        const double2 v = V2[k * ld2 + p];
        acc[k] += v.x * wi.x + v.y * wi.y;
      }
    acc[J] += wi.x * wi.x + wi.y * wi.y;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const double wi = w[n - 1];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) acc[k] += V[k * ldv + n - 1] * wi;
    acc[J] += wi * wi;
  }
#pragma unroll
  for (int k = 0; k < J; ++k)
    if (k == j) acc[k] = acc[J];
  block_reduce_write<J + 1>(acc, j + 1, partial + (int64_t)blockIdx.x * (j + 1));
}

// w -= sum_k h[k] V_k; then (dots=1) slots <V_k, w>, <w,w> or (dots=0) slot 0 = <w,w>
template <int J>
__global__ __launch_bounds__(THREADS) void k_update_dots2(const double *__restrict__ V,
                                                          int64_t ldv, int j,
                                                          double *__restrict__ w, int64_t n,
                                                          const double *__restrict__ h,
                                                          double *__restrict__ partial, int dots,
                                                          const double *w0p, const double *w1p)
{
  // conditional form (w0p != null): only when the first projection removed
  // more than half of the norm^2; the whole grid takes the same branch
  if (w0p && (w1p[0] > 0.5 * w0p[0])) return;
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = k < j ? h[k] : 0.0;
  __syncthreads();
  const double2 *__restrict__ V2 = reinterpret_cast<const double2 *>(V);
  double2 *__restrict__ w2 = reinterpret_cast<double2 *>(w);
  const int64_t ld2 = ldv >> 1, n2 = n >> 1;
  double acc[J + 1];
#pragma unroll
  for (int k = 0; k <= J; ++k) acc[k] = 0;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < n2;
       p += (int64_t)gridDim.x * THREADS) {
    double2 vk[J];
    double2 wi = w2[p];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        vk[k] = V2[k * ld2 + p];
        wi.x -= hk[k] * vk[k].x;
        wi.y -= hk[k] * vk[k].y;
      }
    w2[p] = wi;
    if (dots) {
#pragma unroll
      for (int k = 0; k < J; ++k)
        if (k < j) acc[k] += vk[k].x * wi.x + vk[k].y * wi.y;
    }
    acc[J] += wi.x * wi.x + wi.y * wi.y;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double vk[J];
    double wi = w[n - 1];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) {
        vk[k] = V[k * ldv + n - 1];
        wi -= hk[k] * vk[k];
      }
    w[n - 1] = wi;
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
    block_reduce_write<J + 1>(acc, j + 1, partial + (int64_t)blockIdx.x * (j + 1));
  } else {
    acc[0] = acc[J];
    block_reduce_write<J + 1>(acc, 1, partial + (int64_t)blockIdx.x);
  }
}

__global__ void k_normalize2(double *__restrict__ w, int64_t n, const double *__restrict__ nrm2)
{
  const double inv = 1.0 / sqrt(nrm2[0]);
  double2 *__restrict__ w2 = reinterpret_cast<double2 *>(w);
  const int64_t n2 = n >> 1;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n2;
       p += (int64_t)gridDim.x * blockDim.x) {
    double2 v = w2[p];
    v.x *= inv;
    v.y *= inv;
    w2[p] = v;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) w[n - 1] *= inv;
}

// Close a Krylov step in one sweep: w <- (w - c V h2) / beta, with c = 1
// when the DGKS correction is due (w1 <= w0 / 2) and beta^2 = nrm2[0] as set
// by k_step_finish (w1 - |h2|^2 after a correction, w1 otherwise).
template <int J>
__global__ __launch_bounds__(THREADS) void k_update_normalize(const double *__restrict__ V,
                                                              int64_t ldv, int j,
                                                              double *__restrict__ w, int64_t n,
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
      if (k < jj) wi -= hk[k] * V[k * ldv + i];
    w[i] = wi * inv;
  }
}

template <int J>
__global__ __launch_bounds__(THREADS) void k_update_normalize2(const double *__restrict__ V,
                                                               int64_t ldv, int j,
                                                               double *__restrict__ w, int64_t n,
                                                               const double *__restrict__ h2,
                                                               const double *w0p,
                                                               const double *w1p,
                                                               const double *__restrict__ nrm2,
                                                               int force)
{
  const bool re = force || !(w1p[0] > 0.5 * w0p[0]);
  const double inv = 1.0 / sqrt(nrm2[0]);
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = (re && k < j) ? h2[k] : 0.0;
  __syncthreads();
  const int jj = re ? j : 0;
  const double2 *__restrict__ V2 = reinterpret_cast<const double2 *>(V);
  double2 *__restrict__ w2 = reinterpret_cast<double2 *>(w);
  const int64_t ld2 = ldv >> 1, n2 = n >> 1;
  for (int64_t p = blockIdx.x * (int64_t)THREADS + threadIdx.x; p < n2;
       p += (int64_t)gridDim.x * THREADS) {
    double2 wi = w2[p];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < jj) {
        const double2 v = V2[k * ld2 + p];
        wi.x -= hk[k] * v.x;
        wi.y -= hk[k] * v.y;
      }
    wi.x *= inv;
    wi.y *= inv;
    w2[p] = wi;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double wi = w[n - 1];
    for (int k = 0; k < jj; ++k) wi -= hk[k] * V[k * ldv + n - 1];
    w[n - 1] = wi * inv;
  }
}

// One thread closes step j from the two projections alone: column j of the
// projected matrix = h1 + c h2 (c as above), beta^2 = w1 - |h2|^2 after a
// correction (Pythagoras; V orthonormal) or w1; nrm[0] = beta^2.  flags[0]
// counts corrections; flags[1] records the first step j whose beta is below
// 1e-13 |h| or whose correction removed more than half of w1 again (a
// numerically invariant subspace: the host redoes the cycle from j with
// explicit norms).
__global__ void k_step_close(int j, const double *__restrict__ hb1,
                             const double *__restrict__ hb2, double *__restrict__ Hd, int ldh,
                             double *__restrict__ nrm, double *__restrict__ flags)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
  const double beta = sqrt(fmax(beta2, 0.0));
  Hd[nk + (int64_t)j * ldh] = beta;
  nrm[0] = beta2;
  if (re) flags[0] += 1.0;
  const bool unsafe = re && !(h22 <= 0.5 * w1);
  if ((unsafe || beta <= 1e-13 * sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
}

// Symmetric operator, step j > first step of the cycle: T v_j couples to
// v_{j-1}, v_j, v_{j+1} only.  hL[0..j] = (0, .., 0, beta_{j-1}, alpha) with
// beta_{j-1} = Hd[j, j-1] and alpha = <v_j, T v_j> (hA[0]).
__global__ void k_lanczos_coeffs(int j, const double *__restrict__ hA,
                                 const double *__restrict__ Hd, int ldh, double *__restrict__ hL)
{
  for (int k = threadIdx.x; k <= j; k += blockDim.x)
    hL[k] = k == j ? hA[0] : (k == j - 1 ? Hd[j + (int64_t)(j - 1) * ldh] : 0.0);
}

// Close a Lanczos step: column j = hL + h2 (the full reorthogonalisation pass
// always runs), beta^2 = |w1|^2 - |h2|^2 (Pythagoras).  flags as k_step_close;
// unsafe when the reorthogonalisation removed more than half of |w1|^2.
__global__ void k_step_close_lanczos(int j, const double *__restrict__ hL,
                                     const double *__restrict__ hb2, double *__restrict__ Hd,
                                     int ldh, double *__restrict__ nrm,
                                     double *__restrict__ flags)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
  const double beta = sqrt(fmax(beta2, 0.0));
  Hd[nk + (int64_t)j * ldh] = beta;
  nrm[0] = beta2;
  flags[0] += 1.0;
  const bool unsafe = !(h22 <= 0.5 * w1);
  if ((unsafe || beta <= 1e-13 * sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
}

// DGKS correction only when the first projection removed more than half of
// the norm^2 (w1 <= w0 / 2, or not finite); every block reads the same two
// device values, so the whole grid takes the same branch.
__device__ __forceinline__ bool need_reorth(const double *w0p, const double *w1p)
{
  return !(w1p[0] > 0.5 * w0p[0]);
}

template <int J>
__global__ __launch_bounds__(THREADS) void k_update_norm_cond(const double *__restrict__ V,
                                                              int64_t ldv, int j,
                                                              double *__restrict__ w, int64_t n,
                                                              const double *__restrict__ h,
                                                              double *__restrict__ partial,
                                                              const double *w0p, const double *w1p)
{
  if (!need_reorth(w0p, w1p)) return;
  __shared__ double hk[J];
  for (int k = threadIdx.x; k < J; k += blockDim.x) hk[k] = k < j ? h[k] : 0.0;
  __syncthreads();
  double acc[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * THREADS) {
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < J; ++k)
      if (k < j) wi -= hk[k] * V[k * ldv + i];
    w[i] = wi;
    acc[0] += wi * wi;
  }
  block_reduce_write<1>(acc, 1, partial + (int64_t)blockIdx.x);
}

__global__ __launch_bounds__(64) void k_reduce_partials_cond(const double *__restrict__ partial,
                                                             int nblk, double *__restrict__ out,
                                                             const double *w0p, const double *w1p)
{
  if (!need_reorth(w0p, w1p)) return;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += 64) s += partial[b];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[0] = s;
}

// Close Krylov step j on the device (one thread): column j of the projected
// matrix (ld ldh) = first projection + correction, the subdiagonal
// beta = ||w||, nrm[0] = beta^2 for the normalisation, flags[0] counts
// corrections, flags[1] records the first step whose beta fell below
// 1e-13 ||h|| (an invariant subspace; the host redoes the cycle from there).
__global__ void k_step_finish(int j, const double *__restrict__ hb1,
                              const double *__restrict__ hb2, const double *__restrict__ hb3,
                              double *__restrict__ Hd, int ldh, double *__restrict__ nrm,
                              double *__restrict__ flags)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nk = j + 1;
  const bool re = need_reorth(hb1 + nk, hb2 + nk);
  double hn2 = 0;
  for (int k = 0; k < nk; ++k) {
    const double hk = hb1[k] + (re ? hb2[k] : 0.0);
    Hd[k + (int64_t)j * ldh] = hk;
    hn2 += hk * hk;
  }
  const double beta2 = re ? hb3[0] : hb2[nk];
  const double beta = sqrt(fmax(beta2, 0.0));
  Hd[nk + (int64_t)j * ldh] = beta;
  nrm[0] = beta2;
  if (re) flags[0] += 1.0;
  if ((beta <= 1e-13 * sqrt(hn2) || !(beta > 0)) && flags[1] < 0) flags[1] = j;
}

__global__ void k_normalize(double *__restrict__ w, int64_t n, const double *__restrict__ nrm2)
{
  const double inv = 1.0 / sqrt(nrm2[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] *= inv;
}

template <int M>
__global__ __launch_bounds__(THREADS) void k_gemm_vy(const double *__restrict__ Vin, int64_t ldv,
                                                     int m, const double *__restrict__ Y,
                                                     int ldy, int k, double *Vout, int64_t ldo,
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
    for (int r = 0; r < M; ++r) v[r] = r < m ? Vin[r * ldv + i] : 0.0;
    // all reads of row i happen before any write of row i (Vout may alias)
    for (int c = 0; c < k; ++c) {
      double s = 0;
#pragma unroll
      for (int r = 0; r < M; ++r) s += v[r] * sY[r + c * M];
      Vout[c * ldo + i] = s;
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
__global__ void k_random(double *v, int64_t n, uint64_t seed)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t r = splitmix64(splitmix64(seed) ^ (uint64_t)i);
    v[i] = ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
}

__global__ void k_copy(const double *a, double *b, int64_t n)
{
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
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

template <int J>
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
  block_reduce_write<J>(acc, k, partial + (int64_t)blockIdx.x * k);
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
// the template argument, so keep it close to the actual count.
#define TM_JDISPATCH(KERNEL, jj, grid, ...)                                                  \
  do {                                                                                       \
    if ((jj) <= 4) hipLaunchKernelGGL(KERNEL<4>, grid, dim3(THREADS), 0, s, __VA_ARGS__);     \
    else if ((jj) <= 8) hipLaunchKernelGGL(KERNEL<8>, grid, dim3(THREADS), 0, s, __VA_ARGS__); \
    else if ((jj) <= 12) hipLaunchKernelGGL(KERNEL<12>, grid, dim3(THREADS), 0, s, __VA_ARGS__); \
    else if ((jj) <= 16) hipLaunchKernelGGL(KERNEL<16>, grid, dim3(THREADS), 0, s, __VA_ARGS__); \
    else if ((jj) <= 24) hipLaunchKernelGGL(KERNEL<24>, grid, dim3(THREADS), 0, s, __VA_ARGS__); \
    else if ((jj) <= 32) hipLaunchKernelGGL(KERNEL<32>, grid, dim3(THREADS), 0, s, __VA_ARGS__); \
    else hipLaunchKernelGGL(KERNEL<64>, grid, dim3(THREADS), 0, s, __VA_ARGS__);              \
    TM_LAUNCH_CHECK();                                                                       \
  } while (0)

bool pairs_ok(const void *a, const void *b, int64_t ld)
{
  return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0 &&
         (ld & 1) == 0;
}

void multidot(const double *V, int64_t ldv, int j, const double *w, int64_t n, double *partial,
              double *out, hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool v2 = pairs_ok(V, w, ldv);
  const unsigned nb = v2 ? nblocks(n / 2) : nblocks(n);
  if (v2) TM_JDISPATCH(k_multidot2, j, dim3(nb), V, ldv, j, w, n, partial);
  else TM_JDISPATCH(k_multidot, j, dim3(nb), V, ldv, j, w, n, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(64), 0, s, partial, (int)nb, j + 1, out);
  TM_LAUNCH_CHECK();
}

void update_dots(const double *V, int64_t ldv, int j, double *w, int64_t n, const double *h,
                 double *partial, double *out, hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool v2 = pairs_ok(V, w, ldv);
  const unsigned nb = v2 ? nblocks(n / 2) : nblocks(n);
  if (v2)
    TM_JDISPATCH(k_update_dots2, j, dim3(nb), V, ldv, j, w, n, h, partial, 1, nullptr, nullptr);
  else
    TM_JDISPATCH(k_update_dots, j, dim3(nb), V, ldv, j, w, n, h, partial, 1);
  hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(64), 0, s, partial, (int)nb, j + 1, out);
  TM_LAUNCH_CHECK();
}

void update_norm(const double *V, int64_t ldv, int j, double *w, int64_t n, const double *h,
                 double *partial, double *out, hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool v2 = pairs_ok(V, w, ldv);
  const unsigned nb = v2 ? nblocks(n / 2) : nblocks(n);
  if (v2)
    TM_JDISPATCH(k_update_dots2, j, dim3(nb), V, ldv, j, w, n, h, partial, 0, nullptr, nullptr);
  else
    TM_JDISPATCH(k_update_dots, j, dim3(nb), V, ldv, j, w, n, h, partial, 0);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, s, partial, (int)nb, 1, out);
  TM_LAUNCH_CHECK();
}

void update_norm_cond(const double *V, int64_t ldv, int j, double *w, int64_t n, const double *h,
                      double *partial, double *out, const double *w0, const double *w1,
                      hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  const bool v2 = pairs_ok(V, w, ldv);
  const unsigned nb = v2 ? nblocks(n / 2) : nblocks(n);
  if (v2)
    TM_JDISPATCH(k_update_dots2, j, dim3(nb), V, ldv, j, w, n, h, partial, 0, w0, w1);
  else
    TM_JDISPATCH(k_update_norm_cond, j, dim3(nb), V, ldv, j, w, n, h, partial, w0, w1);
  hipLaunchKernelGGL(k_reduce_partials_cond, dim3(1), dim3(64), 0, s, partial, (int)nb, out, w0,
                     w1);
  TM_LAUNCH_CHECK();
}

void step_close(int j, const double *hb1, const double *hb2, double *Hd, int ldh, double *nrm,
                double *flags, hipStream_t s)
{
  hipLaunchKernelGGL(k_step_close, dim3(1), dim3(64), 0, s, j, hb1, hb2, Hd, ldh, nrm, flags);
  TM_LAUNCH_CHECK();
}

void update_normalize(const double *V, int64_t ldv, int j, double *w, int64_t n,
                      const double *h2, const double *w0, const double *w1, const double *nrm2,
                      int force, hipStream_t s)
{
  TM_CHECK(j <= MAXJ - 1, TM_EINVAL, "Krylov basis too large for the device kernels");
  if (pairs_ok(V, w, ldv))
    TM_JDISPATCH(k_update_normalize2, j, dim3(nblocks(n / 2)), V, ldv, j, w, n, h2, w0, w1, nrm2,
                 force);
  else
    TM_JDISPATCH(k_update_normalize, j, dim3(nblocks(n)), V, ldv, j, w, n, h2, w0, w1, nrm2,
                 force);
}

void lanczos_coeffs(int j, const double *hA, const double *Hd, int ldh, double *hL, hipStream_t s)
{
  hipLaunchKernelGGL(k_lanczos_coeffs, dim3(1), dim3(64), 0, s, j, hA, Hd, ldh, hL);
  TM_LAUNCH_CHECK();
}

void step_close_lanczos(int j, const double *hL, const double *hb2, double *Hd, int ldh,
                        double *nrm, double *flags, hipStream_t s)
{
  hipLaunchKernelGGL(k_step_close_lanczos, dim3(1), dim3(64), 0, s, j, hL, hb2, Hd, ldh, nrm,
                     flags);
  TM_LAUNCH_CHECK();
}

void step_finish(int j, const double *hb1, const double *hb2, const double *hb3, double *Hd,
                 int ldh, double *nrm, double *flags, hipStream_t s)
{
  hipLaunchKernelGGL(k_step_finish, dim3(1), dim3(64), 0, s, j, hb1, hb2, hb3, Hd, ldh, nrm,
                     flags);
  TM_LAUNCH_CHECK();
}

void normalize(double *w, int64_t n, const double *nrm2, hipStream_t s)
{
  if ((reinterpret_cast<uintptr_t>(w) & 15) == 0)
    hipLaunchKernelGGL(k_normalize2, dim3(nblocks(n / 2) * 4), dim3(THREADS), 0, s, w, n, nrm2);
  else
    hipLaunchKernelGGL(k_normalize, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, w, n, nrm2);
  TM_LAUNCH_CHECK();
}

void gemm_vy(const double *Vin, int64_t ldv, int m, const double *Y, int ldy, int k,
             double *Vout, int64_t ldo, int64_t n, hipStream_t s)
{
  TM_CHECK(m <= MAXJ && k <= m, TM_EINVAL, "restart block too large");
  const unsigned nb = nblocks(n) * 4;
  if (m <= 8)
    hipLaunchKernelGGL(k_gemm_vy<8>, dim3(nb), dim3(THREADS), 0, s, Vin, ldv, m, Y, ldy, k, Vout,
                       ldo, n);
  else if (m <= 16)
    hipLaunchKernelGGL(k_gemm_vy<16>, dim3(nb), dim3(THREADS), 0, s, Vin, ldv, m, Y, ldy, k,
                       Vout, ldo, n);
  else if (m <= 32)
    hipLaunchKernelGGL(k_gemm_vy<32>, dim3(nb), dim3(THREADS), 0, s, Vin, ldv, m, Y, ldy, k,
                       Vout, ldo, n);
  else
    hipLaunchKernelGGL(k_gemm_vy<64>, dim3(nb), dim3(THREADS), 0, s, Vin, ldv, m, Y, ldy, k,
                       Vout, ldo, n);
  TM_LAUNCH_CHECK();
}

void random_uniform(double *v, int64_t n, uint64_t seed, hipStream_t s)
{
  hipLaunchKernelGGL(k_random, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, v, n, seed);
  TM_LAUNCH_CHECK();
}

void copy(const double *a, double *b, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, a, b, n);
  TM_LAUNCH_CHECK();
}

void to_f32(const double *a, float *b, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(k_to_f32, dim3(nblocks(n) * 4), dim3(THREADS), 0, s, a, b, n);
  TM_LAUNCH_CHECK();
}

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
  const unsigned nb = nblocks(n);
  for (int row = 0; row < k; ++row) {
    TM_JDISPATCH(k_gram_row, k, dim3(nb), A, B, ld, k, n, row, partial);
    // G[row + c*k] = <A_row, B_c>
    hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(64), 0, s, partial, (int)nb, k,
                       G + row * k);
    TM_LAUNCH_CHECK();
  }
}

}  // namespace kry
}  // namespace ltm
