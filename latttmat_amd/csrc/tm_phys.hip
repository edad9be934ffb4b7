// tm_phys.hip -- observables of an eigenvector pair (the X_phys.dat figures)
// as one device reduction over the state or orbit index.
//
// Per index i with state s (i itself, or the orbit representative reps[i]):
//   weight  w = q_i^2 (symmetric operators, TMat/tmatphys.cpp:13-57) or
//           ql_i qr_i (the others, TMat/tmatphys.cpp:71-98, 106-132)
//   m       = 2 popcount(s') - L, s' the state (or the low layer of a
//             two-layer state, or the representative's high layer,
//             TMat/tmatreduce.cpp:882-909)
//   sums    sum w, sum w m^k (k = 1..4), and the extra observables:
//             node count n, n^2 (TMat/tmatphys.cpp:39-57,
//             TMat/tmatreduce.cpp:593-613);
//             |m_st|, |m|, m_st^2, m^2 with m_st the magnetisation after
//             flipping every odd site (TMat/tmatreduce.cpp:1289-1320).
// One grid-stride pass per block writes 9 partial sums; one block adds the
// partials in a fixed order, so the result does not depend on timing.
#include "tm_device.hpp"

namespace ltm {
namespace {

constexpr int PH_THREADS = 256;
constexpr int PH_MAXBLK = 1024;
constexpr int PH_N = 9;  // sum w, sum w m^k (4), extras (4)

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Node count of TMat/utility.cpp:33-47 (open boundary: the mask nspin-1 as
// written there).
__device__ __forceinline__ int nnode(uint64_t s, int L, int bc)
{
  uint64_t sh = s >> 1;
  if (!bc) sh |= (s & 1ull) << (L - 1);
  else sh |= (s & (uint64_t)(L - 1)) << 1;
  return __popcll(s ^ sh);
}

template <int PK>
__global__ __launch_bounds__(PH_THREADS) void k_phys_sums(const double *__restrict__ ql,
                                                          const double *__restrict__ qr,
                                                          const uint64_t *__restrict__ reps,
                                                          int64_t n, int L, int bc,
                                                          double *__restrict__ partial)
{
  double acc[PH_N];
#pragma unroll
  for (int k = 0; k < PH_N; ++k) acc[k] = 0.0;
  const uint64_t lmask = L >= 64 ? ~0ull : (1ull << L) - 1;
  for (int64_t i = blockIdx.x * (int64_t)PH_THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * PH_THREADS) {
    const uint64_t s = reps ? reps[i] : (uint64_t)i;
    const double a = ql[i];
    const double w = (PK == PHYS_SYM2D || PK == PHYS_SYM3D || PK == PHYS_DNNI) ? a * a : a * qr[i];
    const uint64_t sl = PK == PHYS_GEN_LOW ? (s & lmask) : PK == PHYS_GEN_HIGH ? (s >> L) : s;
    const double m = 2.0 * __popcll(sl) - L;
    acc[0] += w;
    double mm = m;
#pragma unroll
    for (int k = 1; k <= 4; ++k) {
      acc[k] += mm * w;
      mm *= m;
    }
    if (PK == PHYS_SYM2D) {
      const double nd = nnode(s, L, bc);
      acc[5] += nd * w;
      acc[6] += nd * nd * w;
    }
    if (PK == PHYS_DNNI) {
      const double ms = 2.0 * __popcll((s ^ 0xaaaaaaaaaaaaaaaaull) & lmask) - L;
      acc[5] += fabs(ms) * w;
      acc[6] += fabs(m) * w;
      acc[7] += ms * ms * w;
      acc[8] += m * m * w;
    }
  }
  __shared__ double red[PH_N][PH_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < PH_N; ++k) {
    const double t = wave_sum(acc[k]);
    if (lane == 0) red[k][wv] = t;
  }
  __syncthreads();
  if (threadIdx.x < PH_N) {
    double t = 0;
    for (int w = 0; w < PH_THREADS / 64; ++w) t += red[threadIdx.x][w];
    partial[(int64_t)blockIdx.x * PH_N + threadIdx.x] = t;
  }
}

// out[k] = sum over blocks of partial[b * PH_N + k]: each of 64 lanes adds a
// strided subset in order, then a fixed shuffle tree.
__global__ __launch_bounds__(64) void k_phys_final(const double *__restrict__ partial, int nblk,
                                                   double *__restrict__ out)
{
  const int k = blockIdx.x;
  double t = 0;
  for (int b = threadIdx.x; b < nblk; b += 64) t += partial[(int64_t)b * PH_N + k];
  t = wave_sum(t);
  if (threadIdx.x == 0) out[k] = t;
}

template <int PK>
void launch_sums(const double *ql, const double *qr, const uint64_t *reps, int64_t n, int L,
                 int bc, double *partial, int nblk, hipStream_t s)
{
  hipLaunchKernelGGL(k_phys_sums<PK>, dim3(nblk), dim3(PH_THREADS), 0, s, ql, qr, reps, n, L, bc,
                     partial);
  TM_LAUNCH_CHECK();
}

}  // namespace

size_t phys_work_bytes() { return sizeof(double) * (PH_N * (size_t)PH_MAXBLK + PH_N); }

void phys_sums(int pk, const double *ql, const double *qr, const uint64_t *reps, int64_t n, int L,
               int bc, void *work, double sums[9], hipStream_t s)
{
  const int64_t want = (n + PH_THREADS - 1) / PH_THREADS;
  const int nblk = (int)(want < PH_MAXBLK ? (want > 0 ? want : 1) : PH_MAXBLK);
  double *partial = static_cast<double *>(work);
  double *out = partial + (size_t)PH_N * PH_MAXBLK;
  switch (pk) {
    case PHYS_SYM2D: launch_sums<PHYS_SYM2D>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    case PHYS_SYM3D: launch_sums<PHYS_SYM3D>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    case PHYS_DNNI: launch_sums<PHYS_DNNI>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    case PHYS_GEN: launch_sums<PHYS_GEN>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    case PHYS_GEN_LOW: launch_sums<PHYS_GEN_LOW>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    case PHYS_GEN_HIGH: launch_sums<PHYS_GEN_HIGH>(ql, qr, reps, n, L, bc, partial, nblk, s); break;
    default: throw Error(TM_EINVAL, "unknown observable kind");
  }
  hipLaunchKernelGGL(k_phys_final, dim3(PH_N), dim3(64), 0, s, partial, nblk, out);
  TM_LAUNCH_CHECK();
  copy_sync(sums, out, sizeof(double) * PH_N, hipMemcpyDeviceToHost, s);
}

}  // namespace ltm
