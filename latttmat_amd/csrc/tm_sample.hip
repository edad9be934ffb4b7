// tm_sample.hip -- planting equilibrium configurations from the leading
// eigenvector (the reference's dumpconfig program) as device Markov chains.
//
// A chain starts at an index drawn with weight vec_i^2 (SymDumpConfig::initq,
// DumpConfig/dumpconfig.cpp:24-35) and moves by draws from
//   XAT  (DumpConfig/dumpconfig.cpp:88-136):  w(c) = exp(self(a) + cross(a, c)) vec_c
//        over every layer state c, a the current state;
//   XAM, XDM (DumpConfig/dumpreduce.cpp:47-122, 516-590): w(o) = (T e_a)_o vec_o over
//        the orbits o, a the current orbit (one operator apply on a unit
//        vector), then an orbit member s drawn with weight exp(cross(a_state, s)).
// Each draw is upper_bound(cumsum(w), u * sum(w)): the first index whose
// running sum exceeds u times the total, u one of the caller's uniforms.
//
// A weighted draw is two kernels: every workgroup sums the weights of one
// contiguous chunk (coalesced), then one workgroup scans the chunk totals,
// re-sums the chosen chunk in 1024 contiguous segments, scans those and walks
// the chosen segment.  The whole chain stays on the device (the current state
// is read by the next step's kernels), so a chain of any length is one queue
// of launches with no host round trip.
#include <algorithm>

#include "tm_device.hpp"

namespace ltm {
namespace {

constexpr int SEL_THREADS = 256;
constexpr int SEL_MAXBLK = 1024;
constexpr int FIND_THREADS = 1024;

enum SelMode { SEL_SQUARE = 0, SEL_XAT = 1, SEL_PRODUCT = 2 };

struct SelArgs {
  int mode;
  const double *vec;     // |eigenvector|
  const double *y;       // SEL_PRODUCT: T e_a
  int64_t n;
  int64_t chunk;         // elements per workgroup (multiple of SEL_THREADS)
  int L;
  double j0, j1, j2;     // betaJ, -kappa betaJ, h betaJ (dumpconfig.cpp:61-63)
  const uint64_t *cur;   // SEL_XAT: the current state
};

__device__ __forceinline__ uint64_t rotr1(uint64_t s, int L) { return (s >> 1) | ((s & 1ull) << (L - 1)); }
__device__ __forceinline__ uint64_t rotr2(uint64_t s, int L) { return (s >> 2) | ((s & 3ull) << (L - 2)); }

// XAT2DDumpConfig::selfintr (dumpconfig.cpp:73-82): field, NN and NNN in
// the layer, periodic.
__device__ __forceinline__ double xat_self(uint64_t a, const SelArgs &p)
{
  const int L = p.L;
  double t = p.j2 * (L - 2.0 * __popcll(a));
  t += p.j0 * (L - 2.0 * __popcll(a ^ rotr1(a, L)));
  t += p.j1 * (L - 2.0 * __popcll(a ^ rotr2(a, L)));
  return t;
}

// Boltzmann factor table of the XAT move from a: e[k] = exp(self(a) + j0 (L - 2k))
// for the k sites where a and c differ (crossintr, dumpconfig.cpp:83-86).
__device__ void xat_table(const SelArgs &p, double *e)
{
  const uint64_t a = *p.cur;
  const double s = xat_self(a, p);
  for (int k = threadIdx.x; k <= p.L; k += blockDim.x) e[k] = exp(s + p.j0 * (p.L - 2.0 * k));
}

__device__ __forceinline__ double weight(const SelArgs &p, int64_t i, const double *e, uint64_t a)
{
  const double v = p.vec[i];
  if (p.mode == SEL_SQUARE) return v * v;
  if (p.mode == SEL_XAT) return e[__popcll(a ^ (uint64_t)i)] * v;
  return p.y[i] * v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__global__ __launch_bounds__(SEL_THREADS) void k_sel_partials(SelArgs p, double *__restrict__ partial)
{
  __shared__ double e[65];
  __shared__ double red[SEL_THREADS / 64];
  uint64_t a = 0;
  if (p.mode == SEL_XAT) {
    xat_table(p, e);
    a = *p.cur;
    __syncthreads();
  }
  const int64_t b0 = (int64_t)blockIdx.x * p.chunk;
  const int64_t b1 = b0 + p.chunk < p.n ? b0 + p.chunk : p.n;
  double acc = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += SEL_THREADS) acc += weight(p, i, e, a);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < SEL_THREADS / 64; ++w) t += red[w];
    partial[blockIdx.x] = t;
  }
}

// Inclusive scan of s[0..FIND_THREADS) in LDS (fixed order).
__device__ void block_scan(double *s)
{
  for (int off = 1; off < FIND_THREADS; off <<= 1) {
    const double v = threadIdx.x >= off ? s[threadIdx.x - off] : 0.0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
}

// First position i in [0, m) with base + incl[i] > t; if none (rounding at
// the very top), the last position with a positive increment (0 if none).
__device__ int first_above(const double *incl, int m, double base, double t, int *slot)
{
  if (threadIdx.x == 0) *slot = m;
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += blockDim.x)
    if (base + incl[i] > t) atomicMin(slot, i);
  __syncthreads();
  int r = *slot;
  __syncthreads();
  if (r < m) return r;
  if (threadIdx.x == 0) {
    int last = 0;
    for (int i = m - 1; i > 0; --i)
      if (incl[i] > incl[i - 1]) { last = i; break; }
    *slot = last;
  }
  __syncthreads();
  r = *slot;
  __syncthreads();
  return r;
}

// One workgroup: the index drawn with probability w_i / sum(w) for the
// uniform u[ui], written to out[0].
__global__ __launch_bounds__(FIND_THREADS) void k_sel_find(SelArgs p, const double *__restrict__ partial,
                                                           int nblk, const double *__restrict__ u,
                                                           int64_t ui, uint64_t *__restrict__ out)
{
  __shared__ double s[FIND_THREADS];
  __shared__ double e[65];
  __shared__ int slot;
  uint64_t a = 0;
  if (p.mode == SEL_XAT) {
    xat_table(p, e);
    a = *p.cur;
  }
  s[threadIdx.x] = threadIdx.x < nblk ? partial[threadIdx.x] : 0.0;
  __syncthreads();
  block_scan(s);
  const double total = s[nblk - 1];
  const double t = u[ui] * total;
  const int b = first_above(s, nblk, 0.0, t, &slot);
  const double base = b ? s[b - 1] : 0.0;
  __syncthreads();
  // the chosen chunk in FIND_THREADS contiguous segments
  const int64_t c0 = (int64_t)b * p.chunk;
  const int64_t c1 = c0 + p.chunk < p.n ? c0 + p.chunk : p.n;
  const int64_t seg = (c1 - c0 + FIND_THREADS - 1) / FIND_THREADS;
  const int64_t s0 = c0 + threadIdx.x * seg;
  const int64_t s1 = s0 + seg < c1 ? s0 + seg : c1;
  double acc = 0;
  for (int64_t i = s0; i < s1; ++i) acc += weight(p, i, e, a);
  s[threadIdx.x] = acc;
  __syncthreads();
  block_scan(s);
  const int g = first_above(s, FIND_THREADS, base, t, &slot);
  if (threadIdx.x == 0) {
    double run = base + (g ? s[g - 1] : 0.0);
    const int64_t g0 = c0 + g * seg;
    const int64_t g1 = g0 + seg < c1 ? g0 + seg : c1;
    int64_t pick = -1, lastpos = g0 < c1 ? g0 : c1 - 1;
    for (int64_t i = g0; i < g1; ++i) {
      const double w = weight(p, i, e, a);
      if (w > 0) lastpos = i;
      run += w;
      if (run > t) {
        pick = i;
        break;
      }
    }
    out[0] = (uint64_t)(pick >= 0 ? pick : lastpos);
  }
}

// e[idx] = val, idx read on the device.
__global__ void k_set_unit(double *__restrict__ e, const uint64_t *__restrict__ idx, double val)
{
  e[*idx] = val;
}

// The member of orbit o = *qa drawn with weight exp(cross(a, s)), members in
// ascending order, a = *prev; written to *out.  Members: the rotations of the
// representative and of its mirror image (periodic) or the two mirror images
// (open), each with its spin flip when the flip is folded.
__global__ __launch_bounds__(64) void k_orbit_pick(const uint64_t *__restrict__ qa,
                                                   const uint64_t *__restrict__ reps,
                                                   const uint16_t *__restrict__ dups, int L,
                                                   int bc, int censym, int dnni, double j0,
                                                   double j1, const uint64_t *__restrict__ prev,
                                                   const double *__restrict__ u, int64_t ui,
                                                   uint64_t *__restrict__ out, int *__restrict__ err)
{
  __shared__ uint64_t cand[256];
  __shared__ uint64_t mem[256];
  __shared__ double cum[256];
  __shared__ bool first[256];
  const uint64_t mask = (L >= 64) ? ~0ull : (1ull << L) - 1;
  const uint64_t rep = reps[*qa];
  uint64_t rev = 0;
  for (int i = 0; i < L; ++i) rev |= ((rep >> i) & 1ull) << (L - 1 - i);
  const int nbase = bc == 0 ? 2 * L : 2;
  const int nc = nbase * (censym ? 2 : 1);
  for (int c = threadIdx.x; c < nc; c += 64) {
    const int b = c % nbase;
    uint64_t s;
    if (bc == 0) {
      const uint64_t src = b < L ? rep : rev;
      const int r = b % L;
      s = r ? (((src >> r) | (src << (L - r))) & mask) : src;
    } else {
      s = b ? rev : rep;
    }
    if (c >= nbase) s = ~s & mask;
    cand[c] = s;
  }
  __syncthreads();
  // distinct members in ascending order: rank among the first occurrences
  for (int c = threadIdx.x; c < nc; c += 64) {
    bool f = true;
    for (int d = 0; d < c; ++d)
      if (cand[d] == cand[c]) { f = false; break; }
    first[c] = f;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nc; c += 64) {
    if (!first[c]) continue;
    int rank = 0;
    for (int d = 0; d < nc; ++d) rank += (first[d] && cand[d] < cand[c]);
    mem[rank] = cand[c];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int c = 0; c < nc; ++c) m += first[c];
    if (m != (int)dups[*qa]) *err = 1;
    const uint64_t a = *prev;
    double tot = 0;
    for (int i = 0; i < m; ++i) {
      const uint64_t s = mem[i];
      double x = j0 * (L - 2.0 * __popcll(a ^ s));
      if (dnni) {  // XDM2DDumpConfig::crossintr (dumpreduce.cpp:509-514)
        const uint64_t sr = ((s >> 1) | (s << (L - 1))) & mask;  // offsetState(s, 1)
        const uint64_t sl = ((s << 1) | (s >> (L - 1))) & mask;  // lffsetState(s, 1)
        x += j1 * (L - 2.0 * __popcll(a ^ sr));
        x += j1 * (L - 2.0 * __popcll(a ^ sl));
      }
      tot += exp(x);
      cum[i] = tot;
    }
    const double t = u[ui] * tot;
    int pick = m - 1;
    for (int i = 0; i < m; ++i)
      if (cum[i] > t) {
        pick = i;
        break;
      }
    out[0] = mem[pick];
  }
}

__global__ void k_rep_of(const uint64_t *__restrict__ qa, const uint64_t *__restrict__ reps,
                         uint64_t *__restrict__ out)
{
  out[0] = reps[*qa];
}

}  // namespace

void sample_chains(tm_op *op, const double *vec, const double *u, int nchain, int length,
                   uint64_t *states, int *err)
{
  hipStream_t s = op->stream;
  const int64_t n = op->ncol;
  const bool red = op->kind != TM_KIND_XAT;
  int64_t nblk = (n + SEL_THREADS * 4 - 1) / (SEL_THREADS * 4);
  nblk = std::max<int64_t>(1, std::min<int64_t>(nblk, SEL_MAXBLK));
  int64_t chunk = (n + nblk - 1) / nblk;
  chunk = (chunk + SEL_THREADS - 1) / SEL_THREADS * SEL_THREADS;
  nblk = (n + chunk - 1) / chunk;
  DevBuf partial(sizeof(double) * SEL_MAXBLK), qa(sizeof(uint64_t)), e, y;
  if (red) {
    e.alloc(sizeof(double) * n);
    y.alloc(sizeof(double) * n);
    TM_HIP(hipMemsetAsync(e.get(), 0, sizeof(double) * n, s));
  }
  const tm_params &pp = op->p;
  SelArgs a{};
  a.vec = vec;
  a.n = n;
  a.chunk = chunk;
  a.L = op->L;
  a.j0 = pp.betaJ;
  a.j1 = -pp.kappa * pp.betaJ;
  a.j2 = pp.h * pp.betaJ;
  auto draw = [&](int mode, const uint64_t *cur, const double *yv, int64_t ui, uint64_t *out) {
    a.mode = mode;
    a.cur = cur;
    a.y = yv;
    hipLaunchKernelGGL(k_sel_partials, dim3((unsigned)nblk), dim3(SEL_THREADS), 0, s, a,
                       partial.as<double>());
    TM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sel_find, dim3(1), dim3(FIND_THREADS), 0, s, a, partial.as<double>(),
                       (int)nblk, u, ui, out);
    TM_LAUNCH_CHECK();
  };
  const int per = red ? 2 * length - 1 : length;
  for (int c = 0; c < nchain; ++c) {
    uint64_t *st = states + (int64_t)c * length;
    const int64_t ub = (int64_t)c * per;
    if (!red) {
      draw(SEL_SQUARE, nullptr, nullptr, ub, st);
      for (int j = 1; j < length; ++j) draw(SEL_XAT, st + j - 1, nullptr, ub + j, st + j);
      continue;
    }
    uint64_t *q = qa.as<uint64_t>();
    draw(SEL_SQUARE, nullptr, nullptr, ub, q);
    hipLaunchKernelGGL(k_rep_of, dim3(1), dim3(1), 0, s, q, op->reps.as<uint64_t>(), st);
    TM_LAUNCH_CHECK();
    for (int j = 1; j < length; ++j) {
      hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, s, e.as<double>(), q, 1.0);
      TM_LAUNCH_CHECK();
      op->apply_dev(e.get(), y.get(), n, 1, TM_F64, TM_F64, 0, nullptr, s);
      hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, s, e.as<double>(), q, 0.0);
      TM_LAUNCH_CHECK();
      draw(SEL_PRODUCT, nullptr, y.as<double>(), ub + 2 * j - 1, q);
      hipLaunchKernelGGL(k_orbit_pick, dim3(1), dim3(64), 0, s, q, op->reps.as<uint64_t>(),
                         op->dups.as<uint16_t>(), op->L, op->bc & 1, op->censym,
                         op->kind == TM_KIND_XDM ? 1 : 0, a.j0, a.j1, st + j - 1, u, ub + 2 * j,
                         st + j, err);
      TM_LAUNCH_CHECK();
    }
  }
}

}  // namespace ltm
