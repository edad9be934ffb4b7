// tm_sample.hip -- planting equilibrium configurations from the leading
// eigenvector (the reference's dumpconfig program) as device Markov chains.
//
// A chain starts at an index drawn with weight vec_i^2 (SymDumpConfig::initq,
// DumpConfig/dumpconfig.cpp:24-35) or vec_i vecl_i (GenDumpConfig::initq,
// :44-56) and moves by draws from
//   XAT  (DumpConfig/dumpconfig.cpp:88-136):  w(c) = exp(self(a) + cross(a, c)) vec_c
//        over every layer state c, a the current state;
//   YAT  (DumpConfig/dumpconfig.cpp:171-222): the two-layer state (b, a) moves
//        to (c, b) with w(c) = exp(self(a, b) + cross(a, c)) vec_{(c, b)};
//   XAM, XDM, ZBM (DumpConfig/dumpreduce.cpp:47-122, 516-590, 413-470):
//        w(o) = (T e_a)_o vec_o over the orbits o, a the current orbit (one
//        operator apply on a unit vector), then an orbit member s drawn with
//        weight exp(cross(a_state, s));
//   YAM, YBM (DumpConfig/dumpreduce.cpp:156-233, 279-356): the same, the
//        member restricted to those whose low layer is the current high layer.
// Each draw is upper_bound(cumsum(w), u * sum(w)): the first index whose
// running sum exceeds u times the total, u one of the caller's uniforms.
//
// A weighted draw is two kernels: every workgroup sums the weights of one
// contiguous chunk of 2048 or 4096 elements (coalesced), then one workgroup
// finds the index in three levels -- groups of consecutive chunk totals, the
// chunks of the chosen group, 2 or 4 elements per thread of the chosen
// chunk -- each a 1024-wide
// scan in LDS after a few independent loads per thread.  The whole chain stays on the device (the current state
// is read by the next step's kernels), so a chain of any length is one queue
// of launches with no host round trip.
#include <algorithm>

#include "tm_device.hpp"

namespace ltm {
namespace {

constexpr int SEL_THREADS = 256;
constexpr int FIND_THREADS = 1024;
// elements per thread of the chunk sums: 16 for large draws (4096-element
// chunks, fewer and longer streams), 8 below SEL_BIG (enough workgroups)
constexpr int64_t SEL_BIG = int64_t(1) << 23;
constexpr int64_t SEL_MAXN = int64_t(SEL_THREADS) * 16 * FIND_THREADS * FIND_THREADS;  // 2^32

enum SelMode { SEL_SQUARE = 0, SEL_XAT = 1, SEL_PRODUCT = 2, SEL_PAIR = 3, SEL_YAT = 4 };

struct SelArgs {
  int mode;
  const double *vec;     // |eigenvector|
  const double *vecl;    // SEL_PAIR: |left eigenvector|
  const double *y;       // SEL_PRODUCT: T e_a
  int64_t n;
  int L;
  double j0, j1, j2;     // betaJ, -kappa betaJ, h betaJ (dumpconfig.cpp:61-63, 141-143)
  const uint64_t *cur;   // SEL_XAT, SEL_YAT: the current state
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

// YAT2DDumpConfig::selfintr (dumpconfig.cpp:154-163): field and NN in the
// low layer a, NN bonds to the high layer b.
__device__ __forceinline__ double yat_self(uint64_t a, uint64_t b, const SelArgs &p)
{
  const int L = p.L;
  double t = p.j2 * (L - 2.0 * __popcll(a));
  t += p.j0 * (L - 2.0 * __popcll(a ^ rotr1(a, L)));
  t += p.j0 * (L - 2.0 * __popcll(a ^ b));
  return t;
}

// Boltzmann factor table of a move: e[k] = exp(self + J (L - 2k)) for the k
// sites where the current and the new layer differ (XAT: crossintr
// dumpconfig.cpp:83-86 with J = j0; YAT: :165-168 with J = j1).
__device__ void move_table(const SelArgs &p, double *e)
{
  const uint64_t a = *p.cur;
  const uint64_t lm = (1ull << p.L) - 1;
  double s, J;
  if (p.mode == SEL_XAT) {
    s = xat_self(a, p);
    J = p.j0;
  } else {
    s = yat_self(a & lm, a >> p.L, p);
    J = p.j1;
  }
  for (int k = threadIdx.x; k <= p.L; k += blockDim.x) e[k] = exp(s + J * (p.L - 2.0 * k));
}

// Weight of index i; for SEL_YAT, i is the new high layer c and a the
// current two-layer state (b << L) | a_low.
__device__ __forceinline__ double weight(const SelArgs &p, int64_t i, const double *e, uint64_t a)
{
  if (p.mode == SEL_YAT) {
    const uint64_t lm = (1ull << p.L) - 1;
    return e[__popcll((a & lm) ^ (uint64_t)i)] * p.vec[((uint64_t)i << p.L) | (a >> p.L)];
  }
  const double v = p.vec[i];
  if (p.mode == SEL_SQUARE) return v * v;
  if (p.mode == SEL_PAIR) return v * p.vecl[i];
  if (p.mode == SEL_XAT) return e[__popcll(a ^ (uint64_t)i)] * v;
  return p.y[i] * v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int PER>
__global__ __launch_bounds__(SEL_THREADS) void k_sel_partials(SelArgs p, double *__restrict__ partial)
{
  constexpr int SEL_CHUNK = SEL_THREADS * PER;
  __shared__ double e[65];
  __shared__ double red[SEL_THREADS / 64];
  uint64_t a = 0;
  if (p.mode == SEL_XAT || p.mode == SEL_YAT) {
    move_table(p, e);
    a = *p.cur;
    __syncthreads();
  }
  const int64_t b0 = (int64_t)blockIdx.x * SEL_CHUNK;
  double w[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t i = b0 + k * SEL_THREADS + threadIdx.x;
    w[k] = i < p.n ? weight(p, i, e, a) : 0.0;
  }
  double acc = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) acc += w[k];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < SEL_THREADS / 64; ++q) t += red[q];
    partial[blockIdx.x] = t;
  }
}

// Inclusive scan of s[0..FIND_THREADS) in LDS (fixed order): each wave
// scans its 64 values with shuffles, wave 0 scans the 16 wave totals, every
// wave adds the total of the waves before it.
__device__ void block_scan(double *s)
{
  __shared__ double wsum[FIND_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double v = s[threadIdx.x];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (w == 0) {
    double x = lane < FIND_THREADS / 64 ? wsum[lane] : 0.0;
#pragma unroll
    for (int o = 1; o < FIND_THREADS / 64; o <<= 1) {
      const double t = __shfl_up(x, o, 64);
      if (lane >= o) x += t;
    }
    if (lane < FIND_THREADS / 64) wsum[lane] = x;
  }
  __syncthreads();
  if (w > 0) v += wsum[w - 1];
  s[threadIdx.x] = v;
  __syncthreads();
}

// First position i in [0, m) with base + incl[i] > t; if none (rounding at
// the very top), the last position with a positive increment (0 if none).
// incl is a running sum of non-negative weights, so base + incl[i] never
// decreases and the first position above t is the count of positions at or
// below it (one LDS atomic per wave).
__device__ int first_above(const double *incl, int m, double base, double t, int *slot)
{
  if (threadIdx.x == 0) *slot = 0;
  __syncthreads();
  int cnt = 0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) cnt += (base + incl[i] <= t) ? 1 : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(slot, cnt);
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
template <int PER>
__global__ __launch_bounds__(FIND_THREADS) void k_sel_find(SelArgs p, const double *__restrict__ partial,
                                                           int64_t nblk, const double *__restrict__ u,
                                                           int64_t ui, uint64_t *__restrict__ out)
{
  constexpr int SEL_CHUNK = SEL_THREADS * PER;
  constexpr int FIND_EPT = SEL_CHUNK / FIND_THREADS;  // elements per thread in the last level
  __shared__ double s[FIND_THREADS];
  __shared__ double e[65];
  __shared__ int slot;
  uint64_t a = 0;
  if (p.mode == SEL_XAT || p.mode == SEL_YAT) {
    move_table(p, e);
    a = *p.cur;
  }
  const int j = threadIdx.x;
  // level 1: groups of pp consecutive chunk totals
  const int64_t pp = (nblk + FIND_THREADS - 1) / FIND_THREADS;
  {
    double acc = 0;
    const int64_t g0 = (int64_t)j * pp;
    for (int64_t q = 0; q < pp; q += 8) {
      double v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int64_t c = g0 + q + r;
        v[r] = (q + r < pp && c < nblk) ? partial[c] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) acc += v[r];
    }
    s[j] = acc;
  }
  __syncthreads();
  block_scan(s);
  const double total = s[FIND_THREADS - 1];
  const double t = u[ui] * total;
  const int g = first_above(s, FIND_THREADS, 0.0, t, &slot);
  double base = g ? s[g - 1] : 0.0;
  __syncthreads();
  // level 2: the chunks of group g
  {
    const int64_t c = (int64_t)g * pp + j;
    s[j] = (j < pp && c < nblk) ? partial[c] : 0.0;
  }
  __syncthreads();
  block_scan(s);
  const int64_t b = (int64_t)g * pp + first_above(s, FIND_THREADS, base, t, &slot);
  {
    const int ib = (int)(b - (int64_t)g * pp);
    base += ib ? s[ib - 1] : 0.0;
  }
  __syncthreads();
  // level 3: FIND_EPT consecutive elements per thread in chunk b
  const int64_t i0 = b * SEL_CHUNK + (int64_t)FIND_EPT * j;
  double w[FIND_EPT];
  double acc = 0;
#pragma unroll
  for (int r = 0; r < FIND_EPT; ++r) {
    w[r] = i0 + r < p.n ? weight(p, i0 + r, e, a) : 0.0;
    acc += w[r];
  }
  s[j] = acc;
  __syncthreads();
  block_scan(s);
  const int k = first_above(s, FIND_THREADS, base, t, &slot);
  if (j == k) {
    double run = base + (k ? s[k - 1] : 0.0);
    int64_t pick = -1, lastp = i0;
#pragma unroll
    for (int r = 0; r < FIND_EPT; ++r) {
      if (w[r] > 0) lastp = i0 + r;
      run += w[r];
      if (pick < 0 && w[r] > 0 && run > t) pick = i0 + r;
    }
    const int64_t rr = pick >= 0 ? pick : lastp;
    // YAT: (b, a) -> (c, b)
    out[0] = p.mode == SEL_YAT ? (((uint64_t)rr << p.L) | (a >> p.L)) : (uint64_t)rr;
  }
}

// e[idx] = val, idx read on the device.
__global__ void k_set_unit(double *__restrict__ e, const uint64_t *__restrict__ idx, double val)
{
  e[*idx] = val;
}

enum OrbitGroup { GRP_X = 0, GRP_Z = 1, GRP_Y = 2 };
enum CrossKind { CR_XAM = 0, CR_XDM = 1, CR_ZBM = 2, CR_Y = 3 };

__device__ __forceinline__ uint64_t rot_r(uint64_t s, int r, int L, uint64_t mask)
{
  return r ? (((s >> r) | (s << (L - r))) & mask) : s;
}

__device__ __forceinline__ uint64_t brev(uint64_t s, int L)
{
  uint64_t r = 0;
  for (int i = 0; i < L; ++i) r |= ((s >> i) & 1ull) << (L - 1 - i);
  return r;
}

// The exponent of a member's weight, a the current state, s the member.
__device__ double member_cross(int ck, uint64_t a, uint64_t s, int L, uint64_t lm, double c0,
                               double c1, double c2, bool *allowed)
{
  *allowed = true;
  if (ck == CR_Y) {  // YAM / YBM crossintr (dumpreduce.cpp:151-154, 274-277)
    if ((s & lm) != (a >> L)) {
      *allowed = false;
      return 0.0;
    }
    return c1 * (L - 2.0 * __popcll((a & lm) ^ (s >> L)));
  }
  if (ck == CR_ZBM) {  // ZBM2DDumpConfig::crossintr (dumpreduce.cpp:389-407)
    double x = c2 * (L - 2.0 * __popcll(a ^ s));
    x += c1 * (L - 2.0 * __popcll(a ^ rot_r(s, 2, L, lm)));
    x += c1 * (L - 2.0 * __popcll(a ^ rot_r(s, L - 2, L, lm)));
    // even sites i of a against sites i -+ 1 of s
    const uint64_t sl = ((s << 1) | (s >> (L - 1))) & lm;  // bit i = s_{i-1}
    const uint64_t sr = rot_r(s, 1, L, lm);                // bit i = s_{i+1}
    uint64_t even = 0;
    for (int i = 0; i < L; i += 2) even |= 1ull << i;
    const int nanti = __popcll((a ^ sl) & even) + __popcll((a ^ sr) & even);
    return x + c0 * (L - 2.0 * nanti);
  }
  double x = c0 * (L - 2.0 * __popcll(a ^ s));
  if (ck == CR_XDM) {  // XDM2DDumpConfig::crossintr (dumpreduce.cpp:509-514)
    x += c1 * (L - 2.0 * __popcll(a ^ rot_r(s, 1, L, lm)));      // offsetState(s, 1)
    x += c1 * (L - 2.0 * __popcll(a ^ rot_r(s, L - 1, L, lm)));  // lffsetState(s, 1)
  }
  return x;
}

// The member of orbit o = *qa drawn with weight exp(cross(a, s)), members in
// ascending order, a = *prev; written to *out.  Members: the images of the
// representative under the operator's group (X: rotations and mirror image,
// or the two mirror images with open boundary; Z: rotations by two sites and
// the mirror i -> -i; Y: both layers of a 2L-bit state rotated and mirrored
// together), each with its spin flip when the flip is folded.  Members a
// draw cannot reach (Y: low layer other than the current high layer) weigh 0.
__global__ __launch_bounds__(64) void k_orbit_pick(const uint64_t *__restrict__ qa,
                                                   const uint64_t *__restrict__ reps,
                                                   const uint16_t *__restrict__ dups, int L,
                                                   int grp, int bc, int censym, int ck, double c0,
                                                   double c1, double c2,
                                                   const uint64_t *__restrict__ prev,
                                                   const double *__restrict__ u, int64_t ui,
                                                   uint64_t *__restrict__ out, int *__restrict__ err)
{
  __shared__ uint64_t cand[256];
  __shared__ uint64_t mem[256];
  __shared__ double cum[256];
  __shared__ bool first[256];
  __shared__ int nmem;
  __shared__ double tgt;
  const int nbits = grp == GRP_Y ? 2 * L : L;
  const uint64_t lm = (1ull << L) - 1;
  const uint64_t full = nbits >= 64 ? ~0ull : (1ull << nbits) - 1;
  const uint64_t rep = reps[*qa];
  uint64_t mir;
  int nbase;
  if (grp == GRP_Y) {
    mir = (brev(rep >> L, L) << L) | brev(rep & lm, L);
    nbase = 2 * L;
  } else if (grp == GRP_Z) {
    mir = brev(rot_r(rep, 1, L, lm), L);
    nbase = L;  // L/2 rotations by two sites of each
  } else {
    mir = brev(rep, L);
    nbase = bc == 0 ? 2 * L : 2;
  }
  const int nc = nbase * (censym ? 2 : 1);
  for (int c = threadIdx.x; c < nc; c += 64) {
    const int b = c % nbase;
    uint64_t s;
    if (grp == GRP_Y) {
      const uint64_t src = b < L ? rep : mir;
      const int r = b % L;
      s = (rot_r(src >> L, r, L, lm) << L) | rot_r(src & lm, r, L, lm);
    } else if (grp == GRP_Z) {
      const int h = L / 2;
      s = rot_r(b < h ? rep : mir, 2 * (b % h), L, lm);
    } else if (bc == 0) {
      s = rot_r(b < L ? rep : mir, b % L, L, lm);
    } else {
      s = b ? mir : rep;
    }
    if (c >= nbase) s = ~s & full;
    cand[c] = s;
  }
  __syncthreads();
  // distinct members in ascending order: rank among the first occurrences
  // (loops without early exit, unrolled, so that the LDS reads overlap)
  for (int c = threadIdx.x; c < nc; c += 64) {
    const uint64_t v = cand[c];
    bool dup = false;
#pragma unroll 8
    for (int d = 0; d < c; ++d) dup |= cand[d] == v;
    first[c] = !dup;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nc; c += 64) {
    if (!first[c]) continue;
    const uint64_t v = cand[c];
    int rank = 0;
#pragma unroll 8
    for (int d = 0; d < nc; ++d) rank += (first[d] && cand[d] < v);
    mem[rank] = v;
  }
  if (threadIdx.x == 0) {
    int m = 0;
    for (int c = 0; c < nc; ++c) m += first[c];
    nmem = m;
  }
  __syncthreads();
  const int m = nmem;
  // member weights on every lane, summed in ascending order by one lane
  const uint64_t a = *prev;
  for (int i = threadIdx.x; i < m; i += 64) {
    bool ok;
    const double x = member_cross(ck, a, mem[i], L, lm, c0, c1, c2, &ok);
    cum[i] = ok ? exp(x) : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (m != (int)dups[*qa]) *err = 1;
    double tot = 0;
#pragma unroll 8
    for (int i = 0; i < m; ++i) {
      tot += cum[i];
      cum[i] = tot;
    }
    tgt = u[ui] * tot;
  }
  __syncthreads();
  // first i with cum[i] > t (cum never decreases: the count of entries at or
  // below t); else the last with a positive increment
  int cnt = 0, last = 0;
  for (int i = threadIdx.x; i < m; i += 64) {
    cnt += cum[i] <= tgt ? 1 : 0;
    if (cum[i] > (i ? cum[i - 1] : 0.0)) last = i;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cnt += __shfl_down(cnt, o, 64);
    last = max(last, __shfl_down(last, o, 64));
  }
  if (threadIdx.x == 0) out[0] = mem[cnt < m ? cnt : last];
}

__global__ void k_rep_of(const uint64_t *__restrict__ qa, const uint64_t *__restrict__ reps,
                         uint64_t *__restrict__ out)
{
  out[0] = reps[*qa];
}

}  // namespace

void sample_chains(tm_op *op, const double *vec, const double *vecl, const double *u, int nchain,
                   int length, uint64_t *states, int *err)
{
  hipStream_t s = op->stream;
  const tm_kind kind = op->kind;
  const int64_t n = op->ncol;
  const bool red = is_reduced(kind);
  const bool yat = kind == TM_KIND_YAT;
  const int64_t mmax = std::max<int64_t>(n, kind == TM_KIND_YAT ? (int64_t(1) << op->L) : 0);
  if (mmax > SEL_MAXN) throw Error(TM_EUNSUPPORTED, "sampling supports up to 2^32 states");
  // chunk totals of the largest draw (8 elements per thread below SEL_BIG)
  const int64_t pmax = std::max<int64_t>((mmax + SEL_THREADS * 16 - 1) / (SEL_THREADS * 16),
                                         (std::min<int64_t>(mmax, SEL_BIG) + SEL_THREADS * 8 - 1) /
                                             (SEL_THREADS * 8));
  DevBuf partial(sizeof(double) * pmax), qa(sizeof(uint64_t)), e, y;
  if (red) {
    e.alloc(sizeof(double) * n);
    y.alloc(sizeof(double) * n);
    TM_HIP(hipMemsetAsync(e.get(), 0, sizeof(double) * n, s));
  }
  const tm_params &pp = op->p;
  SelArgs a{};
  a.vec = vec;
  a.vecl = vecl;
  a.L = op->L;
  a.j0 = pp.betaJ;
  a.j1 = -pp.kappa * pp.betaJ;
  a.j2 = pp.h * pp.betaJ;
  // one weighted draw over [0, m)
  auto draw = [&](int mode, int64_t m, const uint64_t *cur, const double *yv, int64_t ui,
                  uint64_t *out) {
    const bool big = m >= SEL_BIG;
    const int64_t chunk = SEL_THREADS * (big ? 16 : 8);
    const int64_t nblk = (m + chunk - 1) / chunk;
    a.mode = mode;
    a.n = m;
    a.cur = cur;
    a.y = yv;
    if (big) {
      hipLaunchKernelGGL(k_sel_partials<16>, dim3((unsigned)nblk), dim3(SEL_THREADS), 0, s, a,
                         partial.as<double>());
      TM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sel_find<16>, dim3(1), dim3(FIND_THREADS), 0, s, a,
                         partial.as<double>(), nblk, u, ui, out);
    } else {
      hipLaunchKernelGGL(k_sel_partials<8>, dim3((unsigned)nblk), dim3(SEL_THREADS), 0, s, a,
                         partial.as<double>());
      TM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_sel_find<8>, dim3(1), dim3(FIND_THREADS), 0, s, a,
                         partial.as<double>(), nblk, u, ui, out);
    }
    TM_LAUNCH_CHECK();
  };
  const int start = vecl ? SEL_PAIR : SEL_SQUARE;
  // orbit group and member weights of the reduced kinds
  int grp = GRP_X, ck = CR_XAM;
  double c0 = pp.betaJ, c1 = -pp.kappa * pp.betaJ, c2 = 0.0;
  if (kind == TM_KIND_XDM) ck = CR_XDM;
  if (kind == TM_KIND_ZBM) {
    grp = GRP_Z;
    ck = CR_ZBM;
    c2 = -pp.kappa2 * pp.betaJ;  // dumpreduce.cpp:367-370
  }
  if (kind == TM_KIND_YAM || kind == TM_KIND_YBM) {
    grp = GRP_Y;
    ck = CR_Y;
  }
  const int per = red ? 2 * length - 1 : length;
  for (int c = 0; c < nchain; ++c) {
    uint64_t *st = states + (int64_t)c * length;
    const int64_t ub = (int64_t)c * per;
    if (!red) {
      draw(start, n, nullptr, nullptr, ub, st);
      for (int j = 1; j < length; ++j)
        draw(yat ? SEL_YAT : SEL_XAT, yat ? (int64_t(1) << op->L) : n, st + j - 1, nullptr,
             ub + j, st + j);
      continue;
    }
    uint64_t *q = qa.as<uint64_t>();
    draw(start, n, nullptr, nullptr, ub, q);
    hipLaunchKernelGGL(k_rep_of, dim3(1), dim3(1), 0, s, q, op->reps.as<uint64_t>(), st);
    TM_LAUNCH_CHECK();
    for (int j = 1; j < length; ++j) {
      hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, s, e.as<double>(), q, 1.0);
      TM_LAUNCH_CHECK();
      op->apply_dev(e.get(), y.get(), n, 1, TM_F64, TM_F64, 0, nullptr, s);
      hipLaunchKernelGGL(k_set_unit, dim3(1), dim3(1), 0, s, e.as<double>(), q, 0.0);
      TM_LAUNCH_CHECK();
      draw(SEL_PRODUCT, n, nullptr, y.as<double>(), ub + 2 * j - 1, q);
      hipLaunchKernelGGL(k_orbit_pick, dim3(1), dim3(64), 0, s, q, op->reps.as<uint64_t>(),
                         op->dups.as<uint16_t>(), op->L, grp, op->bc & 1, op->censym, ck, c0, c1,
                         c2, st + j - 1, u, ub + 2 * j, st + j, err);
      TM_LAUNCH_CHECK();
    }
  }
}

}  // namespace ltm
