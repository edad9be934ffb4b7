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

enum SelMode { SEL_SQUARE = 0, SEL_XAT = 1, SEL_PRODUCT = 2, SEL_PAIR = 3, SEL_YAT = 4 };

struct SelArgs {
  int mode;
  const double *vec;     // |eigenvector|
  const double *vecl;    // SEL_PAIR: |left eigenvector|
  const double *y;       // SEL_PRODUCT: T e_a
  int64_t n;
  int64_t chunk;         // elements per workgroup (multiple of SEL_THREADS)
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

__global__ __launch_bounds__(SEL_THREADS) void k_sel_partials(SelArgs p, double *__restrict__ partial)
{
  __shared__ double e[65];
  __shared__ double red[SEL_THREADS / 64];
  uint64_t a = 0;
  if (p.mode == SEL_XAT || p.mode == SEL_YAT) {
    move_table(p, e);
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
  if (p.mode == SEL_XAT || p.mode == SEL_YAT) {
    move_table(p, e);
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
    const uint64_t r = (uint64_t)(pick >= 0 ? pick : lastpos);
    // YAT: (b, a) -> (c, b)
    out[0] = p.mode == SEL_YAT ? ((r << p.L) | (a >> p.L)) : r;
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
      bool ok;
      const double x = member_cross(ck, a, mem[i], L, lm, c0, c1, c2, &ok);
      if (ok) tot += exp(x);
      cum[i] = tot;
    }
    const double t = u[ui] * tot;
    int pick = -1, lastpos = 0;
    for (int i = 0; i < m; ++i) {
      if (cum[i] > (i ? cum[i - 1] : 0.0)) lastpos = i;
      if (pick < 0 && cum[i] > t) pick = i;
    }
    out[0] = mem[pick >= 0 ? pick : lastpos];
  }
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
  DevBuf partial(sizeof(double) * SEL_MAXBLK), qa(sizeof(uint64_t)), e, y;
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
    int64_t nblk = (m + SEL_THREADS * 4 - 1) / (SEL_THREADS * 4);
    nblk = std::max<int64_t>(1, std::min<int64_t>(nblk, SEL_MAXBLK));
    int64_t chunk = (m + nblk - 1) / nblk;
    chunk = (chunk + SEL_THREADS - 1) / SEL_THREADS * SEL_THREADS;
    nblk = (m + chunk - 1) / chunk;
    a.mode = mode;
    a.n = m;
    a.chunk = chunk;
    a.cur = cur;
    a.y = yv;
    hipLaunchKernelGGL(k_sel_partials, dim3((unsigned)nblk), dim3(SEL_THREADS), 0, s, a,
                       partial.as<double>());
    TM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sel_find, dim3(1), dim3(FIND_THREADS), 0, s, a, partial.as<double>(),
                       (int)nblk, u, ui, out);
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
