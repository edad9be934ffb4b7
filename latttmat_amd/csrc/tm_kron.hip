// tm_kron.hip -- full-space Kronecker transfer-matrix operator on gfx950.
//
//   y = D o (x)_{i<L} K o D o x,   K = [[e^bJ, e^-bJ], [e^-bJ, e^bJ]]
//
// is XATMatVecProd2D::perform_op (TMat/tmat.cpp:71-102) with the normfactor
// split as sqrt(nf) into both diagonals, and the new full-space 3D operator.
// The diagonal D(s) is rebuilt from three popcounts of the state and a
// 192-entry table staged in LDS, so the only HBM traffic is the vector.
//
// The L butterfly stages are grouped into passes.  A pass owns bits
// [b0, b0+nb) of the index: a workgroup loads a tile of 2^nb elements along
// those bits times W = 2^logW contiguous elements (coalescing), applies the
// nb stages in LDS with radix-16 register groups, and stores the tile back.
// The first pass applies D on load, the last applies D on store.
#include <type_traits>

#include "tm_device.hpp"

namespace ltm {
namespace {

// Threads per workgroup: 512, or 1024 for the largest tile (one workgroup
// per CU then still keeps 16 wavefronts resident).  A tile candidate encodes
// the tile exponent in bits 0-7 and the 1024-thread choice in bit 8.
constexpr int KRON_WIDE = 256;
// F_TWO: the prologue and epilogue diagonals differ (two LDS tables)
enum { F_PRO = 1, F_EPI = 2, F_SIN = 4, F_SOUT = 8, F_TWO = 16 };

template <typename C, typename TI, typename TO, int NT>
// in and out may be the same array (passes after the first work in place):
// each workgroup reads its whole tile before the barrier and writes the same
// tile after it, so neither pointer is __restrict__.
__global__ __launch_bounds__(NT) void k_kron_pass(
    const TI *in, TO *out, int b0, int nb, int logW, int flags,
    DiagSpec dg, const C *__restrict__ tab, DiagSpec dge, const C *__restrict__ tabe, C r, C f,
    const double *__restrict__ scale)
{
  extern __shared__ __align__(16) unsigned char smem[];
  const int tileLog = nb + logW;
  const int T = 1 << tileLog;
  const int W = 1 << logW;
  C *lds = reinterpret_cast<C *>(smem);
  C *stab = lds + T;
  C *stabe = (flags & F_TWO) ? stab + 192 : stab;
  if (flags & F_TWO) {
    if (flags & F_PRO)
      for (int i = threadIdx.x; i < 192; i += blockDim.x) stab[i] = tab[i];
    if (flags & F_EPI)
      for (int i = threadIdx.x; i < 192; i += blockDim.x) stabe[i] = tabe[i];
  } else if (flags & (F_PRO | F_EPI)) {
    for (int i = threadIdx.x; i < 192; i += blockDim.x) stab[i] = tab[i];
  }

  const uint64_t tile = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  const uint64_t innerBlocks = (1ull << b0) >> logW;
  const uint64_t o = tile / innerBlocks, ib = tile % innerBlocks;
  const uint64_t base = (o << (b0 + nb)) | (ib << logW);
  C sin = 1, sout = 1;
  if (scale && (flags & (F_SIN | F_SOUT))) {
    const double sc = *scale;
    if (flags & F_SIN) sin = (C)sc;
    if (flags & F_SOUT) sout = (C)(1.0 / sc);
  }
  __syncthreads();

  // UNR independent loads per thread before the LDS stores (several HBM
  // requests in flight per wave)
  constexpr int UNR = 8;
  for (int e0 = threadIdx.x; e0 < T; e0 += blockDim.x * UNR) {
    C v[UNR];
#pragma unroll
    for (int uu = 0; uu < UNR; ++uu) {
      const int e = e0 + uu * blockDim.x;
      v[uu] = 0;
      if (e < T) {
        const uint64_t g = base | ((uint64_t)(e >> logW) << b0) | (uint64_t)(e & (W - 1));
        v[uu] = (C)in[g] * sin;
        if (flags & F_PRO) v[uu] *= dev::diag_value<C>(dg, stab, g);
      }
    }
#pragma unroll
    for (int uu = 0; uu < UNR; ++uu) {
      const int e = e0 + uu * blockDim.x;
      if (e < T) lds[dev::sw(e)] = v[uu];
    }
  }
  __syncthreads();

  // nb stages: LDS index of element (p, w) is p * W + w, so bit j of p is
  // bit (j + logW) of the LDS index; the w bits ride along as lo bits.
  if (logW == 0) {
    // contiguous tile (the first pass): bit groups at compile-time positions
    dev::lds_kron(lds, nb, r);
  } else {
    for (int bg = 0; bg < nb; bg += 4) {
      const int rr = nb - bg < 4 ? nb - bg : 4;
      switch (rr) {
        case 1: dev::lds_kron_group<1>(lds, tileLog, bg + logW, r); break;
        case 2: dev::lds_kron_group<2>(lds, tileLog, bg + logW, r); break;
        case 3: dev::lds_kron_group<3>(lds, tileLog, bg + logW, r); break;
        default: dev::lds_kron_group<4>(lds, tileLog, bg + logW, r); break;
      }
      __syncthreads();
    }
  }

  // f = a^nb of K = a [[1, r], [r, 1]] over this pass's bits
  const C fs = f * sout;
  for (int e = threadIdx.x; e < T; e += blockDim.x) {
    const uint64_t g = base | ((uint64_t)(e >> logW) << b0) | (uint64_t)(e & (W - 1));
    C v = lds[dev::sw(e)] * fs;
    if (flags & F_EPI) v *= dev::diag_value<C>(dge, stabe, g);
    out[g] = (TO)v;
  }
}

// Register-resident pass: a tile of 2^T elements (T = TB + 5, 2^TB threads)
// lives in registers, 32 values per thread.  In each round the five register
// bits are five tile bits [p0, p0 + 5) and get their butterflies in
// registers (only the bits set in act); the thread's own bits fill the other
// tile positions from the lowest up.  Between rounds the tile goes through
// LDS once (write in one layout, read in the next: 2 LDS accesses per element
// per round instead of 2 per 4 bits), and the global load and store use
// layouts whose low tile bits are lane bits (contiguous runs).  Tile index e
// maps to the global index as in k_kron_pass: run bits [0, logW), pass bits
// [b0, b0 + nb).
struct RRound { int p0, act; };
struct RPlan {
  int T, logW, b0, nb, nr;
  int pair;  // 64-byte runs: the two tiles of each 128-byte line on one XCD, back to back
  RRound r[3];
};
constexpr int KRON_RPASS = 512;  // autotuner candidate flag: register-resident plan

template <typename C, typename TI, typename TO, int TB, int DK>
__global__ __launch_bounds__(1 << TB) void k_kron_rpass(
    const TI *in, TO *out, RPlan pl, int flags, DiagSpec dg, const C *__restrict__ tab,
    DiagSpec dge, const C *__restrict__ tabe, C r, C f, const double *__restrict__ scale)
{
  constexpr int RB = 5, NV = 1 << RB;
  extern __shared__ __align__(16) unsigned char smem[];
  const int T = pl.T, logW = pl.logW, b0 = pl.b0, nb = pl.nb;
  C *lds = reinterpret_cast<C *>(smem);
  C *stab = lds + (1 << T);
  C *stabe = (flags & F_TWO) ? stab + 192 : stab;
  if (flags & F_TWO) {
    if (flags & F_PRO)
      for (int i = threadIdx.x; i < 192; i += blockDim.x) stab[i] = tab[i];
    if (flags & F_EPI)
      for (int i = threadIdx.x; i < 192; i += blockDim.x) stabe[i] = tabe[i];
  } else if (flags & (F_PRO | F_EPI)) {
    for (int i = threadIdx.x; i < 192; i += blockDim.x) stab[i] = tab[i];
  }
  // 32-bit element indices (the launcher checks Ltot <= 31)
  uint32_t tile = blockIdx.x + blockIdx.y * gridDim.x;
  if (pl.pair) {
    // workgroup b runs on XCD b mod 8; with 64-byte runs tiles 2t and 2t+1
    // read the two halves of the same lines, so give them workgroups b and
    // b + 8 (one XCD, one L2, launched one after the other)
    const uint32_t x = tile & 7, k = tile >> 3;
    tile = ((((k >> 1) << 3) | x) << 1) | (k & 1);
  }
  const uint32_t innerBlocks = (1u << b0) >> logW;
  const uint32_t o = tile / innerBlocks, ib = tile % innerBlocks;
  const uint32_t base = (o << (b0 + nb)) | (ib << logW);
  const uint32_t W = 1u << logW;
  auto glob = [&](uint32_t e) -> uint32_t {
    return base | ((e >> logW) << b0) | (e & (W - 1));
  };
  // global stride of register bit j in a round with register bits at tile
  // positions [p0, p0 + 5): the element of register index i is glob of the
  // thread's part plus the strides of i's bits (disjoint bits: OR = ADD)
  auto gstride = [&](int q) -> uint32_t { return q < logW ? (1u << q) : (1u << (b0 + q - logW)); };
  const int t = threadIdx.x;
  auto thr = [&](int p0) { return (t & ((1 << p0) - 1)) | ((t >> p0) << (p0 + RB)); };
  C sin = 1, sout = 1;
  if (scale && (flags & (F_SIN | F_SOUT))) {
    const double sc = *scale;
    if (flags & F_SIN) sin = (C)sc;
    if (flags & F_SOUT) sout = (C)(1.0 / sc);
  }
  __syncthreads();  // the tables

  C v[NV];
  {
    const int p0 = pl.r[0].p0;
    const uint32_t g0 = glob(thr(p0));
    uint32_t st[RB];
#pragma unroll
    for (int j = 0; j < RB; ++j) st[j] = gstride(p0 + j);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      uint32_t g = g0;
#pragma unroll
      for (int j = 0; j < RB; ++j)
        if (i & (1 << j)) g += st[j];
      // a 32-bit byte offset from a uniform base: one VGPR per address
      v[i] = (C)*reinterpret_cast<const TI *>(reinterpret_cast<const char *>(in) +
                                              g * (uint32_t)sizeof(TI));
    }
    if (flags & F_SIN) {
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] *= sin;
    }
    if (flags & F_PRO) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        uint32_t g = g0;
#pragma unroll
        for (int j = 0; j < RB; ++j)
          if (i & (1 << j)) g += st[j];
        v[i] *= dev::diag_value_k<DK, C>(dg, stab, g);
      }
    }
  }
  for (int k = 0; k < pl.nr; ++k) {
    if (k > 0) {
      const int pa = pl.r[k - 1].p0, pb = pl.r[k].p0;
      const int sa = dev::sw(thr(pa)), sb = dev::sw(thr(pb));
      if (k > 1) __syncthreads();  // the previous exchange's reads are done
      // sw is XOR-linear: sw(part | i << p) = sw(part) ^ sw(i << p)
#pragma unroll
      for (int i = 0; i < NV; ++i) lds[sa ^ dev::sw(i << pa)] = v[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NV; ++i) v[i] = lds[sb ^ dev::sw(i << pb)];
    }
    const int act = pl.r[k].act;
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      if (!(act & (1 << j))) continue;
      const int m = 1 << j;
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        if (!(i & m)) {
          const C x0 = v[i], x1 = v[i | m];
          v[i] = x0 + r * x1;
          v[i | m] = x1 + r * x0;
        }
      }
    }
  }
  // f = a^nb of K = a [[1, r], [r, 1]] over this pass's bits
  const C fs = f * sout;
  const int p0 = pl.r[pl.nr - 1].p0;
  const uint32_t g0 = glob(thr(p0));
  uint32_t st[RB];
#pragma unroll
  for (int j = 0; j < RB; ++j) st[j] = gstride(p0 + j);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    uint32_t g = g0;
#pragma unroll
    for (int j = 0; j < RB; ++j)
      if (i & (1 << j)) g += st[j];
    C x = v[i] * fs;
    if (flags & F_EPI) x *= dev::diag_value_k<DK, C>(dge, stabe, g);
    *reinterpret_cast<TO *>(reinterpret_cast<char *>(out) + g * (uint32_t)sizeof(TO)) = (TO)x;
  }
}

struct Pass { int b0, nb, logW; };

// Tiles of 64 KB (2^13 FP64, 2^14 FP32, 2^12 dual elements) plus the tables
// need the gfx950 LDS size; fall back to half-size tiles on a smaller LDS.
// Passes after the first move 128-byte contiguous runs (whole cache lines:
// with 64-byte runs the other half of each line is read by a tile that
// usually runs on another XCD, measured 5-8% slower for FP64).
int log2_es(int es) { return es == 4 ? 2 : (es == 8 ? 3 : 4); }

// default tile exponent, and the autotuner's candidates (default first)
int kron_tile_default(int es)
{
  const bool big = g_lds_per_block >= (size_t)(96 * 1024);
  return (big ? 16 : 15) - log2_es(es);
}
std::vector<int> kron_tile_candidates(int es)
{
  const int d = kron_tile_default(es);
  std::vector<int> c{d, d - 1};
  if (((size_t(1) << (d + 1)) + 384) * (size_t)es <= g_lds_per_block) {
    c.push_back(d + 1);
    c.push_back((d + 1) | KRON_WIDE);
  }
  return c;
}

// Passes over bits [b0, b0 + nbits): a contiguous first pass when b0 = 0,
// then passes of nb bits times W = 2^logW contiguous elements (>= 64 bytes
// where the lower bits allow it).
std::vector<Pass> plan_range(int b0, int nbits, int es, int tl)
{
  std::vector<Pass> ps;
  const int wl = 7 - log2_es(es);
  const int end = b0 + nbits;
  int b = b0;
  if (b0 == 0) {
    const int nb0 = nbits < tl ? nbits : tl;
    ps.push_back({0, nb0, 0});
    b = nb0;
  }
  while (b < end) {
    int nb = end - b < tl - wl ? end - b : tl - wl;
    int logW = tl - nb;
    if (logW > b) logW = b;
    ps.push_back({b, nb, logW});
    b += nb;
  }
  return ps;
}
std::vector<Pass> plan_passes(int L, int es, int tl) { return plan_range(0, L, es, tl); }

// Register-resident plan over all L bits: 2^TB threads (1024 for 4-byte,
// 512 for 8-byte elements) x 32 values = tiles of T = 15 / 14 bits (128 KB).
// The first pass is contiguous over [0, T) with rounds on bits 5-9, 0-4 and
// then the rest (its load and store layouts keep bits 0-4 on the lanes);
// later passes take the next bits with contiguous runs of the low bits
// (128 bytes, or 64 where that saves a pass), rounds from the top bits down.
// Empty if L < T.
int rpass_tb(int es) { return es == 4 ? 10 : 9; }
std::vector<RPlan> plan_rpass(int L, int es)
{
  const int T = rpass_tb(es) + 5;
  std::vector<RPlan> ps;
  if (L < T) return ps;
  RPlan f{};
  f.T = T;
  f.logW = 0;
  f.b0 = 0;
  f.nb = T;
  f.nr = 3;
  f.pair = 0;
  f.r[0] = {5, 31};
  f.r[1] = {0, 31};
  // bits [10, T) in the last round: register bits from T - 5 up
  f.r[2] = {T - 5, 31 & ~((1 << (10 - (T - 5))) - 1)};
  ps.push_back(f);
  const int l2 = es == 4 ? 2 : (es == 8 ? 3 : 4);
  const int w128 = 7 - l2, w64 = 6 - l2;
  int b = T;
  while (b < L) {
    const int rem = L - b;
    int nb, logW;
    if (rem <= T - w64) {
      nb = rem;
      logW = T - rem;
    } else {
      nb = T - w128;
      logW = w128;
    }
    RPlan q{};
    q.T = T;
    q.logW = logW;
    q.b0 = b;
    q.nb = nb;
    // rounds from the top: register bits [p0, p0 + 5) inside the tile, the
    // active ones being the pass bits in [max(logW, p0), hi)
    int hi = T, nr = 0;
    while (hi > logW) {
      const int p0 = std::min(T - 5, std::max(logW, hi - 5));
      const int lo = std::max(logW, p0);
      q.r[nr++] = {p0, ((1 << (hi - p0)) - 1) & ~((1 << (lo - p0)) - 1)};
      hi = lo;
    }
    q.nr = nr;
    q.pair = ((es << logW) < 128 && L - T >= 4) ? 1 : 0;
    ps.push_back(q);
    b += nb;
  }
  return ps;
}

// whether the register-resident plan can run an operator: a vector of at
// least one tile, 32-bit offsets, a diagonal kind it is instantiated for,
// and the 128 KB tile + tables in LDS
bool rpass_ok(const KronPlan &plan, int es)
{
  if (plan_rpass(plan.L, es).empty() || plan.L > 30 || ((uint64_t)es << plan.L) > (1ull << 32))
    return false;
  if (g_lds_per_block < ((size_t(1) << (rpass_tb(es) + 5)) + 192) * (size_t)es) return false;
  if (!plan.prologue && !plan.epilogue) return true;
  switch (plan.diag.kind) {
    case DIAG_X2D_PBC: case DIAG_X2D_OBC: case DIAG_X2D_C5: case DIAG_X2D_C9:
    case DIAG_X2D_C13: case DIAG_X2D_C17: case DIAG_X3D: return true;
    default: return false;
  }
}

template <typename C, typename TI, typename TO>
void run_rpass(const RPlan &p, int Ltot, const void *in, void *out, int flags, const DiagSpec &dg,
               const C *tab, const DiagSpec &dge, const C *tabe, double ka, double kb,
               const double *scale, hipStream_t s)
{
  const auto kc = dev::kron_coef<C>(ka, 0.0, kb, 0.0, p.nb);
  const uint64_t ntile = 1ull << (Ltot - p.T);
  dim3 grid;
  if (ntile <= 65535) grid = dim3((unsigned)ntile, 1, 1);
  else grid = dim3(65536, (unsigned)(ntile / 65536), 1);
  const size_t shm = ((size_t(1) << p.T) + ((flags & F_TWO) ? 384 : 192)) * sizeof(C);
  TM_CHECK(p.nr >= 1 && p.nr <= 3, TM_EINVAL, "register-resident pass plan out of range");
  TM_CHECK(((uint64_t)sizeof(C) << Ltot) <= (1ull << 32) && Ltot <= 30, TM_EINVAL,
           "register-resident Kronecker pass needs 32-bit element and byte offsets");
  TM_CHECK(!(flags & F_TWO) && dg.kind == dge.kind, TM_EINVAL,
           "register-resident Kronecker pass takes one diagonal kind");
  constexpr int TB = sizeof(C) == 4 ? 10 : 9;
  auto go = [&](auto kind) {
    constexpr int DK = decltype(kind)::value;
    hipLaunchKernelGGL((k_kron_rpass<C, TI, TO, TB, DK>), grid, dim3(1 << TB), shm, s,
                       static_cast<const TI *>(in), static_cast<TO *>(out), p, flags, dg, tab,
                       dge, tabe, kc.r, kc.f, scale);
  };
  const int kind = (flags & (F_PRO | F_EPI)) ? dg.kind : -1;
  switch (kind) {
    case -1: go(std::integral_constant<int, -1>()); break;
    case DIAG_X2D_PBC: go(std::integral_constant<int, DIAG_X2D_PBC>()); break;
    case DIAG_X2D_OBC: go(std::integral_constant<int, DIAG_X2D_OBC>()); break;
    case DIAG_X2D_C5: go(std::integral_constant<int, DIAG_X2D_C5>()); break;
    case DIAG_X2D_C9: go(std::integral_constant<int, DIAG_X2D_C9>()); break;
    case DIAG_X2D_C13: go(std::integral_constant<int, DIAG_X2D_C13>()); break;
    case DIAG_X2D_C17: go(std::integral_constant<int, DIAG_X2D_C17>()); break;
    case DIAG_X3D: go(std::integral_constant<int, DIAG_X3D>()); break;
    default: TM_CHECK(false, TM_EINVAL, "register-resident Kronecker pass: unsupported diagonal");
  }
  TM_LAUNCH_CHECK();
}


// One pass over an Ltot-bit index; K = ka [[1, r], [r, 1]]: the pass applies
// r in its butterflies and ka^nb on store (derivatives dka, dkb for dual
// arithmetic).  dg/tab: prologue diagonal, dge/tabe: epilogue diagonal.
template <typename C, typename TI, typename TO>
void run_pass(const Pass &p, int Ltot, const void *in, void *out, int flags, const DiagSpec &dg,
              const C *tab, const DiagSpec &dge, const C *tabe, double ka, double kb, double dka,
              double dkb, const double *scale, hipStream_t s, bool wide)
{
  const auto kc = dev::kron_coef<C>(ka, dka, kb, dkb, p.nb);
  const uint64_t ntile = 1ull << (Ltot - p.nb - p.logW);
  dim3 grid;
  if (ntile <= 65535) grid = dim3((unsigned)ntile, 1, 1);
  else grid = dim3(65536, (unsigned)(ntile / 65536), 1);
  const size_t shm =
      ((size_t(1) << (p.nb + p.logW)) + ((flags & F_TWO) ? 384 : 192)) * sizeof(C);
  if (wide)
    hipLaunchKernelGGL((k_kron_pass<C, TI, TO, 1024>), grid, dim3(1024), shm, s,
                       static_cast<const TI *>(in), static_cast<TO *>(out), p.b0, p.nb, p.logW,
                       flags, dg, tab, dge, tabe, kc.r, kc.f, scale);
  else
    hipLaunchKernelGGL((k_kron_pass<C, TI, TO, 512>), grid, dim3(512), shm, s,
                       static_cast<const TI *>(in), static_cast<TO *>(out), p.b0, p.nb, p.logW,
                       flags, dg, tab, dge, tabe, kc.r, kc.f, scale);
  TM_LAUNCH_CHECK();
}

template <typename C>
void kron_range_typed(const KronRange &kr, const void *in, void *out, void *mid, tm_dtype io_in,
                      tm_dtype io_out, const double *scale, bool sin, bool sout, hipStream_t s,
                      int cand)
{
  const bool wide = (cand & KRON_WIDE) != 0;
  auto ps = plan_range(kr.b0, kr.nbits, (int)sizeof(C), cand & 0xFF);
  const int np = (int)ps.size();
  const C *tp = static_cast<const C *>(kr.tpro), *te = static_cast<const C *>(kr.tepi);
  const bool two = kr.pro && kr.epi && (tp != te || kr.dpro.kind != kr.depi.kind);
  // one table: the kernel reads (dpro, tpro) for both ends
  if (!two && !kr.pro) tp = te;
  if (!two && kr.pro) te = tp;
  const DiagSpec &dsel = kr.pro ? kr.dpro : kr.depi;
  const int L = kr.Ltot;
  for (int i = 0; i < np; ++i) {
    int flags = two ? F_TWO : 0;
    if (i == 0 && kr.pro) flags |= F_PRO;
    if (i == np - 1 && kr.epi) flags |= F_EPI;
    if (i == 0 && scale && sin) flags |= F_SIN;
    if (i == np - 1 && scale && sout) flags |= F_SOUT;
    const void *src = (i == 0) ? in : mid;
    void *dst = (i == np - 1) ? out : mid;
    const bool src_io = (i == 0);
    const bool dst_io = (i == np - 1);
    const double *sc = (flags & (F_SIN | F_SOUT)) ? scale : nullptr;
    const DiagSpec &dp = two ? kr.dpro : dsel, &de = two ? kr.depi : dsel;
    const double ka = kr.ka, kb = kr.kb;
    if (src_io && dst_io) {
      if (io_in == TM_F64 && io_out == TM_F64) run_pass<C, double, double>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
      else if (io_in == TM_F64) run_pass<C, double, float>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
      else if (io_out == TM_F64) run_pass<C, float, double>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
      else run_pass<C, float, float>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
    } else if (src_io) {
      if (io_in == TM_F64) run_pass<C, double, C>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
      else run_pass<C, float, C>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
    } else if (dst_io) {
      if (io_out == TM_F64) run_pass<C, C, double>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
      else run_pass<C, C, float>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, sc, s, wide);
    } else {
      run_pass<C, C, C>(ps[i], L, src, dst, flags, dp, tp, de, te, ka, kb, 0.0, 0.0, nullptr, s, wide);
    }
  }
}

// The register-resident plan over all L bits (one diagonal table for both
// ends; the first pass applies the prologue and the input scale, the last
// the epilogue and 1/scale).
template <typename C>
void kron_rpass_typed(const KronPlan &plan, const void *in, void *out, void *mid, tm_dtype io_in,
                      tm_dtype io_out, const double *scale, hipStream_t s, const C *tab)
{
  const auto ps = plan_rpass(plan.L, (int)sizeof(C));
  const int np = (int)ps.size();
  TM_CHECK(np > 0, TM_EINVAL, "register-resident Kronecker plan needs L >= its tile bits");
  const int L = plan.L;
  for (int i = 0; i < np; ++i) {
    int flags = 0;
    if (i == 0 && plan.prologue) flags |= F_PRO;
    if (i == np - 1 && plan.epilogue) flags |= F_EPI;
    if (i == 0 && scale) flags |= F_SIN;
    if (i == np - 1 && scale) flags |= F_SOUT;
    const void *src = (i == 0) ? in : mid;
    void *dst = (i == np - 1) ? out : mid;
    const bool src_io = (i == 0), dst_io = (i == np - 1);
    const double *sc = (flags & (F_SIN | F_SOUT)) ? scale : nullptr;
    const DiagSpec &d = plan.diag;
    const double ka = plan.ka, kb = plan.kb;
    if (src_io && dst_io) {
      if (io_in == TM_F64 && io_out == TM_F64) run_rpass<C, double, double>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
      else if (io_in == TM_F64) run_rpass<C, double, float>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
      else if (io_out == TM_F64) run_rpass<C, float, double>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
      else run_rpass<C, float, float>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
    } else if (src_io) {
      if (io_in == TM_F64) run_rpass<C, double, C>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
      else run_rpass<C, float, C>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
    } else if (dst_io) {
      if (io_out == TM_F64) run_rpass<C, C, double>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
      else run_rpass<C, C, float>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, sc, s);
    } else {
      run_rpass<C, C, C>(ps[i], L, src, dst, flags, d, tab, d, tab, ka, kb, nullptr, s);
    }
  }
}

template <typename C>
void kron_typed(const KronPlan &plan, const void *in, void *out, tm_dtype io_in, tm_dtype io_out,
                void *work, const double *scale, hipStream_t s, const C *tab, int cand)
{
  // Intermediate buffer: the output itself when it has the compute type,
  // else the workspace.
  const bool out_is_c = (sizeof(C) == 8) == (io_out == TM_F64);
  if (cand & KRON_RPASS) {
    kron_rpass_typed<C>(plan, in, out, out_is_c ? out : work, io_in, io_out, scale, s, tab);
    return;
  }
  const KronRange kr{plan.L, 0, plan.L, plan.prologue, plan.epilogue, plan.diag, plan.diag,
                     tab, tab, plan.ka, plan.kb};
  kron_range_typed<C>(kr, in, out, out_is_c ? out : work, io_in, io_out, scale, true, true, s,
                      cand);
}

}  // namespace

size_t kron_work_bytes(int L, tm_dtype comp)
{
  return comp == TM_F32 ? (sizeof(float) << L) : 0;
}

void launch_kron_dual(const KronPlan &plan, const double *in, void *out, const void *tab_dual,
                      double dka, double dkb, hipStream_t s)
{
  using dev::Dual;
  const int L = plan.L;
  const Dual *tab = static_cast<const Dual *>(tab_dual);
  Dual *o = static_cast<Dual *>(out);
  auto run = [&](int cand) {
    const bool wide = (cand & KRON_WIDE) != 0;
    auto ps = plan_passes(L, (int)sizeof(Dual), cand & 0xFF);
    const int np = (int)ps.size();
    for (int i = 0; i < np; ++i) {
      int flags = 0;
      if (i == 0 && plan.prologue) flags |= F_PRO;
      if (i == np - 1 && plan.epilogue) flags |= F_EPI;
      if (i == 0)
        run_pass<Dual, double, Dual>(ps[i], L, in, o, flags, plan.diag, tab, plan.diag, tab,
                                     plan.ka, plan.kb, dka, dkb, nullptr, s, wide);
      else
        run_pass<Dual, Dual, Dual>(ps[i], L, o, o, flags, plan.diag, tab, plan.diag, tab,
                                   plan.ka, plan.kb, dka, dkb, nullptr, s, wide);
    }
  };
  const int es = (int)sizeof(Dual);
  run(autotune_tile(TUNE_KRON_DUAL, 0, L, es, kron_tile_candidates(es), plan.tune && in != out, s,
                  run));
}

void launch_kron(const KronPlan &plan, const void *in, void *out, tm_dtype io_in, tm_dtype io_out,
                 tm_dtype comp, void *work, const double *scale, hipStream_t s)
{
  const int es = comp == TM_F64 ? 8 : 4;
  auto run = [&](int cand) {
    if (comp == TM_F64)
      kron_typed<double>(plan, in, out, io_in, io_out, work, scale, s, plan.tab64, cand);
    else
      kron_typed<float>(plan, in, out, io_in, io_out, work, scale, s, plan.tab32, cand);
  };
  // the register-resident plan first (the default) where the vector holds
  // at least one of its tiles; timing candidates rewrites out from in, so
  // only out-of-place applies tune
  std::vector<int> cands = kron_tile_candidates(es);
  if (rpass_ok(plan, es)) cands.insert(cands.begin(), KRON_RPASS);
  run(autotune_tile(TUNE_KRON, 0, plan.L, es, cands, plan.tune && in != out, s, run));
}

void launch_kron_range(const KronRange &kr, const void *in, void *out, void *mid,
                       tm_dtype io_in, tm_dtype io_out, tm_dtype comp, const double *scale,
                       bool sin, bool sout, bool tune, hipStream_t s)
{
  const int es = comp == TM_F64 ? 8 : 4;
  auto run = [&](int cand) {
    if (comp == TM_F64)
      kron_range_typed<double>(kr, in, out, mid, io_in, io_out, scale, sin, sout, s, cand);
    else
      kron_range_typed<float>(kr, in, out, mid, io_in, io_out, scale, sin, sout, s, cand);
  };
  // a candidate run must leave its input intact to be repeated
  const bool safe = in != out && in != mid;
  const uint32_t id = (uint32_t)kr.b0 * 64u + (uint32_t)kr.nbits;
  run(autotune_tile(TUNE_KRON_RANGE, id, kr.Ltot, es, kron_tile_candidates(es), tune && safe, s,
                    run));
}

}  // namespace ltm
