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

template <typename C>
void kron_typed(const KronPlan &plan, const void *in, void *out, tm_dtype io_in, tm_dtype io_out,
                void *work, const double *scale, hipStream_t s, const C *tab, int cand)
{
  // Intermediate buffer: the output itself when it has the compute type,
  // else the workspace.
  const bool out_is_c = (sizeof(C) == 8) == (io_out == TM_F64);
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
  // timing candidates rewrites out from in, so only out-of-place applies tune
  run(autotune_tile(TUNE_KRON, 0, plan.L, es, kron_tile_candidates(es), plan.tune && in != out, s,
                  run));
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
