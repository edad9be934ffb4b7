// tm_sweep.hip -- auxiliary-bit sweep operators on gfx950: DNNI (XDT) and
// BNNNI/3NN diagonal TM (ZBT and its transpose).
//
// The reference applies L (XDT, TMat/tmat.cpp:344-476) or L/2 (ZBT,
// TMat/tmat.cpp:521-649) "permute_mult_T1" steps, each of which rotates the
// whole state index and mixes the lowest bit(s) with 1 (2) auxiliary bits.
// Written in physical bit positions the rotation disappears: step t is a
// local gate on (aux, bit t) [XDT] or (aux pair, bits 2t,2t+1) [ZBT] whose
// coefficients depend on the next bit(s) of the state (or, for the last step,
// on the "lastbit" value the reference loops over).  The lastbit loop becomes
// a batch dimension, so one sweep over the array
//     Z[lastbit][aux][P],   P = 2^L physical states
// does all of it.  Gates are grouped into passes: a workgroup holds a tile of
// Z spanning (lastbit, aux, the pass's state bits, W contiguous low states)
// in LDS, applies its gates in order, and writes the tile back.  The first
// pass builds Z from x on the fly (Tv^1/2 x placed at the right aux/lastbit)
// and the last pass contracts over (lastbit, aux) and applies Tv^1/2.
#include <type_traits>

#include "tm_device.hpp"

namespace ltm {
namespace {

// Workgroup size: 512 threads, or 1024 (one 128 KB tile per CU then still
// keeps 16 wavefronts resident) -- an autotuner choice for FP64/FP32; a tile
// candidate encodes the tile exponent in bits 0-7, the 1024-thread choice in
// bit 8 and "keep lastbit in the first pass's tile" in bit 9 (the split first
// pass writes 16-byte pieces of each state's aux values, which some widths
// and precisions pay for more than the pass it saves).
constexpr int SWEEP_WIDE = 256;
constexpr int SWEEP_NOSPLIT = 512;

struct SweepArgs {
  int L;
  int model;        // 0 = XDT, 1 = ZBT
  int dir;          // ZBT: 0 = T, 1 = T^T
  int bc;           // XDT: 0 periodic (lastbit batch), 5 / 9 confined
  int nlb, na;      // lastbit values, aux values (log2 in nlbl, nal)
  int nlbl, nal;
  int t0, ng;       // first gate, number of gates in this pass
  int bitlo;        // first state bit owned by the pass
  int nbits;        // state bits owned by the pass
  int logW;
  int first, last;
  int lbsplit;      // the lastbit values in separate workgroups (a first pass
                    // that is not the last: no gate of it reads lastbit)
  int64_t nblk;     // lbsplit: workgroups (tiles x lastbit values) before padding
  DiagSpec dg;
  double b1[2], b2[2], b3[2];
  double outfac;    // XDT: normfactor; ZBT: 1 (folded into the diagonal)
  double db1[2], db2[2], db3[2], doutfac;  // derivatives (dual-number kernels only)
  // ZBT gate coefficients of the FP64 and FP32 kernels, computed on the
  // host (wave-uniform kernel arguments): Q[p][s], kap[s] and rho2 of the
  // ratio form (zbt_gate); the dual-number kernels build theirs on the device
  double zQ[16], zkap[4], zr2;
  float fQ[16], fkap[4], fr2, fb1[2], fb2[2];
};

__device__ __forceinline__ uint32_t ins0(uint32_t q, int pos)
{
  return ((q >> pos) << (pos + 1)) | (q & ((1u << pos) - 1));
}

template <typename C>
struct Coef {
  C b1[2], b2[2], b3[2];
};

// b[i] for a run-time bit i as a select (a dynamically indexed private
// array would be placed in scratch memory)
template <typename C>
__device__ __forceinline__ C pick(const C (&b)[2], int i)
{
  return i ? b[1] : b[0];
}

// XDT gate: in[a][u] (a = aux, u = state bit t), control c.
//   out[p][s] = b1[s^p] b2[c^s] (b2[s] in[0][p] + b2[s^1] in[1][p])
template <typename C>
__device__ __forceinline__ void xdt_gate(C v[2][2], int c, const Coef<C> &k)
{
  C o[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int s = 0; s < 2; ++s)
      o[p][s] = k.b1[s ^ p] * pick(k.b2, c ^ s) * (k.b2[s] * v[0][p] + k.b2[s ^ 1] * v[1][p]);
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int s = 0; s < 2; ++s) v[p][s] = o[p][s];
}

// ZBT gate: in[aq][u] (aq = aux pair (al,ar), u = (bit 2t+1, bit 2t)),
// controls (fl, fr).  Output aux p = (pl,pr) takes the old bit pair, the
// new bit pair is s = (sl, sr):
//   out[p][s] = A[p][s] F_{fl,fr}[s] sum_aq Cm[s][aq] in[aq][p]
//   T   (DIR 0): A = b1[sr^pl] b3[sr^pr] b3[sl^pl], Cm = b1[sr^al] b2[sr^ar] b2[sl^al],
//                F = b2[sr^fr] b2[sl^fl]
//   T^T (DIR 1): A = b1[sl^pr] b3[sr^pr] b3[sl^pl], Cm = b2[sr^ar] b2[sl^al],
//                F = b1[sl^fr] b2[sr^fr] b2[sl^fl]
// The sum over aq factors into two 2x2 stages in ratio form (b[x] = b[0]
// rho^x): w[al][sr] = in[al,sr] + rho2 in[al,1-sr] over ar, then over al
//   DIR 0: acc = w[0][sr] + kap[s] w[1][sr],  kap = rho1^(1-2sr) rho2^(1-2sl)
//   DIR 1: acc = w[sl][sr] + rho2 w[1-sl][sr]
// one FMA per value each, and out = (Q[p][s] F[s]) acc with the
// control-independent Q (A and the pulled-out factors) tabulated once per
// thread.
template <typename C>
__device__ __forceinline__ C cdiv(C a, C b)
{
  return a / b;
}
template <>
__device__ __forceinline__ dev::Dual cdiv<dev::Dual>(dev::Dual a, dev::Dual b)
{
  const double q = a.v / b.v;
  return dev::Dual(q, (a.d - q * b.d) / b.v);
}

template <typename C>
struct ZCoef {
  C Q[4][4], kap[4];
  C r2;
  C b1[2], b2[2];
};

template <typename C, int DIR>
__device__ __forceinline__ ZCoef<C> make_zcoef(const Coef<C> &k)
{
  ZCoef<C> z;
  const C r1 = cdiv(k.b1[1], k.b1[0]), r2 = cdiv(k.b2[1], k.b2[0]);
  const C ir1 = cdiv(C(1.0), r1), ir2 = cdiv(C(1.0), r2);
  const C pre = k.b1[0] * k.b2[0] * k.b2[0];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int sl = s >> 1, sr = s & 1;
    z.kap[s] = (sr ? ir1 : r1) * (sl ? ir2 : r2);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int pl = p >> 1, pr = p & 1;
      if (DIR == 0)
        z.Q[p][s] = k.b1[sr ^ pl] * k.b3[sr ^ pr] * k.b3[sl ^ pl] * pre * (sr ? r1 : C(1.0)) *
                    (sl ? r2 : C(1.0));
      else
        z.Q[p][s] = k.b1[sl ^ pr] * k.b3[sr ^ pr] * k.b3[sl ^ pl] * (k.b2[0] * k.b2[0]);
    }
  }
  z.r2 = r2;
  for (int i = 0; i < 2; ++i) {
    z.b1[i] = k.b1[i];
    z.b2[i] = k.b2[i];
  }
  return z;
}

template <typename C, int DIR>
__device__ __forceinline__ void zbt_gate(C v[4][4], int fl, int fr, const ZCoef<C> &z)
{
  // F[s] for this item's controls
  const C ur0 = fr ? z.b2[1] : z.b2[0], ur1 = fr ? z.b2[0] : z.b2[1];  // b2[sr ^ fr]
  const C ul0 = fl ? z.b2[1] : z.b2[0], ul1 = fl ? z.b2[0] : z.b2[1];  // b2[sl ^ fl]
  C F[4];
  F[0] = ul0 * ur0;
  F[1] = ul0 * ur1;
  F[2] = ul1 * ur0;
  F[3] = ul1 * ur1;
  if (DIR == 1) {
    const C g0 = fr ? z.b1[1] : z.b1[0], g1 = fr ? z.b1[0] : z.b1[1];  // b1[sl ^ fr]
    F[0] *= g0;
    F[1] *= g0;
    F[2] *= g1;
    F[3] *= g1;
  }
  C w[2][2][4];
#pragma unroll
  for (int al = 0; al < 2; ++al)
#pragma unroll
    for (int sr = 0; sr < 2; ++sr)
#pragma unroll
      for (int p = 0; p < 4; ++p) w[al][sr][p] = v[al * 2 + sr][p] + z.r2 * v[al * 2 + (sr ^ 1)][p];
  C o[4][4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int sl = s >> 1, sr = s & 1;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const C acc = DIR == 0 ? w[0][sr][p] + z.kap[s] * w[1][sr][p]
                             : w[sl][sr][p] + z.r2 * w[sl ^ 1][sr][p];
      o[p][s] = (z.Q[p][s] * F[s]) * acc;
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int s = 0; s < 4; ++s) v[p][s] = o[p][s];
}

// One column of a ZBT gate, in place and transposed: the four inputs
// in[aq][p] (aq = (al, ar)) of a fixed old bit pair p become the outputs
// out[p][s] (s = (sl, sr)), written back to the same four slots in s order
// (slot aq holds s = aq afterwards).  qf[s] = Q[p][s] F[s].
template <typename C, int DIR>
__device__ __forceinline__ void zbt_col(C &v0, C &v1, C &v2, C &v3, C q0, C q1, C q2, C q3,
                                        C r2, C k0, C k1, C k2, C k3)
{
  const C w00 = v0 + r2 * v1, w01 = v1 + r2 * v0;  // w[al = 0][sr]
  const C w10 = v2 + r2 * v3, w11 = v3 + r2 * v2;  // w[al = 1][sr]
  if (DIR == 0) {
    v0 = q0 * (w00 + k0 * w10);
    v1 = q1 * (w01 + k1 * w11);
    v2 = q2 * (w00 + k2 * w10);
    v3 = q3 * (w01 + k3 * w11);
  } else {
    v0 = q0 * (w00 + r2 * w10);
    v1 = q1 * (w01 + r2 * w11);
    v2 = q2 * (w10 + r2 * w00);
    v3 = q3 * (w11 + r2 * w01);
  }
}

template <typename C>
struct ZArg;
template <>
struct ZArg<double> {
  static __device__ __forceinline__ double Q(const SweepArgs &a, int i) { return a.zQ[i]; }
  static __device__ __forceinline__ double kap(const SweepArgs &a, int i) { return a.zkap[i]; }
  static __device__ __forceinline__ double r2(const SweepArgs &a) { return a.zr2; }
  static __device__ __forceinline__ double b1(const SweepArgs &a, int i)
  {
    return i ? a.b1[1] : a.b1[0];
  }
  static __device__ __forceinline__ double b2(const SweepArgs &a, int i)
  {
    return i ? a.b2[1] : a.b2[0];
  }
};
template <>
struct ZArg<float> {
  static __device__ __forceinline__ float Q(const SweepArgs &a, int i) { return a.fQ[i]; }
  static __device__ __forceinline__ float kap(const SweepArgs &a, int i) { return a.fkap[i]; }
  static __device__ __forceinline__ float r2(const SweepArgs &a) { return a.fr2; }
  static __device__ __forceinline__ float b1(const SweepArgs &a, int i)
  {
    return i ? a.fb1[1] : a.fb1[0];
  }
  static __device__ __forceinline__ float b2(const SweepArgs &a, int i)
  {
    return i ? a.fb2[1] : a.fb2[0];
  }
};

// F[s] = b2[sr ^ fr] b2[sl ^ fl] (times b1[sl ^ fr] for T^T) for run-time
// controls (fl, fr)
template <typename C, int DIR>
__device__ __forceinline__ void zbt_F(C (&F)[4], int fl, int fr, const SweepArgs &a)
{
  using Z = ZArg<C>;
  const C ur0 = Z::b2(a, fr), ur1 = Z::b2(a, fr ^ 1);
  const C ul0 = Z::b2(a, fl), ul1 = Z::b2(a, fl ^ 1);
  F[0] = ul0 * ur0;
  F[1] = ul0 * ur1;
  F[2] = ul1 * ur0;
  F[3] = ul1 * ur1;
  if (DIR == 1) {
    const C g0 = Z::b1(a, fr), g1 = Z::b1(a, fr ^ 1);
    F[0] *= g0;
    F[1] *= g0;
    F[2] *= g1;
    F[3] *= g1;
  }
}

// One ZBT gate on a 16-value item v[aux][bits] with run-time controls; on
// return v[i][j] holds bits = i, aux = j.
template <typename C, int DIR>
__device__ __forceinline__ void zbt_gate_single(C (&v)[4][4], int fl, int fr, const SweepArgs &a)
{
  using Z = ZArg<C>;
  C F[4];
  zbt_F<C, DIR>(F, fl, fr, a);
  const C r2 = Z::r2(a);
  const C k0 = Z::kap(a, 0), k1 = Z::kap(a, 1), k2 = Z::kap(a, 2), k3 = Z::kap(a, 3);
#pragma unroll
  for (int p = 0; p < 4; ++p)
    zbt_col<C, DIR>(v[0][p], v[1][p], v[2][p], v[3][p], Z::Q(a, p * 4 + 0) * F[0],
                    Z::Q(a, p * 4 + 1) * F[1], Z::Q(a, p * 4 + 2) * F[2],
                    Z::Q(a, p * 4 + 3) * F[3], r2, k0, k1, k2, k3);
}

// LDS bank swizzle of the tile: position e is stored at e ^ f(e), with f
// GF(2)-linear from the tile bits [lo, lo + 12) into bits [0, 5): column j
// is the image of bit lo + j (lo = 4 for 8- and 16-byte elements, 5 for
// 4-byte ones; the map is triangular, so e -> e ^ f(e) is a bijection that
// keeps every bit from 5 up).  The columns were chosen so that the access
// patterns of the passes -- lanes on consecutive states (load, store,
// contraction) and lanes on consecutive gate items of a gate at any even
// tile position -- reach 16 (8-byte) / 32 (4-byte) distinct banks per lane
// group on tiles of 2^11 and more elements.  f is applied through two
// 64-entry tables over its 6-bit input chunks, built in LDS per workgroup.
constexpr int SWZ_TAB = 128;
constexpr uint32_t SWZ_COLS_8[12] = {0x6, 0xb, 0x13, 0x2, 0x12, 0x18, 0x12, 0x8, 0x4, 0x1a, 0xa, 0x5};
constexpr uint32_t SWZ_COLS_4[12] = {0x6, 0x1b, 0x4, 0x15, 0x3, 0x15, 0x16, 0x1f, 0x15, 0x7, 0x0, 0x0};

template <int MODEL, int DIR, typename C, typename TI, typename TO, int NT>
// zin and zout are the same aux array (the passes work in place): each
// workgroup reads its tile before the barrier and writes it back after it,
// so they are not __restrict__.
__global__ __launch_bounds__(NT) void k_sweep_pass(const TI *__restrict__ x,
                                                               const C *zin,
                                                               C *zout,
                                                               TO *__restrict__ y, SweepArgs a,
                                                               const C *__restrict__ tab,
                                                               const double *__restrict__ scale)
{
  extern __shared__ __align__(16) unsigned char smem[];
  // lastbit bits held in the tile (none when split over workgroups)
  const int nlbt = a.lbsplit ? 0 : a.nlbl;
  const int tileLog = nlbt + a.nal + a.nbits + a.logW;
  const int T = 1 << tileLog;
  C *lds = reinterpret_cast<C *>(smem);
  char *ldsb = reinterpret_cast<char *>(smem);
  C *stab = lds + T;
  // split: workgroup -> (tile, lastbit) with the lastbit values of a tile on
  // one XCD (their interleaved writes to the state-major aux array then meet
  // in the same L2)
  uint64_t tile = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
  int wlb = 0;
  if (a.lbsplit) {
    if (tile >= (uint64_t)((a.nblk + 7) / 8 * 8)) return;  // grid padding
    const int64_t q = dev::xcd_remap((int64_t)tile, a.nblk);
    if (q < 0) return;
    wlb = (int)(q & (a.nlb - 1));
    tile = (uint64_t)q >> a.nlbl;
  }
  if (a.first || a.last)
    for (int i = threadIdx.x; i < 192; i += blockDim.x) stab[i] = tab[i];
  int *swzt = reinterpret_cast<int *>(stab + 192);
  constexpr int slo = sizeof(C) == 4 ? 5 : 4;
  for (int i = threadIdx.x; i < SWZ_TAB; i += blockDim.x) {
    const int chunk = i >> 6, idx = i & 63;
    int v = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b)
      if ((idx >> b) & 1) v ^= (int)(sizeof(C) == 4 ? SWZ_COLS_4 : SWZ_COLS_8)[chunk * 6 + b];
    swzt[i] = v;
  }
  Coef<C> k;
  for (int i = 0; i < 2; ++i) {
    k.b1[i] = dev::mk<C>(a.b1[i], a.db1[i]);
    k.b2[i] = dev::mk<C>(a.b2[i], a.db2[i]);
    k.b3[i] = dev::mk<C>(a.b3[i], a.db3[i]);
  }
  // dual-number kernels: gate coefficients on the device (FP64 and FP32
  // kernels take theirs from the kernel arguments)
  ZCoef<C> zk{};
  if constexpr (sizeof(C) > 8) zk = make_zcoef<C, DIR>(k);
  const int L = a.L;
  const uint64_t innerBlocks = (1ull << a.bitlo) >> a.logW;
  const uint64_t o = tile / innerBlocks, ib = tile % innerBlocks;
  const uint64_t base = (o << (a.bitlo + a.nbits)) | (ib << a.logW);
  const int W = 1 << a.logW;
  const int posAux = a.logW + a.nbits;
  const int posLb = posAux + a.nal;
  // The aux array is state-major in HBM: Z[(P << lgz) | z], z = (lastbit,
  // aux), so the 2^lgz values of a state are contiguous and a tile's runs
  // are W * 2^lgz elements long.  In the tile, (aux, lastbit) are the high
  // bits (from posAux) of the position, stored swizzled (swz).
  const int lgz = a.nlbl + a.nal;
  const int lgt = nlbt + a.nal;  // aux (and lastbit) bits in the tile
  const int nz = 1 << lgt;
  const int zoff = wlb << a.nal;  // this workgroup's lastbit in the HBM z index
  auto swz = [&](int e) __attribute__((always_inline)) {
    return e ^ swzt[(e >> slo) & 63] ^ swzt[64 + ((e >> (slo + 6)) & 63)];
  };
  C sin = 1, sout = 1;
  if (scale) {
    if (a.first) sin = (C)(*scale);
    if (a.last) sout = (C)(1.0 / *scale);
  }
  __syncthreads();

  // ---- load / build the tile ----
  // One state per thread and step: its nz values are contiguous in HBM
  // (Z[(P << lgz) | z]) and sit at swz(es) ^ szt[z] in LDS (es = the state's
  // tile position; swz is linear, szt[z] = swz(z << posAux)).  A lane group
  // walks consecutive states.  The first pass writes the single nonzero
  // value of a state (at z = (lastbit, aux) of its own bits) and zeros.
  const int NS = T >> lgt;
  int szt[16];
#pragma unroll
  for (int z = 0; z < 16; ++z) szt[z] = z < nz ? swz(z << posAux) : 0;
  auto state_P = [&](int es) __attribute__((always_inline)) {
    return base | ((uint64_t)(es >> a.logW) << a.bitlo) | (uint64_t)(es & (W - 1));
  };
  // Passes after the first move each state's nz values as CB-byte chunks on
  // consecutive lanes (LAM lanes per state), so one load or store
  // instruction of a wave covers whole contiguous 128-byte lines instead of
  // one 16-byte piece of 64 different lines.
  auto fill = [&](auto nzc) __attribute__((always_inline)) {
    constexpr int NZ = decltype(nzc)::value;
    if (a.first) {
#pragma unroll 2
      for (int es = threadIdx.x; es < NS; es += blockDim.x) {
        const uint64_t P = state_P(es);
        const int ses = swz(es);
        int zi = -1;
        C val = 0;
        if (MODEL == 0 && a.bc != 0) {
          // confined ends, TMat/tmat.cpp:440-458
          const int top = (int)(P >> (L - 1));
          const int low = (int)(P & 1) ^ (a.bc == 5 ? 0 : 1);
          zi = a.bc == 5 ? 0 : 1;
          val = (C)x[P] * sin * dev::diag_value<C>(a.dg, stab, P) * pick(k.b2, low) *
                pick(k.b2, top);
        } else {
          // periodic: lastbit = the low bit(s), aux = the top bit(s)
          const int lb = MODEL == 0 ? (int)(P & 1) : (int)(P & 3);
          const int aux = MODEL == 0 ? (int)((P >> (L - 1)) & 1) : (int)((P >> (L - 2)) & 3);
          if (nlbt == 0) zi = lb == wlb ? aux : -1;
          else zi = (lb << a.nal) | aux;
          if (zi >= 0) val = (C)x[P] * sin * dev::diag_value<C>(a.dg, stab, P);
        }
#pragma unroll
        for (int z = 0; z < NZ; ++z) lds[ses ^ szt[z]] = z == zi ? val : C(0);
      }
    } else {
      constexpr int BS = NZ * (int)sizeof(C);
      constexpr int CB = BS < 16 ? BS : 16;
      constexpr int LAM = BS / CB, VPC = CB / (int)sizeof(C);
#pragma unroll 4
      for (int i = threadIdx.x; i < NS * LAM; i += blockDim.x) {
        const int es = i / LAM, j = i % LAM;
        const uint64_t P = state_P(es);
        const C *src = static_cast<const C *>(__builtin_assume_aligned(
            zin + ((P << lgz) | (uint64_t)zoff) + j * VPC, CB));
        C v[VPC];
#pragma unroll
        for (int u = 0; u < VPC; ++u) v[u] = src[u];
        const int sa = swz(es) ^ swz((j * VPC) << posAux);
#pragma unroll
        for (int u = 0; u < VPC; ++u) lds[sa ^ szt[u]] = v[u];
      }
    }
  };
  if (MODEL == 1 && nz == 16) fill(std::integral_constant<int, 16>{});
  else if (nz == 4) fill(std::integral_constant<int, 4>{});
  else fill(std::integral_constant<int, 2>{});
  __syncthreads();

  // ---- gates ----
  for (int g = 0; g < a.ng; ++g) {
    const int t = a.t0 + g;
    // swz is linear over XOR and an item's base b0 has zeros where the gate's
    // offsets have their bits: each LDS address is swz(b0) ^ swz(offset),
    // with the offsets' swizzles computed once per gate
    if (MODEL == 0) {
      const int jb = a.logW + (t - a.bitlo);  // tile position of state bit t
      const int nitems = T >> 2;
      int offB[2][2];  // byte offsets of the item's values
#pragma unroll
      for (int ax = 0; ax < 2; ++ax)
#pragma unroll
        for (int u = 0; u < 2; ++u) offB[ax][u] = swz((ax << posAux) | (u << jb)) * (int)sizeof(C);
      // the control (state bit t + 1): a tile bit of the item while the
      // next gate is in this pass, else a bit of the workgroup's base (or
      // lastbit / the confined end for the last gate)
      const bool cin = g + 1 < a.ng;
      const int cbase = t + 1 < L ? (int)((base >> (t + 1)) & 1) : 0;
      for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
        // insert zeros at the two varying positions (ascending)
        const uint32_t b0 = ins0(ins0((uint32_t)q, jb), posAux);
        int c;
        if (cin) c = (int)((b0 >> (jb + 1)) & 1);
        else if (t + 1 < L) c = cbase;
        else c = (a.bc == 0) ? (a.lbsplit ? wlb : (int)(b0 >> posLb)) : 0;
        int sbB = swz((int)b0) * (int)sizeof(C);
        // keep the byte scaling on the item base (one XOR per access; the
        // compiler would otherwise scale each access after the XOR)
        asm volatile("" : "+v"(sbB));
        C v[2][2];
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
          for (int u = 0; u < 2; ++u) v[ax][u] = *reinterpret_cast<C *>(ldsb + (sbB ^ offB[ax][u]));
        xdt_gate(v, c, k);
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
          for (int u = 0; u < 2; ++u) *reinterpret_cast<C *>(ldsb + (sbB ^ offB[ax][u])) = v[ax][u];
      }
    } else {
      const int jb = a.logW + (2 * t - a.bitlo);  // tile position of bit 2t
      const int nitems = T >> 4;
      int offB[4][4];  // byte offsets of the item's values
#pragma unroll
      for (int ax = 0; ax < 4; ++ax)
#pragma unroll
        for (int u = 0; u < 4; ++u) offB[ax][u] = swz((ax << posAux) | (u << jb)) * (int)sizeof(C);
      // the controls (state bits 2t+2, 2t+3): tile bits of the item while
      // the next gate is in this pass, else bits of the workgroup's base (or
      // lastbit for the last gate)
      const bool cin = g + 1 < a.ng;
      const int frb = 2 * t + 2 < L ? (int)((base >> (2 * t + 2)) & 1) : 0;
      const int flb = 2 * t + 2 < L ? (int)((base >> (2 * t + 3)) & 1) : 0;
      for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
        uint32_t b0 = ins0((uint32_t)q, jb);
        b0 = ins0(b0, jb + 1);
        b0 = ins0(b0, posAux);
        b0 = ins0(b0, posAux + 1);
        int fl, fr;
        if (cin) {
          fr = (int)((b0 >> (jb + 2)) & 1);
          fl = (int)((b0 >> (jb + 3)) & 1);
        } else if (2 * t + 2 < L) {
          fr = frb;
          fl = flb;
        } else {
          const int lb = a.lbsplit ? wlb : (int)(b0 >> posLb);
          fl = lb >> 1;
          fr = lb & 1;
        }
        int sbB = swz((int)b0) * (int)sizeof(C);
        asm volatile("" : "+v"(sbB));  // one XOR per access (see the XDT loop)
        auto at = [&](int ax, int u) -> C & {
          return *reinterpret_cast<C *>(ldsb + (sbB ^ offB[ax][u]));
        };
        C v[4][4];
#pragma unroll
        for (int ax = 0; ax < 4; ++ax)
#pragma unroll
          for (int u = 0; u < 4; ++u) v[ax][u] = at(ax, u);
        if constexpr (sizeof(C) > 8) {
          zbt_gate<C, DIR>(v, fl, fr, zk);
#pragma unroll
          for (int ax = 0; ax < 4; ++ax)
#pragma unroll
            for (int u = 0; u < 4; ++u) at(ax, u) = v[ax][u];
        } else {
          zbt_gate_single<C, DIR>(v, fl, fr, a);
          // v[i][j]: bits = i, aux = j
#pragma unroll
          for (int ax = 0; ax < 4; ++ax)
#pragma unroll
            for (int u = 0; u < 4; ++u) at(ax, u) = v[u][ax];
        }
      }
    }
    __syncthreads();
  }

  // ---- store ----
  if (!a.last) {
    auto drain = [&](auto nzc) __attribute__((always_inline)) {
      constexpr int NZ = decltype(nzc)::value;
      constexpr int BS = NZ * (int)sizeof(C);
      constexpr int CB = BS < 16 ? BS : 16;
      constexpr int LAM = BS / CB, VPC = CB / (int)sizeof(C);
#pragma unroll 4
      for (int i = threadIdx.x; i < NS * LAM; i += blockDim.x) {
        const int es = i / LAM, j = i % LAM;
        const uint64_t P = state_P(es);
        const int sa = swz(es) ^ swz((j * VPC) << posAux);
        C *dst = static_cast<C *>(
            __builtin_assume_aligned(zout + ((P << lgz) | (uint64_t)zoff) + j * VPC, CB));
#pragma unroll
        for (int u = 0; u < VPC; ++u) dst[u] = lds[sa ^ szt[u]];
      }
    };
    if (MODEL == 1 && nz == 16) drain(std::integral_constant<int, 16>{});
    else if (nz == 4) drain(std::integral_constant<int, 4>{});
    else drain(std::integral_constant<int, 2>{});
  } else {
    const C fac = dev::mk<C>(a.outfac, a.doutfac) * sout;
    for (int es = threadIdx.x; es < NS; es += blockDim.x) {
      const uint64_t P = state_P(es);
      const int ses = swz(es);
      C acc = 0;
#pragma unroll
      for (int z = 0; z < 16; ++z)
        if (z < nz) acc += lds[ses ^ szt[z]];
      y[P] = (TO)(fac * dev::diag_value<C>(a.dg, stab, P) * acc);
    }
  }
}

struct SPass { int t0, ng, bitlo, nbits, logW, split; };

// Tiles of 128 KB (2^14 FP64, 2^15 FP32, 2^13 dual elements) on the gfx950
// LDS, half that on a smaller one.  Passes after the first move runs of W
// consecutive states times their contiguous 2^(lastbit + aux bits) values,
// at least 128 bytes.  A first pass that is not the last leaves lastbit to
// the grid (no gate before the last reads it), which frees its bits for one
// more gate: ZBT at L = 22 then takes two passes instead of three.
int sweep_log_es(int es) { return es == 4 ? 2 : (es == 8 ? 3 : 4); }
// default tile exponent first, then the autotuner's other candidates.  At
// small widths the default tile leaves too few workgroups to fill the chip
// (2^(L + lgz) values in tiles of 2^d), so smaller tiles are offered down
// to about 2048 workgroups.
std::vector<int> sweep_tile_candidates(int es, int L, int lgz)
{
  const bool big = g_lds_per_block >= (size_t)(136 * 1024);
  const int d = (big ? 17 : 16) - sweep_log_es(es);
  std::vector<int> c;
  if (es <= 8)
    c = {d, d - 1, d | SWEEP_WIDE, d | SWEEP_NOSPLIT, (d - 1) | SWEEP_NOSPLIT,
         d | SWEEP_WIDE | SWEEP_NOSPLIT};
  else
    c = {d, d - 1, d | SWEEP_NOSPLIT};
  for (int t = d - 2; t >= 8 && L + lgz - t <= 11; --t) c.push_back(t);
  return c;
}

std::vector<SPass> plan_sweep(int model, int L, int bc, int es, int tl, bool split_ok)
{
  const int les = sweep_log_es(es);
  const int fixed = model == 0 ? (bc == 0 ? 2 : 1) : 4;  // log2(nlb * na)
  const int fixa = model == 0 ? 1 : 2;                     // log2(na)
  const int wl = 7 - les - fixed > 0 ? 7 - les - fixed : 0;
  const int bpg = model == 0 ? 1 : 2;                      // state bits per gate
  const int ngates = model == 0 ? L : L / 2;
  std::vector<SPass> ps;
  int t = 0;
  if (split_ok && fixed > fixa && ngates * bpg + fixed > tl) {
    // first pass with lastbit in the grid; at least one gate left over
    int ng = (tl - fixa) / bpg;
    if (ng > ngates - 1) ng = ngates - 1;
    ps.push_back({0, ng, 0, ng * bpg, 0, 1});
    t = ng;
  }
  while (t < ngates) {
    const int bitlo = t * bpg;
    int budget = tl - fixed;
    int logW = 0;
    if (bitlo > 0) {
      logW = wl < bitlo ? wl : bitlo;
      budget -= logW;
    }
    int ng = budget / bpg;
    if (ng > ngates - t) ng = ngates - t;
    if (ng < 1) ng = 1;
    const int nbits = ng * bpg;
    // widen W with the leftover budget on the last pass
    if (bitlo > 0) {
      int extra = tl - fixed - nbits - logW;
      while (extra > 0 && logW < bitlo) { ++logW; --extra; }
    }
    ps.push_back({t, ng, bitlo, nbits, logW, 0});
    t += ng;
  }
  return ps;
}

template <int MODEL, int DIR, typename C, typename TI, typename TO>
void run_sweep_pass(const SPass &p, SweepArgs a, const void *x, const C *zin, C *zout, void *y,
                    const C *tab, const double *scale, hipStream_t s, bool wide)
{
  a.t0 = p.t0;
  a.ng = p.ng;
  a.bitlo = p.bitlo;
  a.nbits = p.nbits;
  a.logW = p.logW;
  a.lbsplit = p.split;
  const int tileLog = (p.split ? 0 : a.nlbl) + a.nal + p.nbits + p.logW;
  uint64_t ntile = 1ull << (a.L - p.nbits - p.logW);
  if (p.split) {
    // tiles x lastbit values, padded to a multiple of 8 for the XCD remap
    a.nblk = (int64_t)(ntile << a.nlbl);
    ntile = (uint64_t)((a.nblk + 7) / 8 * 8);
  }
  dim3 grid;
  if (ntile <= 65535) grid = dim3((unsigned)ntile);
  else grid = dim3(65536, (unsigned)((ntile + 65535) / 65536));
  const size_t shm = ((size_t(1) << tileLog) + 192) * sizeof(C) + SWZ_TAB * sizeof(int);
  if constexpr (sizeof(C) <= 8) {
    if (wide) {
      hipLaunchKernelGGL((k_sweep_pass<MODEL, DIR, C, TI, TO, 1024>), grid, dim3(1024), shm, s,
                         static_cast<const TI *>(x), zin, zout, static_cast<TO *>(y), a, tab,
                         scale);
      TM_LAUNCH_CHECK();
      return;
    }
  }
  hipLaunchKernelGGL((k_sweep_pass<MODEL, DIR, C, TI, TO, 512>), grid, dim3(512), shm, s,
                     static_cast<const TI *>(x), zin, zout, static_cast<TO *>(y), a, tab, scale);
  TM_LAUNCH_CHECK();
}

template <int MODEL, int DIR>
SweepArgs make_args(const SweepPlan &plan)
{
  SweepArgs a{};
  a.L = plan.L;
  a.model = MODEL;
  a.dir = DIR;
  a.bc = plan.bc;
  if (MODEL == 0) {
    a.nlb = plan.bc == 0 ? 2 : 1;
    a.na = 2;
  } else {
    a.nlb = 4;
    a.na = 4;
  }
  a.nlbl = a.nlb == 1 ? 0 : (a.nlb == 2 ? 1 : 2);
  a.nal = a.na == 2 ? 1 : 2;
  a.dg = plan.diag;
  for (int i = 0; i < 2; ++i) {
    a.b1[i] = plan.b1[i];
    a.b2[i] = plan.b2[i];
    a.b3[i] = plan.b3[i];
    a.db1[i] = plan.db1[i];
    a.db2[i] = plan.db2[i];
    a.db3[i] = plan.db3[i];
  }
  a.outfac = MODEL == 0 ? plan.nf : 1.0;
  a.doutfac = MODEL == 0 ? plan.dnf : 0.0;
  if (MODEL == 1) {
    // the ratio form of the ZBT gate (see zbt_gate)
    const double *b1 = plan.b1, *b2 = plan.b2, *b3 = plan.b3;
    const double r1 = b1[1] / b1[0], r2 = b2[1] / b2[0];
    const double pre = b1[0] * b2[0] * b2[0];
    for (int sidx = 0; sidx < 4; ++sidx) {
      const int sl = sidx >> 1, sr = sidx & 1;
      a.zkap[sidx] = (sr ? 1.0 / r1 : r1) * (sl ? 1.0 / r2 : r2);
      for (int pidx = 0; pidx < 4; ++pidx) {
        const int pl = pidx >> 1, pr = pidx & 1;
        const double q = DIR == 0 ? b1[sr ^ pl] * b3[sr ^ pr] * b3[sl ^ pl] * pre *
                                        (sr ? r1 : 1.0) * (sl ? r2 : 1.0)
                                  : b1[sl ^ pr] * b3[sr ^ pr] * b3[sl ^ pl] * (b2[0] * b2[0]);
        a.zQ[pidx * 4 + sidx] = q;
      }
    }
    a.zr2 = r2;
    for (int i = 0; i < 16; ++i) a.fQ[i] = (float)a.zQ[i];
    for (int i = 0; i < 4; ++i) a.fkap[i] = (float)a.zkap[i];
    a.fr2 = (float)a.zr2;
    for (int i = 0; i < 2; ++i) {
      a.fb1[i] = (float)b1[i];
      a.fb2[i] = (float)b2[i];
    }
  }
  return a;
}

// TI/TO: vector types in/out (TI = TO = the io type, or double -> Dual for
// the dual-number kernels).  The passes work in place in the aux array.
template <int MODEL, int DIR, typename C, typename TI, typename TO>
void sweep_run_tl(const SweepPlan &plan, const void *in, void *out, void *work,
                  const double *scale, hipStream_t s, const C *tab, int cand)
{
  SweepArgs a = make_args<MODEL, DIR>(plan);
  const bool wide = (cand & SWEEP_WIDE) != 0;
  auto ps = plan_sweep(MODEL, plan.L, plan.bc, (int)sizeof(C), cand & 0xFF,
                       (cand & SWEEP_NOSPLIT) == 0);
  C *z = static_cast<C *>(work);
  const int np = (int)ps.size();
  for (int i = 0; i < np; ++i) {
    a.first = (i == 0);
    a.last = (i == np - 1);
    const double *sc = (a.first || a.last) ? scale : nullptr;
    run_sweep_pass<MODEL, DIR, C, TI, TO>(ps[i], a, in, z, z, out, tab, sc, s, wide);
  }
}

template <int MODEL, int DIR, typename C, typename TI, typename TO>
void sweep_run(const SweepPlan &plan, const void *in, void *out, void *work, const double *scale,
               hipStream_t s, const C *tab)
{
  auto run = [&](int tl) {
    sweep_run_tl<MODEL, DIR, C, TI, TO>(plan, in, out, work, scale, s, tab, tl);
  };
  const int es = (int)sizeof(C);
  const uint32_t kid = 4 * MODEL + 2 * DIR + (plan.bc != 0 ? 1 : 0);
  // timing candidates rewrites out from in, so only out-of-place applies tune
  const int lgz = MODEL == 0 ? (plan.bc == 0 ? 2 : 1) : 4;
  run(autotune_tile(TUNE_SWEEP, kid, plan.L, es, sweep_tile_candidates(es, plan.L, lgz),
                    plan.tune && in != out, s, run));
}

template <int MODEL, int DIR, typename C>
void sweep_typed(const SweepPlan &plan, const void *in, void *out, tm_dtype io, void *work,
                 const double *scale, hipStream_t s, const C *tab)
{
  if (io == TM_F64) sweep_run<MODEL, DIR, C, double, double>(plan, in, out, work, scale, s, tab);
  else sweep_run<MODEL, DIR, C, float, float>(plan, in, out, work, scale, s, tab);
}

}  // namespace

size_t sweep_work_bytes(int kind, int L, tm_dtype comp)
{
  const size_t es = comp == TM_F64 ? 8 : 4;
  const size_t mult = kind == TM_KIND_ZBT ? 16 : 4;
  return mult * (size_t(1) << L) * es;
}

size_t sweep_dual_work_bytes(int kind, int L)
{
  const size_t es = sizeof(dev::Dual);
  const size_t mult = kind == TM_KIND_ZBT ? 16 : 4;
  return mult * (size_t(1) << L) * es;
}

void launch_sweep_dual(const SweepPlan &plan, const double *in, void *out, int transpose,
                       void *work, const void *tab_dual, hipStream_t s)
{
  using dev::Dual;
  const Dual *tab = static_cast<const Dual *>(tab_dual);
  if (plan.kind == TM_KIND_XDT)
    sweep_run<0, 0, Dual, double, Dual>(plan, in, out, work, nullptr, s, tab);
  else if (transpose)
    sweep_run<1, 1, Dual, double, Dual>(plan, in, out, work, nullptr, s, tab);
  else
    sweep_run<1, 0, Dual, double, Dual>(plan, in, out, work, nullptr, s, tab);
}

void launch_sweep(const SweepPlan &plan, const void *in, void *out, tm_dtype io, tm_dtype comp,
                  int transpose, void *work, const double *scale, hipStream_t s)
{
  if (plan.kind == TM_KIND_XDT) {
    if (comp == TM_F64) sweep_typed<0, 0, double>(plan, in, out, io, work, scale, s, plan.tab64);
    else sweep_typed<0, 0, float>(plan, in, out, io, work, scale, s, plan.tab32);
  } else {
    if (comp == TM_F64) {
      if (transpose) sweep_typed<1, 1, double>(plan, in, out, io, work, scale, s, plan.tab64);
      else sweep_typed<1, 0, double>(plan, in, out, io, work, scale, s, plan.tab64);
    } else {
      if (transpose) sweep_typed<1, 1, float>(plan, in, out, io, work, scale, s, plan.tab32);
      else sweep_typed<1, 0, float>(plan, in, out, io, work, scale, s, plan.tab32);
    }
  }
}

}  // namespace ltm
