// tm_device.hpp -- device-side helpers: spin-state bit operations and the
// popcount-indexed Boltzmann diagonal.  Bit i of a state is site i; a set
// bit is a down spin (the reference convention, readme.md:100).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tm_internal.hpp"

namespace ltm {
namespace dev {

// Forward-mode dual number (value, derivative with respect to one model
// parameter).  The operator kernels are templates over their arithmetic
// type; instantiated with Dual they apply T and dT/dtheta in one pass.
struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double a) : v(a), d(0.0) {}
  __host__ __device__ constexpr Dual(double a, double b) : v(a), d(b) {}
};
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b)
{
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__host__ __device__ __forceinline__ Dual &operator*=(Dual &a, Dual b) { return a = a * b; }

// a coefficient of arithmetic type C from its value and derivative
template <typename C>
__host__ __device__ __forceinline__ C mk(double v, double) { return (C)v; }
template <>
__host__ __device__ __forceinline__ Dual mk<Dual>(double v, double d) { return Dual(v, d); }

__device__ __forceinline__ int pop(uint64_t x) { return __popcll(x); }

// rotate right by d within n bits (TMat/utility.hpp:91-94)
__device__ __forceinline__ uint64_t rotr(uint64_t s, int n, int d)
{
  return (s >> d) | ((s & ((1ull << d) - 1)) << (n - d));
}
// rotate left by d within n bits (TMat/utility.hpp:99-102)
__device__ __forceinline__ uint64_t rotl(uint64_t s, int n, int d)
{
  return ((s << d) & ((1ull << n) - 1)) | (s >> (n - d));
}
// reverse the low n bits (same result as TMat/utility.cpp:17-31)
__device__ __forceinline__ uint64_t brev(uint64_t s, int n)
{
  return __brevll(s) >> (64 - n);
}
__device__ __forceinline__ uint32_t brev32(uint32_t s, int n)
{
  return __brev(s) >> (32 - n);
}

// 3D row rotation right by 1 / 2 inside every row of s2 bits
// (TMat/utility.cpp:52-63; in diag_index_w below); whole-layer row shift is
// rotr(s, L, s2) (TMat/utility.cpp:68-73).

// Popcount indices (i1, i2, i0) of a state for the diagonal tables.  Each
// branch reproduces the corresponding genTvec of the reference; the table
// entries carry the coupling constants and constant offsets.  W is the state
// word: 32-bit arithmetic (half the VALU work) when the width allows it.
struct Idx3 { int i1, i2, i0; };

__device__ __forceinline__ int popw(uint32_t x) { return __popc(x); }
__device__ __forceinline__ int popw(uint64_t x) { return __popcll(x); }

template <typename W>
__device__ __forceinline__ W rotrw(W s, int n, int d)
{
  return (s >> d) | ((s & ((W(1) << d) - 1)) << (n - d));
}

// One diagonal kind K known at compile time (K = -1: no diagonal term
// beyond the field popcount i0).
template <int K, typename W>
__device__ __forceinline__ Idx3 diag_index_kw(const DiagSpec &g, W s)
{
  const int L = g.L;
  const W mask = (L >= (int)(8 * sizeof(W))) ? ~W(0) : ((W(1) << L) - 1);
  Idx3 r;
  r.i0 = popw(s);
  r.i1 = r.i2 = 0;
  if constexpr (K == DIAG_X2D_PBC || K == DIAG_ZBT) {
    r.i1 = popw(W(s ^ rotrw<W>(s, L, 1)));
    r.i2 = popw(W(s ^ rotrw<W>(s, L, 2)));
  } else if constexpr (K == DIAG_X2D_OBC) {
    r.i1 = popw(W((s & (mask >> 1)) ^ (s >> 1)));
    r.i2 = popw(W((s & (mask >> 2)) ^ (s >> 2)));
  } else if constexpr (K == DIAG_X2D_C5 || K == DIAG_X2D_C9) {
    const W a = W(s << 1) | (K == DIAG_X2D_C9 ? W(1) : W(0));
    r.i1 = popw(W(s ^ a));
    r.i2 = popw(W((a & mask) ^ (s >> 1)));
  } else if constexpr (K == DIAG_X2D_C13 || K == DIAG_X2D_C17) {
    constexpr bool one = K == DIAG_X2D_C17;
    const W a = W(s << 1) | (one ? W(1) : W(0));
    const W b = W(s << 2) | (one ? W(3) : W(0));
    r.i1 = popw(W(s ^ a));
    r.i2 = popw(W(s ^ b));
  } else if constexpr (K == DIAG_X3D) {
    const W f1 = ((s >> 1) & W(g.cmr1)) | ((s & W(g.cm1)) << (g.s2 - 1));
    const W f2 = ((s >> 2) & W(g.cmr2)) | ((s & W(g.cm2)) << (g.s2 - 2));
    r.i1 = popw(W(s ^ f1)) + popw(W(s ^ rotrw<W>(s, L, g.s2)));
    r.i2 = popw(W(s ^ f2));
  } else if constexpr (K == DIAG_XDT_PBC) {
    r.i1 = popw(W(s ^ rotrw<W>(s, L, 1)));
  } else if constexpr (K == DIAG_XDT_C5) {
    r.i1 = popw(W(W(s << 1) ^ s));
  } else if constexpr (K == DIAG_XDT_C9) {
    r.i1 = popw(W((W(s << 1) | W(1)) ^ s));
  } else if constexpr (K == DIAG_Y2TV) {
    const W lf = s & mask, rt = s >> L;
    r.i1 = popw(W(lf ^ rotrw<W>(lf, L, 1))) + popw(W(rt ^ rotrw<W>(rt, L, 1)));
    r.i2 = popw(W(lf ^ rt));
  } else if constexpr (K == DIAG_Y2D) {
    r.i1 = popw(W((s & mask) ^ (s >> L)));
  }
  return r;
}

template <typename W>
__device__ __forceinline__ Idx3 diag_index_w(const DiagSpec &g, W s)
{
  switch (g.kind) {
    case DIAG_X2D_PBC: return diag_index_kw<DIAG_X2D_PBC, W>(g, s);
    case DIAG_ZBT: return diag_index_kw<DIAG_ZBT, W>(g, s);
    case DIAG_X2D_OBC: return diag_index_kw<DIAG_X2D_OBC, W>(g, s);
    case DIAG_X2D_C5: return diag_index_kw<DIAG_X2D_C5, W>(g, s);
    case DIAG_X2D_C9: return diag_index_kw<DIAG_X2D_C9, W>(g, s);
    case DIAG_X2D_C13: return diag_index_kw<DIAG_X2D_C13, W>(g, s);
    case DIAG_X2D_C17: return diag_index_kw<DIAG_X2D_C17, W>(g, s);
    case DIAG_X3D: return diag_index_kw<DIAG_X3D, W>(g, s);
    case DIAG_XDT_PBC: return diag_index_kw<DIAG_XDT_PBC, W>(g, s);
    case DIAG_XDT_C5: return diag_index_kw<DIAG_XDT_C5, W>(g, s);
    case DIAG_XDT_C9: return diag_index_kw<DIAG_XDT_C9, W>(g, s);
    case DIAG_Y2TV: return diag_index_kw<DIAG_Y2TV, W>(g, s);
    case DIAG_Y2D: return diag_index_kw<DIAG_Y2D, W>(g, s);
    default: return diag_index_kw<-1, W>(g, s);
  }
}

// 32-bit words up to 30 state bits: the open-boundary branches shift left by
// up to two bits without masking and count bits L and L+1 as well
__device__ __forceinline__ Idx3 diag_index(const DiagSpec &g, uint64_t s)
{
  const int bits = g.kind >= DIAG_Y2TV ? 2 * g.L : g.L;
  return bits <= 30 ? diag_index_w<uint32_t>(g, (uint32_t)s) : diag_index_w<uint64_t>(g, s);
}

template <typename C>
__device__ __forceinline__ C diag_value(const DiagSpec &g, const C *tab, uint64_t s)
{
  Idx3 i = diag_index(g, s);
  return tab[i.i1] * tab[64 + i.i2] * tab[128 + i.i0];
}
// the same for a compile-time kind and a state of at most 30 bits
template <int K, typename C>
__device__ __forceinline__ C diag_value_k(const DiagSpec &g, const C *tab, uint32_t s)
{
  Idx3 i = diag_index_kw<K, uint32_t>(g, s);
  return tab[i.i1] * tab[64 + i.i2] * tab[128 + i.i0];
}

// Half the two-layer Boltzmann weight W(c, p) of the symmetrised axial
// transfer (the YAM/YBM Tvec without 1/sqrt(dup), TMat/tmatreduce.cpp:717-780;
// also the YAT diagonal Tv(c) K(c, p)), c the current layer, p the previous
// one, from a popcount table: [0,64) NN bonds in c, [64,128) down spins of c,
// [128,192) bonds between c and p; with ybm also [192,256) NNN bonds in c and
// [256,320) the two diagonal bond sets between c and p.
template <typename C>
__device__ __forceinline__ C w_layers(uint32_t c, uint32_t p, int L, const C *tab, int ybm)
{
  const uint32_t m = (L >= 32) ? ~0u : ((1u << L) - 1);
  const uint32_t r1 = ((c >> 1) | (c << (L - 1))) & m;
  C w = tab[__popc(c ^ r1)] * tab[64 + __popc(c)] * tab[128 + __popc(c ^ p)];
  if (ybm) {
    const uint32_t r2 = ((c >> 2) | (c << (L - 2))) & m;
    const uint32_t pr = ((p >> 1) | (p << (L - 1))) & m, pl = ((p << 1) | (p >> (L - 1))) & m;
    w *= tab[192 + __popc(c ^ r2)] * tab[256 + __popc(c ^ pr) + __popc(c ^ pl)];
  }
  return w;
}

// LDS swizzle for 8-byte elements: XOR the low 5 index bits with the next 5.
// A thread that owns 16 consecutive elements (or 16 elements strided by 16)
// then touches distinct banks across its wavefront.  Involution, identity
// for index < 32.
__host__ __device__ __forceinline__ int sw(int e) { return e ^ ((e >> 5) & 31); }

// One radix-2^R Kronecker stage group on R register values: v[i] indexed by
// R bits, K' = [[1, r], [r, 1]] applied on each of the R bits (one FMA per
// output).  K = [[a, b], [b, a]] = a K' with r = b / a: the caller applies
// the factor a^(bits transformed) once per element.
template <int R, typename C>
__device__ __forceinline__ void kron_regs(C *v, C r)
{
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int m = 1 << j;
#pragma unroll
    for (int i = 0; i < (1 << R); ++i) {
      if (!(i & m)) {
        C x0 = v[i], x1 = v[i | m];
        v[i] = x0 + r * x1;
        v[i | m] = x1 + r * x0;
      }
    }
  }
}

// Transform of an LDS array of 2^nbits elements by K' = [[1, r], [r, 1]] on
// bits [bit0, nbits), using every thread of the block (the caller scales by
// a^(nbits - bit0) for K = a K').  Element index p is stored at lds[sw(p)].
// Ends with a block barrier.
template <int R, typename C>
__device__ __forceinline__ void lds_kron_group(C *lds, int nbits, int bg, C r)
{
  const int nitems = 1 << (nbits - R);
  for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
    const int lo = q & ((1 << bg) - 1);
    const int hi = q >> bg;
    const int base = (hi << (bg + R)) | lo;
    C v[1 << R];
#pragma unroll
    for (int i = 0; i < (1 << R); ++i) v[i] = lds[sw(base | (i << bg))];
    kron_regs<R>(v, r);
#pragma unroll
    for (int i = 0; i < (1 << R); ++i) lds[sw(base | (i << bg))] = v[i];
  }
}

// lds_kron_group with the bit group BG known at compile time: sw is linear
// over XOR, so every address is sw(base) ^ sw(i << BG) -- one XOR with an
// immediate per element instead of recomputing the swizzle.
template <int R, int BG, typename C>
__device__ __forceinline__ void lds_kron_group_c(C *lds, int nbits, C r)
{
  const int nitems = 1 << (nbits - R);
  for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
    const int lo = q & ((1 << BG) - 1);
    const int hi = q >> BG;
    const int sb = sw((hi << (BG + R)) | lo);
    C v[1 << R];
#pragma unroll
    for (int i = 0; i < (1 << R); ++i) v[i] = lds[sb ^ sw(i << BG)];
    kron_regs<R>(v, r);
#pragma unroll
    for (int i = 0; i < (1 << R); ++i) lds[sb ^ sw(i << BG)] = v[i];
  }
}

template <int BG, typename C>
__device__ __forceinline__ void lds_kron_step(C *lds, int nbits, C r)
{
  if (nbits <= BG) return;
  switch (nbits - BG) {
    case 1: lds_kron_group_c<1, BG>(lds, nbits, r); break;
    case 2: lds_kron_group_c<2, BG>(lds, nbits, r); break;
    case 3: lds_kron_group_c<3, BG>(lds, nbits, r); break;
    default: lds_kron_group_c<4, BG>(lds, nbits, r); break;
  }
  __syncthreads();
}

// Transform of all nbits <= 16 bits of an LDS array (see lds_kron_group).
// Inlined into every caller: as a called function its registers would be
// outside the caller's register budget (and cost the call ABI).
template <typename C>
__device__ __forceinline__ void lds_kron(C *lds, int nbits, C r)
{
  lds_kron_step<0>(lds, nbits, r);
  lds_kron_step<4>(lds, nbits, r);
  lds_kron_step<8>(lds, nbits, r);
  lds_kron_step<12>(lds, nbits, r);
}

// r = b / a and f = a^nb (with derivatives for dual arithmetic) of the
// butterfly K = [[a, b], [b, a]] = a [[1, r], [r, 1]] over nb bits.
template <typename C>
struct KronCoef {
  C r, f;
};
template <typename C>
inline KronCoef<C> kron_coef(double a, double da, double b, double db, int nb)
{
  const double rv = b / a, rd = (db * a - b * da) / (a * a);
  double fv = 1.0;
  for (int i = 0; i < nb; ++i) fv *= a;
  const double fd = nb > 0 ? nb * (fv / a) * da : 0.0;
  return {mk<C>(rv, rd), mk<C>(fv, fd)};
}

// XCD-aware block remap: hardware dispatches workgroup i to XCD i % 8.
// Launch 8 * ceil(n / 8) workgroups; workgroup i takes logical block
// (i % 8) * ceil(n / 8) + i / 8, so every XCD walks a contiguous range of
// logical blocks (neighbours share that XCD's L2).  A bijection onto
// [0, 8 * ceil(n / 8)); returns -1 for the padding ids >= n.
__device__ __forceinline__ int64_t xcd_remap(int64_t i, int64_t nlogical)
{
  const int64_t per = (nlogical + 7) / 8;
  const int64_t q = (i % 8) * per + i / 8;
  return q < nlogical ? q : -1;
}

}  // namespace dev
}  // namespace ltm
