// tm_axial.hip -- ANNNI transfer along the axial (NNN) direction on gfx950.
//
// A state holds two layers, index (a << L) | b (a the high layer, b the low
// one), 4^L states.  The reference applies L single-bit steps, each a 2x2
// gate on one bit followed by a layer rotation (TMat/tmat.cpp:126-302).  The
// L steps compose to a Kronecker product over one layer, so here:
//
//   YAT   (YATMatVecProd2D, tmat.cpp:207-242):
//         y[(a<<L)|b] = nf Tv(b) K(b, a) z[(b<<L)|a],   z = (x)_{high} K2 x
//   YAT^T (YATMatVecProd2D_T, tmat.cpp:126-183):
//         y = (x)_{high} K2 w,   w[(a<<L)|b] = nf Tv(a) K(a, b) x[(b<<L)|a]
//   Y2AT  (Y2ATMatVecProd2D, tmat.cpp:265-302): each step's 4x4 gate on the
//         bit pair (i, L+i) factors into K2' on both bits and the NN weight
//         between the old low and the new high bit, so
//         y = Tv (x)_{low} K2' D1 (x)_{high} K2' Tv x.
//
// The Kronecker parts are the generic Kronecker passes over a bit range of
// the 2L-bit index (launch_kron_range); the layer swap with its diagonal is
// one LDS-tiled transpose here (64 x 64 tiles: coalesced 512-byte FP64 rows
// in and out).  Tv, K and nf come from a 192-entry popcount table.
#include "tm_device.hpp"

namespace ltm {
namespace {

constexpr int SW_T = 64;         // tile edge
constexpr int SW_THREADS = 256;

// out[(a << L) | b] = f(a, b) in[(b << L) | a] over one 64 x 64 tile (or the
// whole 2^L x 2^L block when L < 6), with f = w_layers(b, a) (wsel 0),
// w_layers(a, b) (wsel 1) or 1 (wsel 2).
template <typename C, typename TI, typename TO>
__global__ __launch_bounds__(SW_THREADS) void k_layer_swap(
    const TI *__restrict__ in, TO *__restrict__ out, int L, const C *__restrict__ tab, int wsel,
    int ybm, const double *__restrict__ scale, int sflags)
{
  __shared__ C tile[SW_T][SW_T + 1];
  __shared__ C stab[320];
  const int ntab = wsel == 2 ? 0 : (ybm ? 320 : 192);
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) stab[i] = tab[i];
  const int n1 = 1 << L;
  const int ts = n1 < SW_T ? n1 : SW_T;
  const int per = n1 / ts;  // tiles per layer index
  const uint32_t ta = blockIdx.x % per, tb = blockIdx.x / per;
  const uint32_t a0 = ta * ts, b0 = tb * ts;
  C sc = 1;
  if (scale) {
    if (sflags & 1) sc = (C)(*scale);
    if (sflags & 2) sc = (C)(1.0 / *scale);
  }
  // load rows b of the input, contiguous in a
  for (int i = threadIdx.x; i < ts * ts; i += blockDim.x) {
    const int rb = i / ts, ca = i % ts;
    tile[rb][ca] = (C)in[((uint64_t)(b0 + rb) << L) | (a0 + ca)];
  }
  __syncthreads();
  // store rows a of the output, contiguous in b
  for (int i = threadIdx.x; i < ts * ts; i += blockDim.x) {
    const int ra = i / ts, cb = i % ts;
    const uint32_t a = a0 + ra, b = b0 + cb;
    C f = sc;
    if (wsel == 0) f *= dev::w_layers<C>(b, a, L, stab, ybm);
    else if (wsel == 1) f *= dev::w_layers<C>(a, b, L, stab, ybm);
    out[((uint64_t)a << L) | b] = (TO)(tile[cb][ra] * f);
  }
}

template <typename C, typename TI, typename TO>
void layer_swap(const void *in, void *out, int L, const C *tab, int wsel, int ybm,
                const double *scale, int sflags, hipStream_t s)
{
  const int n1 = 1 << L;
  const int ts = n1 < SW_T ? n1 : SW_T;
  const uint64_t nblk = (uint64_t)(n1 / ts) * (uint64_t)(n1 / ts);
  hipLaunchKernelGGL((k_layer_swap<C, TI, TO>), dim3((unsigned)nblk), dim3(SW_THREADS), 0, s,
                     static_cast<const TI *>(in), static_cast<TO *>(out), L, tab, wsel, ybm,
                     scale, sflags);
  TM_LAUNCH_CHECK();
}

template <typename C>
void swap_io(const void *in, void *out, tm_dtype ti, tm_dtype to, int L, const C *tab, int wsel,
             int ybm, const double *scale, int sflags, hipStream_t s)
{
  if (ti == TM_F64 && to == TM_F64) layer_swap<C, double, double>(in, out, L, tab, wsel, ybm, scale, sflags, s);
  else if (ti == TM_F64) layer_swap<C, double, float>(in, out, L, tab, wsel, ybm, scale, sflags, s);
  else if (to == TM_F64) layer_swap<C, float, double>(in, out, L, tab, wsel, ybm, scale, sflags, s);
  else layer_swap<C, float, float>(in, out, L, tab, wsel, ybm, scale, sflags, s);
}

template <typename C>
void axial_typed(const AxialPlan &p, const void *in, void *out, tm_dtype io, int transpose,
                 void *work, const double *scale, hipStream_t s, const C *tab, const C *tabb)
{
  const int L = p.L;
  const tm_dtype ct = sizeof(C) == 8 ? TM_F64 : TM_F32;
  DiagSpec none{};
  none.kind = -1;
  if (p.kind == TM_KIND_YAT) {
    const KronRange hi{2 * L, L, L, false, false, none, none, nullptr, nullptr, p.ka, p.kb};
    if (!transpose) {
      // z = (x)_{high} K2 x into the workspace, then the swap with Tv(low)
      launch_kron_range(hi, in, work, work, io, ct, ct, scale, true, false, p.tune, s);
      swap_io<C>(work, out, ct, io, L, tab, 0, 0, scale, 2, s);
    } else {
      // the swap with Tv(high) into the workspace, then (x)_{high} K2
      swap_io<C>(in, work, io, ct, L, tab, 1, 0, scale, 1, s);
      const bool out_c = io == ct;
      launch_kron_range(hi, work, out, out_c ? out : work, ct, io, ct, scale, false, true,
                        p.tune, s);
    }
    return;
  }
  // Y2AT (its transposed instance is the operator itself, TMat/tmat.hpp:382-384)
  DiagSpec tv{};
  tv.kind = DIAG_Y2TV;
  tv.L = L;
  DiagSpec d1{};
  d1.kind = DIAG_Y2D;
  d1.L = L;
  const KronRange hi{2 * L, L, L, true, true, tv, d1, tab, tabb, p.ka, p.kb};
  const KronRange lo{2 * L, 0, L, false, true, none, tv, nullptr, tab, p.ka, p.kb};
  launch_kron_range(hi, in, work, work, io, ct, ct, scale, true, false, p.tune, s);
  const bool out_c = io == ct;
  launch_kron_range(lo, work, out, out_c ? out : work, ct, io, ct, scale, false, true, p.tune, s);
}

}  // namespace

void launch_layer_swap(const void *in, void *out, tm_dtype ti, tm_dtype to, tm_dtype comp, int L,
                       const void *tab, int wsel, int ybm, const double *scale, int sflags,
                       hipStream_t s)
{
  if (comp == TM_F64)
    swap_io<double>(in, out, ti, to, L, static_cast<const double *>(tab), wsel, ybm, scale,
                    sflags, s);
  else
    swap_io<float>(in, out, ti, to, L, static_cast<const float *>(tab), wsel, ybm, scale,
                   sflags, s);
}

size_t axial_work_bytes(int L, tm_dtype comp)
{
  return (comp == TM_F64 ? 8 : 4) * (size_t(1) << (2 * L));
}

void launch_axial(const AxialPlan &plan, const void *in, void *out, tm_dtype io, tm_dtype comp,
                  int transpose, void *work, const double *scale, hipStream_t s)
{
  if (comp == TM_F64)
    axial_typed<double>(plan, in, out, io, transpose, work, scale, s, plan.tab64, plan.tabb64);
  else
    axial_typed<float>(plan, in, out, io, transpose, work, scale, s, plan.tab32, plan.tabb32);
}

}  // namespace ltm
