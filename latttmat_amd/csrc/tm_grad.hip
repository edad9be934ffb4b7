// tm_grad.hip -- Hellmann-Feynman derivatives of an eigenvalue with respect to
// the model parameters theta = (betaJ, kappa, kappa2, h):
//
//     d lambda / d theta = <l, (dT/dtheta) r> / <l, r>
//
// with r the right and l the left eigenvector (l = r for symmetric T).  The
// "dT/dtheta-weighted" operator is the operator kernel itself instantiated
// with dual numbers: every Boltzmann factor, diagonal-table entry and
// normfactor carries (value, d/dtheta), built on the host by the same
// templated code that builds the operator (tm_consts.hpp), so a single pass
// over r returns T r and (dT/dtheta) r exactly.  The reference has no
// derivatives; the operators differentiated are those of TMat/tmat.cpp
// (XAT, XDT, ZBT), TMat/tmatreduce.cpp (XAM) and TMat/tmat3d.cpp (3D).
//
// Reduced operators: T_red = P T P^T with a parameter-independent projection
// (orbit sizes are fixed), so the numerators are computed on the full space
// after expanding l and r (expand_orbits).
#include <memory>
#include <vector>

#include "tm_consts.hpp"
#include "tm_device.hpp"

namespace ltm {
namespace {

using dev::Dual;
constexpr int DOT_THREADS = 256;
constexpr int DOT_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// partial[3b + {0,1,2}] = sum over block b's grid-stride range of
// l*y.v, l*y.d and l*r.
__global__ __launch_bounds__(DOT_THREADS) void k_dual_dots(const double *__restrict__ l,
                                                           const Dual *__restrict__ y,
                                                           const double *__restrict__ r,
                                                           int64_t n, double *__restrict__ partial)
{
  double a0 = 0, a1 = 0, a2 = 0;
  for (int64_t i = blockIdx.x * (int64_t)DOT_THREADS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * DOT_THREADS) {
    const double li = l[i];
    const Dual yi = y[i];
    a0 += li * yi.v;
    a1 += li * yi.d;
    a2 += li * r[i];
  }
  __shared__ double red[3][DOT_THREADS / 64];
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = a0;
    red[1][wv] = a1;
    red[2][wv] = a2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0;
    for (int w = 0; w < DOT_THREADS / 64; ++w) t += red[threadIdx.x][w];
    partial[3 * blockIdx.x + threadIdx.x] = t;
  }
}

// Constants of a full-space operator kind over scalar S.
template <class S>
OpConsts<S> consts_for(const tm_op *op, S bJ, S k, S k2, S h)
{
  switch (op->kind) {
    case TM_KIND_XAT: return consts_xat<S>(op->L, op->bc, bJ, k, h);
    // the reduced operator's energies count both boundaries with L bonds
    // (TMat/tmatreduce.cpp:384-399), i.e. the periodic XAT constants
    case TM_KIND_XAM2D: return consts_xat<S>(op->L, 0, bJ, k, h);
    case TM_KIND_XDT: return consts_xdt<S>(op->L, op->bc, bJ, k, h);
    case TM_KIND_ZBT: return consts_zbt<S>(op->L, bJ, k, k2, h);
    // the reduced circuits carry the periodic full-space constants
    case TM_KIND_XDM: return consts_xdt<S>(op->L, 0, bJ, k, h);
    case TM_KIND_ZBM: return consts_zbt<S>(op->L, bJ, k, k2, h);
    case TM_KIND_XAT3D:
    case TM_KIND_XAM3D: return consts_xat3d<S>(op->L, bJ, k, h);
    default: throw Error(TM_EINVAL, "unknown operator kind");
  }
}

OpConsts<HDual> seeded(const tm_op *op, int theta)
{
  const tm_params &p = op->p;
  return consts_for<HDual>(op, HDual(p.betaJ, theta == 0), HDual(p.kappa, theta == 1),
                           HDual(p.kappa2, theta == 2), HDual(p.h, theta == 3));
}

// Numerators on a full-space operator op whose constants are those of cop
// (cop = op, or the reduced operator op was expanded from).
void full_numerators(const tm_op *cop, tm_op *op, const double *l, const double *r,
                     double num[5])
{
  hipStream_t s = op->stream;
  const int64_t n = op->ncol;
  DevBuf y(sizeof(Dual) * n), tabd(sizeof(Dual) * 192), partial(sizeof(double) * 3 * DOT_BLOCKS);
  DevBuf work;
  if (op->kind == TM_KIND_XDT || op->kind == TM_KIND_ZBT)
    work.alloc(sweep_dual_work_bytes(op->kind, op->L));
  std::vector<double> hp(3 * DOT_BLOCKS);
  double lr = 0;
  const bool has_k2 = cop->kind == TM_KIND_ZBT || cop->kind == TM_KIND_ZBM;
  for (int th = 0; th < 4; ++th) {
    if (th == 2 && !has_k2) {
      num[2] = 0.0;  // kappa2 enters only the diagonal-TM couplings
      continue;
    }
    const OpConsts<HDual> c = seeded(cop, th);
    Dual tab[192];
    for (int i = 0; i < 192; ++i) tab[i] = Dual(c.tab[i].v, c.tab[i].d);
    TM_HIP(hipMemcpyAsync(tabd.get(), tab, sizeof tab, hipMemcpyHostToDevice, s));
    if (op->kind == TM_KIND_XAT || op->kind == TM_KIND_XAT3D) {
      KronPlan kp{};
      kp.L = op->L;
      kp.prologue = kp.epilogue = true;
      kp.diag = op->diag;
      kp.ka = c.b1[0].v;
      kp.kb = c.b1[1].v;
      launch_kron_dual(kp, r, y.get(), tabd.get(), c.b1[0].d, c.b1[1].d, s);
    } else {
      SweepPlan sp{};
      sp.kind = op->kind;
      sp.L = op->L;
      sp.bc = op->bc;
      sp.diag = op->diag;
      for (int i = 0; i < 2; ++i) {
        sp.b1[i] = c.b1[i].v;
        sp.b2[i] = c.b2[i].v;
        sp.b3[i] = c.b3[i].v;
        sp.db1[i] = c.b1[i].d;
        sp.db2[i] = c.b2[i].d;
        sp.db3[i] = c.b3[i].d;
      }
      sp.nf = c.nf.v;
      sp.dnf = c.nf.d;
      launch_sweep_dual(sp, r, y.get(), 0, work.get(), tabd.get(), s);
    }
    hipLaunchKernelGGL(k_dual_dots, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, l,
                       y.as<Dual>(), r, n, partial.as<double>());
    TM_LAUNCH_CHECK();
    TM_HIP(hipMemcpyAsync(hp.data(), partial.get(), sizeof(double) * hp.size(),
                          hipMemcpyDeviceToHost, s));
    TM_HIP(hipStreamSynchronize(s));
    double d = 0;
    lr = 0;
    for (int b = 0; b < DOT_BLOCKS; ++b) {
      d += hp[3 * b + 1];
      lr += hp[3 * b + 2];
    }
    num[th] = d;
  }
  num[4] = lr;
  op->matvecs += has_k2 ? 4 : 3;
}

}  // namespace

void grad_numerators(tm_op *op, const double *l, const double *r, double num[5])
{
  if (!is_reduced(op->kind)) {
    full_numerators(op, op, l, r, num);
    return;
  }
  // reduced: the same numerators on the expanded full-space vectors
  // (<P l, dT P r> = <l, P^T dT P r>, <P l, P r> = <l, r>)
  tm_params q = op->p;
  if (op->kind == TM_KIND_XAM2D) q.axialdir = op->bc ? 9 : 1;
  else if (op->kind == TM_KIND_XDM) q.axialdir = 1;
  else if (op->kind == TM_KIND_ZBM) q.axialdir = 3;
  else q.flags |= TM_FLAG_FULLSPACE;
  q.flags &= ~TM_FLAG_GENERIC_REDUCED;
  std::unique_ptr<tm_op> full(make_op(q));
  const bool two = op->kind != TM_KIND_XAM3D;
  const GroupSpec g{two ? 2 : 3, op->L, op->s1, op->s2, op->kind == TM_KIND_XAM2D ? op->bc : 0,
                    op->censym, op->rotate, op->kind == TM_KIND_ZBM ? 1 : 0};
  const size_t bytes = sizeof(double) * (size_t)full->ncol;
  DevBuf lf(bytes), rf(bytes);
  expand_orbits(g, op->reps.as<uint64_t>(), op->dups.as<uint16_t>(), op->ncol, l,
                lf.as<double>(), full->stream);
  expand_orbits(g, op->reps.as<uint64_t>(), op->dups.as<uint16_t>(), op->ncol, r,
                rf.as<double>(), full->stream);
  full_numerators(op, full.get(), lf.as<double>(), rf.as<double>(), num);
  op->matvecs += 4;
}

void dlnnf(const tm_op *op, double out[4])
{
  for (int th = 0; th < 4; ++th) {
    const OpConsts<HDual> c = seeded(op, th);
    out[th] = c.nf.d / c.nf.v;
  }
}

}  // namespace ltm
