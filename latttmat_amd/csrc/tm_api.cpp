// tm_api.cpp -- the C ABI (include/latttmat.h).  Every entry point maps
// C++ exceptions to a tm_status and a thread-local message.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "tm_internal.hpp"

namespace ltm {
tm_status solve_impl(tm_op *op, const tm_solve_opts &o, tm_solve_out *out);
}

namespace {
thread_local std::string g_err;

tm_status fail(tm_status s, const char *msg)
{
  g_err = msg;
  return s;
}

template <typename F>
tm_status guarded(F &&f)
{
  try {
    return f();
  } catch (const ltm::Error &e) {
    return fail(e.status, e.what());
  } catch (const std::bad_alloc &) {
    return fail(TM_ENOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(TM_EINVAL, e.what());
  }
}
}  // namespace

extern "C" {

const char *tm_last_error(void) { return g_err.c_str(); }
const char *tm_version(void) { return "latttmat-mi355x 0.1 (gfx950)"; }

tm_status tm_op_create(const tm_params *p, tm_op **out)
{
  if (!p || !out) return fail(TM_EINVAL, "null argument");
  *out = nullptr;
  return guarded([&] {
    *out = ltm::make_op(*p);
    return TM_OK;
  });
}

void tm_op_destroy(tm_op *op) { delete op; }

int64_t tm_op_rows(const tm_op *op) { return op ? op->ncol : 0; }
int64_t tm_op_cols(const tm_op *op) { return op ? op->ncol : 0; }
double tm_op_normfactor(const tm_op *op) { return op ? op->nf : std::nan(""); }
int tm_op_is_symmetric(const tm_op *op) { return op && op->symmetric ? 1 : 0; }
tm_kind tm_op_kind(const tm_op *op) { return op ? op->kind : (tm_kind)0; }
int tm_op_suggested_ncv(const tm_op *, int neigs) { return neigs > 4 ? 2 * neigs + 1 : 2 * neigs + 3; }
int tm_op_layer_size(const tm_op *op) { return op ? op->L : 0; }
int64_t tm_op_matvec_count(const tm_op *op) { return op ? op->matvecs : 0; }

tm_status tm_op_apply(const tm_op *cop, const double *x, double *y, int transpose)
{
  if (!cop || !x || !y) return fail(TM_EINVAL, "null argument");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    const size_t bytes = sizeof(double) * op->ncol;
    op->hx.ensure(bytes);
    op->hy.ensure(bytes);
    ltm::copy_large(op->hx.get(), x, bytes, hipMemcpyHostToDevice, op->stream);
    op->apply_dev(op->hx.get(), op->hy.get(), op->ncol, 1, TM_F64, TM_F64, transpose, nullptr,
                  op->stream);
    ltm::copy_large(y, op->hy.get(), bytes, hipMemcpyDeviceToHost, op->stream);
    return TM_OK;
  });
}

tm_status tm_op_apply_dev(const tm_op *cop, const void *x, void *y, int64_t ld, int nvec,
                          tm_dtype dtype, tm_dtype compute, int transpose, void *stream)
{
  if (!cop || !x || !y || nvec < 1 || ld < cop->ncol) return fail(TM_EINVAL, "bad argument");
  if ((dtype != TM_F64 && dtype != TM_F32) || (compute != TM_F64 && compute != TM_F32))
    return fail(TM_EINVAL, "bad dtype");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    hipStream_t s = static_cast<hipStream_t>(stream);
    // The apply uses the operator's workspaces (op->work, px, py), which
    // the operator's own stream also uses (tm_op_apply, tm_solve): order the
    // caller's stream after the operator's pending work, and the operator's
    // stream after this apply.
    ltm::stream_wait(s, op->stream);
    try {
      if (compute == TM_F32) {
        // per-vector power-of-two scale so that max|x| is in [1/2, 1)
        ltm::DevBuf partial(sizeof(double) * ltm::kry::NBLK);
        ltm::DevBuf scale(sizeof(double) * nvec);
        const size_t es = dtype == TM_F64 ? 8 : 4;
        if (dtype == TM_F64) {
          for (int v = 0; v < nvec; ++v) {
            ltm::kry::pow2_scale(static_cast<const double *>(x) + (size_t)v * ld, op->ncol,
                                 partial.as<double>(), scale.as<double>() + v, s);
            op->apply_dev(static_cast<const char *>(x) + (size_t)v * ld * es,
                          static_cast<char *>(y) + (size_t)v * ld * es, ld, 1, dtype, compute,
                          transpose, scale.as<double>() + v, s);
          }
        } else {
          op->apply_dev(x, y, ld, nvec, dtype, compute, transpose, nullptr, s);
        }
        TM_HIP(hipStreamSynchronize(s));  // scratch buffers are released on return
      } else {
        op->apply_dev(x, y, ld, nvec, dtype, compute, transpose, nullptr, s);
      }
    } catch (...) {
      try {
        ltm::stream_wait(op->stream, s);
      } catch (...) {
      }
      throw;
    }
    ltm::stream_wait(op->stream, s);
    return TM_OK;
  });
}

tm_status tm_op_apply_dev_noalloc(const tm_op *cop, const void *x, void *y, int64_t ld, int nvec,
                                  tm_dtype dtype, tm_dtype compute, int transpose, void *stream)
{
  if (!cop || !x || !y || nvec < 1 || ld < cop->ncol) return fail(TM_EINVAL, "bad argument");
  if ((dtype != TM_F64 && dtype != TM_F32) || (compute != TM_F64 && compute != TM_F32))
    return fail(TM_EINVAL, "bad dtype");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    ltm::apply_dev_noalloc(op, x, y, ld, nvec, dtype, compute, transpose,
                           static_cast<hipStream_t>(stream));
    return TM_OK;
  });
}

tm_status tm_op_apply_prec(const tm_op *cop, const double *x, double *y, int transpose,
                           tm_dtype compute)
{
  if (!cop || !x || !y) return fail(TM_EINVAL, "null argument");
  if (compute != TM_F64 && compute != TM_F32) return fail(TM_EINVAL, "bad compute dtype");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    const size_t bytes = sizeof(double) * op->ncol;
    op->hx.ensure(bytes);
    op->hy.ensure(bytes);
    ltm::copy_large(op->hx.get(), x, bytes, hipMemcpyHostToDevice, op->stream);
    ltm::DevBuf partial(sizeof(double) * ltm::kry::NBLK), scale(sizeof(double));
    const double *sc = nullptr;
    if (compute == TM_F32) {
      ltm::kry::pow2_scale(op->hx.as<double>(), op->ncol, partial.as<double>(),
                           scale.as<double>(), op->stream);
      sc = scale.as<double>();
    }
    op->apply_dev(op->hx.get(), op->hy.get(), op->ncol, 1, TM_F64, compute, transpose, sc,
                  op->stream);
    ltm::copy_large(y, op->hy.get(), bytes, hipMemcpyDeviceToHost, op->stream);
    return TM_OK;
  });
}

tm_status tm_op_time_apply(const tm_op *cop, tm_dtype io, tm_dtype compute, int transpose,
                           int warmup, int reps, double *ms_per_apply)
{
  if (!cop || !ms_per_apply || reps < 1 || warmup < 0) return fail(TM_EINVAL, "bad argument");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    const size_t es = io == TM_F64 ? 8 : 4;
    const int64_t n = op->ncol;
    ltm::DevBuf x(es * n), y(es * n), tmp(sizeof(double) * n);
    ltm::kry::random_uniform(tmp.as<double>(), n, 12345, op->stream);
    if (io == TM_F64) ltm::kry::copy(tmp.as<double>(), x.as<double>(), n, op->stream);
    else ltm::kry::to_f32(tmp.as<double>(), x.as<float>(), n, op->stream);
    // the eigensolver's T.x: vectors in the operator's internal order
    for (int i = 0; i < warmup; ++i)
      op->apply_internal(x.get(), y.get(), n, 1, io, compute, transpose, nullptr, op->stream);
    hipEvent_t e0, e1;
    TM_HIP(hipEventCreate(&e0));
    TM_HIP(hipEventCreate(&e1));
    TM_HIP(hipEventRecord(e0, op->stream));
    for (int i = 0; i < reps; ++i)
      op->apply_internal(x.get(), y.get(), n, 1, io, compute, transpose, nullptr, op->stream);
    TM_HIP(hipEventRecord(e1, op->stream));
    TM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_apply = ms / reps;
    return TM_OK;
  });
}

tm_status tm_op_orbits(const tm_op *op, uint64_t *reps, uint16_t *dups)
{
  if (!op || !reps || !dups) return fail(TM_EINVAL, "null argument");
  if (!ltm::is_reduced(op->kind))
    return fail(TM_EINVAL, "operator has no orbit tables (not a reduced operator)");
  return guarded([&] {
    ltm::copy_sync(reps, op->reps.get(), sizeof(uint64_t) * op->ncol, hipMemcpyDeviceToHost, op->stream);
    ltm::copy_sync(dups, op->dups.get(), sizeof(uint16_t) * op->ncol, hipMemcpyDeviceToHost, op->stream);
    return TM_OK;
  });
}

void tm_solve_opts_default(tm_solve_opts *o)
{
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->nev = 1;
  o->ncv = 0;
  o->tol = 1e-10;
  o->maxit = 1000;
  o->maxtrial = 3;
  o->precision = TM_PREC_F64;
  o->seed = 1;
  o->want_left = 0;
  o->want_vectors = 0;
  o->mixed_tol = 2e-7;
  o->verbose = 0;
  o->use_graph = 1;
  o->basis_f32 = 1;
  o->refine_extra = 1;
  o->refine = 0;
  o->verify = 1;
}

tm_status tm_solve(tm_op *op, const tm_solve_opts *opts, tm_solve_out *out)
{
  if (!op || !opts || !out) return fail(TM_EINVAL, "null argument");
  return guarded([&] {
    tm_status st = ltm::solve_impl(op, *opts, out);
    if (st != TM_OK) g_err = "eigensolver did not converge all requested pairs";
    return st;
  });
}

tm_status tm_solve_batch(tm_op *const *ops, int n, const tm_solve_opts *opts, tm_solve_out *outs,
                         int nworkers)
{
  if (n < 0 || (n > 0 && (!ops || !opts || !outs))) return fail(TM_EINVAL, "null argument");
  for (int i = 0; i < n; ++i)
    if (!ops[i]) return fail(TM_EINVAL, "null operator in batch");
  if (n == 0) return TM_OK;
  // largest first: the long solves start at once, the short ones fill in
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return ops[a]->ncol > ops[b]->ncol; });
  const int nw = std::max(1, std::min(nworkers > 0 ? nworkers : 4, n));
  std::atomic<int> next{0};
  std::mutex mu;
  tm_status first = TM_OK;
  std::string first_msg;
  auto worker = [&] {
    for (;;) {
      const int q = next.fetch_add(1);
      if (q >= n) break;
      const int i = order[q];
      const tm_status st = tm_solve(ops[i], &opts[i], &outs[i]);
      // a batch solves each instance once: free its solver buffers now
      // rather than holding every instance's basis until the ops go
      ops[i]->scache.reset();
      if (st != TM_OK) {
        std::lock_guard<std::mutex> lk(mu);
        if (first == TM_OK) {
          first = st;
          first_msg = "instance " + std::to_string(i) + ": " + g_err;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int w = 1; w < nw; ++w) th.emplace_back(worker);
  worker();
  for (auto &t : th) t.join();
  if (first != TM_OK) g_err = first_msg;
  return first;
}

tm_status tm_op_moments(const tm_op *op, const double *ql, const double *qr, double out[8],
                        int *nout)
{
  if (!op || !qr || !out || !nout) return fail(TM_EINVAL, "null argument");
  using namespace ltm;
  int pk, bc = 0;
  switch (op->kind) {
    case TM_KIND_XAT: pk = PHYS_SYM2D; bc = op->bc; break;      // TMat/tmat.cpp:39
    case TM_KIND_XAM2D: pk = PHYS_SYM2D; bc = op->bc; break;    // ext & 1, TMat/tmatreduce.cpp:338
    case TM_KIND_XDT: pk = PHYS_SYM2D; break;                   // periodic, TMat/tmat.cpp:322-323
    case TM_KIND_XAT3D: case TM_KIND_XAM3D: pk = PHYS_SYM3D; break;
    case TM_KIND_XDM: pk = PHYS_DNNI; break;
    case TM_KIND_ZBT: case TM_KIND_ZBM: pk = PHYS_GEN; break;
    case TM_KIND_YAT: pk = PHYS_GEN_LOW; break;
    case TM_KIND_YAM: case TM_KIND_YBM: pk = PHYS_GEN_HIGH; break;
    case TM_KIND_Y2AT: pk = PHYS_ZERO; break;
    default: return fail(TM_EINVAL, "unknown operator kind");
  }
  return guarded([&] {
    const int L = op->L;
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    if (pk == PHYS_ZERO) {
      *nout = 4;
      return TM_OK;
    }
    const bool gen = pk == PHYS_GEN || pk == PHYS_GEN_LOW || pk == PHYS_GEN_HIGH;
    const size_t bytes = sizeof(double) * op->ncol;
    DevBuf dl(bytes), dr(gen ? bytes : 0), work(phys_work_bytes());
    copy_large(dl.get(), ql ? ql : qr, bytes, hipMemcpyHostToDevice, op->stream);
    if (gen) copy_large(dr.get(), qr, bytes, hipMemcpyHostToDevice, op->stream);
    double s[9];
    phys_sums(pk, dl.as<double>(), gen ? dr.as<double>() : nullptr,
              is_reduced(op->kind) ? op->reps.as<uint64_t>() : nullptr, op->ncol, L, bc,
              work.get(), s, op->stream);
    const double div = gen ? L * s[0] : (double)L;
    for (int k = 0; k < 4; ++k) out[k] = s[1 + k] / div;
    if (pk == PHYS_SYM2D) {
      out[4] = s[5];
      out[5] = s[6];
      *nout = 6;
    } else if (pk == PHYS_SYM3D) {
      out[4] = out[5] = NAN;
      *nout = 6;
    } else if (pk == PHYS_DNNI) {
      out[4] = s[5] / L;
      out[5] = s[6] / L;
      out[6] = s[7] / ((double)L * L);
      out[7] = s[8] / ((double)L * L);
      *nout = 8;
    } else {
      *nout = 4;
    }
    return TM_OK;
  });
}

tm_status tm_sample_configs(const tm_op *cop, const double *vec, const double *vecl,
                            const double *u, int nchain, int length, uint64_t *states)
{
  if (!cop || !vec || !u || !states) return fail(TM_EINVAL, "null argument");
  if (nchain < 1 || length < 1) return fail(TM_EINVAL, "nchain and length must be positive");
  const tm_kind k = cop->kind;
  const bool sym = k == TM_KIND_XAT || k == TM_KIND_XAM2D || k == TM_KIND_XDM;
  const bool gen = k == TM_KIND_YAT || k == TM_KIND_ZBM || k == TM_KIND_YAM || k == TM_KIND_YBM;
  if (!sym && !gen)
    return fail(TM_EUNSUPPORTED,
                "configuration sampling is built for the models dumpconfig serves "
                "(XAT, YAT, XAM2D, XDM, ZBM, YAM, YBM)");
  if (k == TM_KIND_XAT && cop->bc != 0)
    return fail(TM_EUNSUPPORTED, "the XAT sampler is periodic (dumpconfig serves axialdir 1)");
  if (gen && !vecl) return fail(TM_EINVAL, "this operator's sampler needs the left eigenvector");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    using namespace ltm;
    const int64_t n = op->ncol;
    const bool onestep = k == TM_KIND_XAT || k == TM_KIND_YAT;
    const int64_t per = onestep ? length : 2 * (int64_t)length - 1;
    const int64_t nu = per * nchain;
    DevBuf dv(sizeof(double) * n), dl(gen ? sizeof(double) * n : 0), du(sizeof(double) * nu),
        ds(sizeof(uint64_t) * nchain * (int64_t)length), derr(sizeof(int));
    // absolute values, as DumpConfig<S>::readfile takes them (dumpconfig.hpp:71)
    std::vector<double> av(vec, vec + n);
    for (double &x : av) x = std::fabs(x);
    copy_large(dv.get(), av.data(), sizeof(double) * n, hipMemcpyHostToDevice, op->stream);
    if (gen) {
      av.assign(vecl, vecl + n);
      for (double &x : av) x = std::fabs(x);
      copy_large(dl.get(), av.data(), sizeof(double) * n, hipMemcpyHostToDevice, op->stream);
    }
    copy_sync(du.get(), u, sizeof(double) * nu, hipMemcpyHostToDevice, op->stream);
    const int zero = 0;
    copy_sync(derr.get(), &zero, sizeof zero, hipMemcpyHostToDevice, op->stream);
    sample_chains(op, dv.as<double>(), gen ? dl.as<double>() : nullptr, du.as<double>(), nchain,
                  length, ds.as<uint64_t>(), derr.as<int>());
    copy_sync(states, ds.get(), sizeof(uint64_t) * nchain * (int64_t)length,
              hipMemcpyDeviceToHost, op->stream);
    int err = 0;
    copy_sync(&err, derr.get(), sizeof err, hipMemcpyDeviceToHost, op->stream);
    if (err) throw Error(TM_EINVAL, "orbit member count disagrees with the orbit size");
    return TM_OK;
  });
}

tm_status tm_grad_hf(const tm_op *cop, const double *l, const double *r, double dlam[4],
                     double dlnnf[4])
{
  if (!cop || !l || !r || !dlam) return fail(TM_EINVAL, "null argument");
  if (cop->kind == TM_KIND_YAT || cop->kind == TM_KIND_Y2AT || cop->kind == TM_KIND_YAM ||
      cop->kind == TM_KIND_YBM)
    return fail(TM_EUNSUPPORTED, "no dual-number kernels for the axial (two-layer) operators");
  tm_op *op = const_cast<tm_op *>(cop);
  return guarded([&] {
    const size_t bytes = sizeof(double) * op->ncol;
    ltm::DevBuf dl(bytes), dr(bytes);
    ltm::copy_large(dl.get(), l, bytes, hipMemcpyHostToDevice, op->stream);
    ltm::copy_large(dr.get(), r, bytes, hipMemcpyHostToDevice, op->stream);
    double num[5];
    ltm::grad_numerators(op, dl.as<double>(), dr.as<double>(), num);
    // num[4] = <l, r>
    for (int i = 0; i < 4; ++i) dlam[i] = num[i] / num[4];
    if (dlnnf) ltm::dlnnf(op, dlnnf);
    return TM_OK;
  });
}

}  // extern "C"
