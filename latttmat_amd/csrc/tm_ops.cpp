// tm_ops.cpp -- operator construction (the pfunc dispatch of the reference,
// TMat/pfunc.cpp:16-106) and device table set-up.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>
#include <map>
#include <tuple>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "tm_consts.hpp"
#include "tm_internal.hpp"

namespace ltm {

namespace {

int pop64(uint64_t x) { return __builtin_popcountll(x); }
uint64_t rotr_h(uint64_t s, int n, int d) { return (s >> d) | ((s & ((1ull << d) - 1)) << (n - d)); }

// 3D column masks for DiagSpec (bits of columns [0,c) of every row)
uint64_t colmask(int s1, int s2, int c)
{
  uint64_t m = 0;
  for (int r = 0; r < s1; ++r) m |= ((1ull << c) - 1) << (r * s2);
  return m;
}

void upload_tables(tm_op *op, const std::vector<double> &t, bool second = false)
{
  std::vector<float> f(t.begin(), t.end());
  DevBuf &d64 = second ? op->tabb64 : op->tab64;
  DevBuf &d32 = second ? op->tabb32 : op->tab32;
  d64.alloc(t.size() * sizeof(double));
  d32.alloc(t.size() * sizeof(float));
  copy_sync(d64.get(), t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, op->stream);
  copy_sync(d32.get(), f.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, op->stream);
}

// normfactor of the X-type operators; exponent (units of betaJ) in *e
double nf_x2d(int L, double bJ, double k, double *e)
{
  *e = nfexp_x2d(L, k);
  return std::exp(bJ * *e);
}
double nf_x3d(int size, double bJ, double k, double *e)
{
  *e = nfexp_x3d(size, k);
  return std::exp(bJ * *e);
}

// Upload the constants of a full-space operator: tables, Boltzmann factors,
// normfactor.
void apply_consts(tm_op *op, const OpConsts<double> &c)
{
  upload_tables(op, std::vector<double>(c.tab, c.tab + 192));
  for (int i = 0; i < 2; ++i) {
    op->bzwj1[i] = c.b1[i];
    op->bzwj2[i] = c.b2[i];
    op->bzwj3[i] = c.b3[i];
  }
  op->nf = c.nf;
}

double energy_half_x2d(uint64_t s, int L, int obc, double bJ, double k, double h)
{
  const uint64_t mask = (1ull << L) - 1;
  int n1, n2;
  if (!obc) {
    n1 = pop64(s ^ rotr_h(s, L, 1));
    n2 = pop64(s ^ rotr_h(s, L, 2));
  } else {
    n1 = pop64((s & (mask >> 1)) ^ (s >> 1));
    n2 = pop64((s & (mask >> 2)) ^ (s >> 2));
  }
  const int n0 = pop64(s);
  const double hb = 0.5 * bJ;
  return -hb * (2 * n1 - L) + hb * k * (2 * n2 - L) - hb * h * (2 * n0 - L);
}

uint64_t rff_h(uint64_t s, int s1, int s2, int d)
{
  const uint64_t m = (1ull << s2) - 1;
  uint64_t r = 0;
  for (int i = 0; i < s1; ++i) {
    const uint64_t row = (s >> (i * s2)) & m;
    r |= rotr_h(row, s2, d) << (i * s2);
  }
  return r;
}

// 3D layer exponent, TMat/tmat3d.cpp:188-199 (without the normfactor term)
double energy_half_x3d(uint64_t s, int s1, int s2, double bJ, double k, double h)
{
  const int size = s1 * s2;
  const double hb = 0.5 * bJ;
  const int n1a = pop64(s ^ rff_h(s, s1, s2, 1));
  const int n1b = pop64(s ^ rotr_h(s, size, s2));
  const int n2 = pop64(s ^ rff_h(s, s1, s2, 2));
  const int n0 = pop64(s);
  return 2 * (-hb) * (n1a + n1b - size) + hb * k * (2 * n2 - size) - hb * h * (2 * n0 - size);
}

void set_bz(tm_op *op)
{
  OpConsts<double> c;
  consts_bz(c, op->p.betaJ, op->p.kappa, op->p.kappa2);
  for (int i = 0; i < 2; ++i) {
    op->bzwj1[i] = c.b1[i];
    op->bzwj2[i] = c.b2[i];
    op->bzwj3[i] = c.b3[i];
  }
}

void init_device()
{
  static bool done = false;
  if (done) return;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    throw Error(TM_EHIP, "no HIP device available: the MI355X kernels cannot run");
  }
  int dev = 0;
  TM_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  TM_HIP(hipGetDeviceProperties(&prop, dev));
  g_lds_per_block = prop.sharedMemPerBlock;
  g_num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  done = true;
}

// Reduced 2D/3D operators: orbit tables + Tvec on the device.  With
// diag_in_full the full-space operator applied between expansion and gather
// carries the layer diagonal and the normfactor itself (XDM, ZBM), so
// Tvec = 1/sqrt(dup) and dup*Tvec = sqrt(dup).
void build_reduced(tm_op *op, const GroupSpec &g, double nfexp, bool want_fused,
                   bool diag_in_full = false)
{
  const bool fused = want_fused && !g.zgroup && g.dim == 2 && g.bc == 0 && (g.L % 2) == 0 &&
                     g.L >= 8 &&
                     // FP64 column + the pass-1 bases (at most 1182 words at L = 28)
                     ((size_t(1) << (g.L / 2)) * sizeof(double) + 8192 <= g_lds_per_block);
  const int64_t M = enumerate_orbits(g, op->reps, op->dups, fused ? nullptr : &op->map, op->stream);
  op->ncol = M;
  std::vector<uint64_t> reps(M);
  std::vector<uint16_t> dups(M);
  copy_sync(reps.data(), op->reps.get(), M * 8, hipMemcpyDeviceToHost, op->stream);
  copy_sync(dups.data(), op->dups.get(), M * 2, hipMemcpyDeviceToHost, op->stream);
  std::vector<double> tv(M), dt(M);
  const double Jn = 0.5 * op->p.betaJ * nfexp;
  for (int64_t o = 0; o < M; ++o) {
    if (diag_in_full) {
      tv[o] = 1.0 / std::sqrt((double)dups[o]);
      dt[o] = std::sqrt((double)dups[o]);
      continue;
    }
    double e;
    if (g.dim == 2) e = energy_half_x2d(reps[o], g.L, g.bc, op->p.betaJ, op->p.kappa, op->p.h);
    else e = energy_half_x3d(reps[o], g.s1, g.s2, op->p.betaJ, op->p.kappa, op->p.h);
    tv[o] = std::exp(e + Jn) / std::sqrt((double)dups[o]);
    dt[o] = dups[o] * tv[o];
  }
  if (fused) {
    op->ft.reset(new FusedTables());
    build_fused_tables(g, op->reps, M, *op->ft, op->stream);
    // the fused kernels work in the internal (row) orbit order
    op->perm = op->ft->perm;
    std::vector<double> a(M), b(M);
    for (int64_t i = 0; i < M; ++i) {
      a[i] = tv[op->perm[i]];
      b[i] = dt[op->perm[i]];
    }
    tv.swap(a);
    dt.swap(b);
  }
  std::vector<float> tv32(tv.begin(), tv.end()), dt32(dt.begin(), dt.end());
  op->tvec.alloc(M * 8);
  op->dt64.alloc(M * 8);
  op->tvec32.alloc(M * 4);
  op->dt32.alloc(M * 4);
  copy_sync(op->tvec.get(), tv.data(), M * 8, hipMemcpyHostToDevice, op->stream);
  copy_sync(op->dt64.get(), dt.data(), M * 8, hipMemcpyHostToDevice, op->stream);
  copy_sync(op->tvec32.get(), tv32.data(), M * 4, hipMemcpyHostToDevice, op->stream);
  copy_sync(op->dt32.get(), dt32.data(), M * 4, hipMemcpyHostToDevice, op->stream);
}

}  // namespace

tm_op *make_op(const tm_params &p)
{
  init_device();
  TM_CHECK(p.model == 'A' || p.model == 'B' || p.model == 'D', TM_EINVAL,
           "invalid model (expected A, B or D)");
  TM_CHECK(p.dim == 2 || p.dim == 3, TM_EINVAL, "invalid dim (expected 2 or 3)");
  std::unique_ptr<tm_op> op(new tm_op());
  // one stream per operator (default priority, blocking): its applies and
  // solves are stream-ordered, and independent operators run concurrently.
  // Nothing in the library uses the legacy null stream, so its implicit
  // synchronisation never reaches a stream that another thread captures.
  TM_HIP(hipStreamCreate(&op->stream));
  op->p = p;
  set_bz(op.get());
  const int multieigs = p.neigs > 1 ? 2 : 0;  // TMat/pfunc.cpp:22
  const double bJ = p.betaJ, k = p.kappa, h = p.h;
  auto make_xat = [&](int L, int bc) {
    TM_CHECK(L >= 4 && L <= 34, TM_EINVAL, "XAT width must be in [4, 34]");
    op->kind = TM_KIND_XAT;
    op->L = L;
    op->bc = bc;
    op->nstate = op->ncol = int64_t(1) << L;
    op->diag = DiagSpec{bc, L, 0, 0, 0, 0, 0, 0};
    apply_consts(op.get(), consts_xat<double>(L, bc, bJ, k, h));  // TMat/tmat.cpp:39-56
  };
  auto make_xam = [&](int L, int ext) {
    TM_CHECK(L >= 4 && L <= 32, TM_EINVAL, "XAM width must be in [4, 32]");
    op->kind = TM_KIND_XAM2D;
    op->L = L;
    op->bc = ext & 1;
    op->censym = (h == 0 && (ext & 2) == 0);  // TMat/tmatreduce.cpp:341
    op->nleadbit = (L < 10 || op->bc) ? 1 : 2 + (int)std::log2((double)L);
    op->nstate = int64_t(1) << L;
    double e;
    op->nf = nf_x2d(L, bJ, k, &e);
    GroupSpec g{2, L, 0, 0, op->bc, op->censym, 0};
    build_reduced(op.get(), g, e, !(p.flags & TM_FLAG_GENERIC_REDUCED));
  };
  auto make_xdt = [&](int L, int bc) {
    TM_CHECK(L >= 4 && L <= 32, TM_EINVAL, "XDT width must be in [4, 32]");
    op->kind = TM_KIND_XDT;
    op->L = L;
    op->bc = bc;
    op->nstate = op->ncol = int64_t(1) << L;
    const int dk = bc == 0 ? DIAG_XDT_PBC : (bc == 5 ? DIAG_XDT_C5 : DIAG_XDT_C9);
    op->diag = DiagSpec{dk, L, 0, 0, 0, 0, 0, 0};
    apply_consts(op.get(), consts_xdt<double>(L, bc, bJ, k, h));  // TMat/tmat.cpp:322-341
  };
  auto make_zbt = [&](int L) {
    TM_CHECK(L >= 4 && L <= 30 && (L % 2) == 0, TM_EINVAL,
             "ZBT needs an even width in [4, 30] (TMat/tmat.cpp:620)");
    op->kind = TM_KIND_ZBT;
    op->symmetric = false;
    op->L = L;
    op->nstate = op->ncol = int64_t(1) << L;
    op->diag = DiagSpec{DIAG_ZBT, L, 0, 0, 0, 0, 0, 0};
    apply_consts(op.get(), consts_zbt<double>(L, bJ, k, p.kappa2, h));  // TMat/tmat.cpp:480-518
  };
  // Reduced operators over a full-space gate circuit: T_red = P^T T P with
  // P(s, o) = [s in o] / sqrt(|o|), the algebra of XDMMatVecProd2D::perform_op
  // (TMat/tmatreduce.cpp:1054-1228) and ZBMMatVecProd2D::perform_op_internal
  // (TMat/tmatreduce.cpp:1457-1746): the scatter of Tvec*x over every image
  // of an orbit, the transfer over the expanded layer, and the gather
  // y *= dup*Tvec at the representative.  Both are periodic only
  // (TMat/tmatreduce.cpp:1047-1049, 1446-1448).
  auto make_xdm = [&](int L, int ext) {
    TM_CHECK(L >= 4 && L <= 30, TM_EINVAL, "XDM width must be in [4, 30]");
    make_xdt(L, 0);  // the full-space DNNI circuit and its constants
    op->kind = TM_KIND_XDM;
    op->fullkind = TM_KIND_XDT;
    op->censym = (h == 0 && (ext & 2) == 0);  // TMat/tmatreduce.cpp:1028
    GroupSpec g{2, L, 0, 0, 0, op->censym, 0, 0};
    build_reduced(op.get(), g, 0.0, false, true);
  };
  auto make_zbm = [&](int L, int ext) {
    TM_CHECK(L >= 4 && L <= 30 && (L % 2) == 0, TM_EINVAL,
             "ZBM needs an even width in [4, 30] (translations by two sites)");
    make_zbt(L);
    op->kind = TM_KIND_ZBM;
    op->fullkind = TM_KIND_ZBT;
    op->censym = (h == 0 && (ext & 2) == 0);  // TMat/tmatreduce.cpp:1406
    GroupSpec g{2, L, 0, 0, 0, op->censym, 0, 1};
    build_reduced(op.get(), g, 0.0, false, true);
  };

  // Axial ANNNI transfer over two-layer states (4^L).  K2 = e^{-bJ k} on
  // equal spins, e^{bJ k} on opposite ones (the NNN coupling along the axis).
  auto make_axial = [&](int L, bool pairs) {
    TM_CHECK(L >= 2 && L <= 15, TM_EINVAL, "axial (Y) width must be in [2, 15] (4^L states)");
    op->kind = pairs ? TM_KIND_Y2AT : TM_KIND_YAT;
    op->symmetric = false;
    op->L = L;
    op->nstate = op->ncol = int64_t(1) << (2 * L);
    const double Ld = L;
    if (!pairs) {
      // TMat/tmat.cpp:187-205: nf; Tvec = genXATmat2D({bJ, 0, 0, h, L}) as
      // Tv(layer) K(layer, previous) (tmat.hpp:361-366, tmat.cpp:668-684)
      op->nf = k < 0.5 ? std::exp(Ld * bJ * (k - 2.0)) : std::exp(-Ld * bJ * (k + 1.0));
      std::vector<double> t(192, 0.0);
      for (int n = 0; n <= L; ++n) {
        t[n] = std::exp(-bJ * (2.0 * n - Ld)) * op->nf;  // NN bonds in the layer
        t[64 + n] = std::exp(-bJ * h * (2.0 * n - Ld));   // field
        t[128 + n] = std::exp(-bJ * (2.0 * n - Ld));      // NN to the previous layer
      }
      upload_tables(op.get(), t);
    } else {
      // TMat/tmat.cpp:246-263, Tv^(1/2) of tmat.hpp:396-416 with sqrt(nf)
      op->nf = k < 0.5 ? std::exp(2 * Ld * bJ * (k - 2.0)) : std::exp(-2 * Ld * bJ * (k + 1.0));
      const double J1 = -0.5 * bJ, hJ = -0.5 * bJ * h, sq = std::sqrt(op->nf);
      std::vector<double> t(192, 0.0), d(192, 1.0);
      for (int n = 0; n <= 2 * L; ++n) {
        t[n] = std::exp(J1 * (2.0 * n - 2 * Ld)) * sq;  // NN bonds in both layers
        t[128 + n] = std::exp(hJ * (2.0 * n - 2 * Ld));  // field on both layers
      }
      for (int n = 0; n <= L; ++n) {
        t[64 + n] = std::exp(J1 * (2.0 * n - Ld));  // NN between the two layers
        d[n] = std::exp(bJ * (Ld - 2.0 * n));       // old low layer to new high layer
      }
      upload_tables(op.get(), t);
      upload_tables(op.get(), d, true);
    }
  };

  // Reduced axial transfer over two-layer states (c current, p previous):
  // T_red = P^T T P with T(n, c; c, p) = W(n, c) K2(n, p) W(c, p) and P over
  // the orbits of rotating / reflecting (/ flipping) both layers together
  // (YAMMatVecProd2D, TMat/tmatreduce.cpp:661-949; YBM = layertype 1,
  // :1017-1021).  W carries half of the layer's bonds and field and of the
  // bonds to the previous layer, sqrt(nf) folded in (genTvec :717-745,
  // genTvec1 :748-780).
  auto make_yam = [&](int L, int ext, bool ybm) {
    TM_CHECK(L >= 4 && L <= 15, TM_EINVAL, "reduced axial (Y) width must be in [4, 15]");
    op->kind = ybm ? TM_KIND_YBM : TM_KIND_YAM;
    op->symmetric = false;
    op->L = L;
    op->nstate = int64_t(1) << (2 * L);
    op->censym = (h == 0.0 && (ext & 2) == 0);  // TMat/tmatreduce.cpp:664
    const double Ld = L, k2 = p.kappa2;
    double nfe;
    if (!ybm) nfe = k < 0.5 ? Ld * (k - 2.0) : -Ld * (k + 1.0);
    else if (k + k2 < 0.5) nfe = Ld * 2.0 * (k + k2 - 1.0);
    else if (k < 0.5 && k2 - k < 0.5) nfe = -Ld;
    else if (k2 < 2 * k) nfe = -Ld * 2.0 * k;
    else nfe = Ld * 2.0 * (k - k2);
    op->nf = std::exp(bJ * nfe);
    const double hb = 0.5 * bJ, sq = std::exp(hb * nfe);
    std::vector<double> t(ybm ? 320 : 192, 0.0);
    for (int n = 0; n <= L; ++n) {
      t[n] = std::exp(hb * (Ld - 2.0 * n)) * sq;      // NN bonds in the layer
      t[64 + n] = std::exp(hb * h * (Ld - 2.0 * n));   // field
      t[128 + n] = std::exp(hb * (Ld - 2.0 * n));      // bonds to the previous layer
      if (ybm) t[192 + n] = std::exp(-hb * k * (Ld - 2.0 * n));  // NNN in the layer
    }
    if (ybm)
      for (int n = 0; n <= 2 * L; ++n) t[256 + n] = std::exp(-hb * k2 * (2 * Ld - 2.0 * n));
    upload_tables(op.get(), t);
    GroupSpec g{2, 2 * L, 0, 0, 0, op->censym, 0, 0, L};
    build_reduced(op.get(), g, 0.0, false, true);
  };

  if (p.model == 'A' && p.dim == 2) {
    const int L = p.size1;
    TM_CHECK(L >= 4, TM_EINVAL, "Invalid size (<4)");  // TMat/read_input.cpp:86-90
    switch (p.axialdir) {
      case 1: make_xat(L, 0); break;
      case 9: make_xat(L, 1); break;
      case 17: make_xat(L, 5); break;
      case 18: make_xat(L, 9); break;
      case 19: make_xat(L, 13); break;
      case 20: make_xat(L, 17); break;
      case 5: make_xam(L, multieigs); break;
      case 13: make_xam(L, multieigs | 1); break;
      case 0: make_axial(L, true); break;   // TMat/pfunc.cpp:30-31
      case 2: make_axial(L, false); break;  // TMat/pfunc.cpp:35-36
      case 6: make_yam(L, multieigs, false); break;  // TMat/pfunc.cpp:41-42
      default:
        throw Error(TM_EINVAL, "Function yet implemented (unknown axialdir)");
    }
  } else if (p.model == 'A' && p.dim == 3) {
    TM_CHECK(p.axialdir == 1, TM_EINVAL, "3D supports axialdir 1 only (TMat/pfunc.cpp:63-75)");
    const int s1 = p.size1, s2 = p.size2, size = s1 * s2;
    TM_CHECK(s1 >= 2 && s2 >= 3 && size <= 32, TM_EINVAL, "3D layer must be s1>=2, s2>=3, s1*s2<=32");
    op->s1 = s1;
    op->s2 = s2;
    op->L = size;
    op->nstate = int64_t(1) << size;
    double e;
    op->nf = nf_x3d(size, bJ, k, &e);
    if (p.flags & TM_FLAG_FULLSPACE) {
      op->kind = TM_KIND_XAT3D;
      op->ncol = op->nstate;
      op->diag = DiagSpec{DIAG_X3D, size, s1, s2, colmask(s1, s2, 1), colmask(s1, s2, s2 - 1),
                          colmask(s1, s2, 2), colmask(s1, s2, s2 - 2)};
      apply_consts(op.get(), consts_xat3d<double>(size, bJ, k, h));
    } else {
      op->kind = TM_KIND_XAM3D;
      op->censym = (h == 0);                    // TMat/tmat3d.cpp:105
      op->rotate = (k == 0) && (s1 == s2);
      op->nleadbit = size < 10 ? 1 : 2 + (int)std::log2((double)size);
      GroupSpec g{3, size, s1, s2, 0, op->censym, op->rotate};
      build_reduced(op.get(), g, e, false);
    }
  } else if (p.model == 'B') {
    TM_CHECK(p.dim == 2, TM_EINVAL, "BNNNI is 2D only");
    switch (p.axialdir) {
      case 3: make_zbt(p.size1); break;
      case 7: make_zbm(p.size1, multieigs); break;
      case 6: make_yam(p.size1, multieigs, true); break;  // TMat/pfunc.cpp:81
      default:
        throw Error(TM_EINVAL, "Function yet implemented (unknown axialdir)");
    }
  } else {  // 'D'
    TM_CHECK(p.dim == 2, TM_EINVAL, "DNNI is 2D only");
    switch (p.axialdir) {
      case 1: case 2: make_xdt(p.size1, 0); break;
      case 17: make_xdt(p.size1, 5); break;
      case 18: make_xdt(p.size1, 9); break;
      case 5: case 6: make_xdm(p.size1, multieigs); break;
      default:
        throw Error(TM_EINVAL, "Function yet implemented (unknown axialdir)");
    }
  }
  op->one.alloc(sizeof(double));
  const double one = 1.0;
  copy_sync(op->one.get(), &one, sizeof one, hipMemcpyHostToDevice, op->stream);
  return op.release();
}

}  // namespace ltm

tm_op::~tm_op()
{
  if (stream) (void)hipStreamDestroy(stream);
}

void tm_op::apply_dev(const void *x, void *y, int64_t ld, int nvec, tm_dtype io, tm_dtype comp,
                      int transpose, const double *scale_in, hipStream_t s,
                      const ltm::Workspace *ws)
{
  using namespace ltm;
  if (perm.empty()) {
    apply_internal(x, y, ld, nvec, io, comp, transpose, scale_in, s, ws);
    return;
  }
  const size_t es = io == TM_F64 ? 8 : 4;
  void *bx, *by;
  if (ws) {
    TM_CHECK(ws->px && ws->py && ws->perm_bytes >= perm_bytes(io), TM_EINVAL,
             "workspace too small for the permutation buffers");
    bx = ws->px;
    by = ws->py;
  } else {
    px.ensure(perm_bytes(io));
    py.ensure(perm_bytes(io));
    bx = px.get();
    by = py.get();
  }
  const uint32_t *pd = ft->olist.as<uint32_t>();
  for (int v = 0; v < nvec; ++v) {
    const void *xv = static_cast<const char *>(x) + (size_t)v * ld * es;
    void *yv = static_cast<char *>(y) + (size_t)v * ld * es;
    permute_to_internal(xv, bx, pd, ncol, io, s);
    apply_internal(bx, by, ncol, 1, io, comp, transpose, scale_in ? scale_in + v : nullptr, s, ws);
    permute_to_external(by, yv, pd, ncol, io, s);
  }
}

size_t tm_op::perm_bytes(tm_dtype io) const
{
  return perm.empty() ? 0 : (io == TM_F64 ? 8 : 4) * (size_t)ncol;
}

ltm::SweepPlan tm_op::sweep_plan() const
{
  ltm::SweepPlan sp{};
  sp.kind = fullkind ? fullkind : kind;
  sp.L = L;
  sp.bc = bc;
  sp.diag = diag;
  sp.tab64 = tab64.as<double>();
  sp.tab32 = tab32.as<float>();
  for (int i = 0; i < 2; ++i) {
    sp.b1[i] = bzwj1[i];
    sp.b2[i] = bzwj2[i];
    sp.b3[i] = bzwj3[i];
  }
  sp.nf = nf;
  sp.tune = !(p.flags & TM_FLAG_NO_AUTOTUNE);
  return sp;
}

namespace {
// The reduced-operator plan of op (sp receives the circuit plan of XDM / ZBM).
ltm::ReducedPlan reduced_plan(const tm_op &op, int transpose, ltm::SweepPlan &sp)
{
  using namespace ltm;
  ReducedPlan rp{};
  rp.M = op.ncol;
  rp.L = op.L;
  rp.tvec = op.tvec.as<double>();
  rp.dt64 = op.dt64.as<double>();
  rp.tvec32 = op.tvec32.as<float>();
  rp.dt32 = op.dt32.as<float>();
  rp.reps = op.reps.as<uint64_t>();
  rp.map = op.map.as<uint32_t>();
  rp.ft = op.ft.get();
  rp.ka = op.bzwj1[0];
  rp.kb = op.bzwj1[1];
  if (op.kind == TM_KIND_XDM || op.kind == TM_KIND_ZBM) {
    sp = op.sweep_plan();
    rp.sweep = &sp;
    rp.transpose = op.kind == TM_KIND_ZBM ? transpose : 0;
  }
  if (op.kind == TM_KIND_YAM || op.kind == TM_KIND_YBM) {
    rp.L = 2 * op.L;
    rp.ylayer = op.L;
    rp.ybm = op.kind == TM_KIND_YBM;
    rp.ka = op.bzwj2[0];  // K2 on the previous layer (permute_mult_TL)
    rp.kb = op.bzwj2[1];
    rp.wtab64 = op.tab64.as<double>();
    rp.wtab32 = op.tab32.as<float>();
    rp.transpose = transpose;
  }
  return rp;
}
}  // namespace

size_t tm_op::work_bytes(tm_dtype comp) const
{
  using namespace ltm;
  switch (kind) {
    case TM_KIND_XAT:
    case TM_KIND_XAT3D: return kron_work_bytes(L, comp);
    case TM_KIND_XDT:
    case TM_KIND_ZBT: return sweep_work_bytes(kind, L, comp);
    case TM_KIND_YAT:
    case TM_KIND_Y2AT: return axial_work_bytes(L, comp);
    default: {
      SweepPlan sp{};
      return reduced_work_bytes(reduced_plan(*this, 0, sp), comp);
    }
  }
}

void tm_op::apply_internal(const void *x, void *y, int64_t ld, int nvec, tm_dtype io,
                           tm_dtype comp, int transpose, const double *scale_in, hipStream_t s,
                           const ltm::Workspace *ws)
{
  using namespace ltm;
  const size_t es = io == TM_F64 ? 8 : 4;
  const bool tune = !(p.flags & TM_FLAG_NO_AUTOTUNE) && !ws;
  // the workspace: the caller's (checked) or the operator's own
  const size_t need = work_bytes(comp);
  void *wk;
  if (ws) {
    TM_CHECK(ws->work_bytes >= need && (need == 0 || ws->work), TM_EINVAL,
             "workspace too small for this operator");
    wk = ws->work;
  } else {
    work.ensure(need);
    wk = work.get();
  }
  for (int v = 0; v < nvec; ++v) {
    const void *xv = static_cast<const char *>(x) + (size_t)v * ld * es;
    void *yv = static_cast<char *>(y) + (size_t)v * ld * es;
    switch (kind) {
      case TM_KIND_XAT:
      case TM_KIND_XAT3D: {
        KronPlan kp{};
        kp.L = L;
        kp.prologue = kp.epilogue = true;
        kp.diag = diag;
        kp.tab64 = tab64.as<double>();
        kp.tab32 = tab32.as<float>();
        kp.ka = bzwj1[0];
        kp.kb = bzwj1[1];
        kp.tune = tune;
        launch_kron(kp, xv, yv, io, io, comp, wk, scale_in, s);
        break;
      }
      case TM_KIND_XDT:
      case TM_KIND_ZBT: {
        SweepPlan sp = sweep_plan();
        sp.tune = tune;
        launch_sweep(sp, xv, yv, io, comp, kind == TM_KIND_ZBT ? transpose : 0, wk, scale_in, s);
        break;
      }
      case TM_KIND_YAT:
      case TM_KIND_Y2AT: {
        AxialPlan ap{};
        ap.kind = kind;
        ap.L = L;
        ap.ka = std::exp(-p.betaJ * p.kappa);
        ap.kb = std::exp(p.betaJ * p.kappa);
        ap.tab64 = tab64.as<double>();
        ap.tab32 = tab32.as<float>();
        ap.tabb64 = tabb64.as<double>();
        ap.tabb32 = tabb32.as<float>();
        ap.tune = tune;
        launch_axial(ap, xv, yv, io, comp, transpose, wk, scale_in, s);
        break;
      }
      case TM_KIND_XAM2D:
      case TM_KIND_XAM3D:
      case TM_KIND_XDM:
      case TM_KIND_ZBM:
      case TM_KIND_YAM:
      case TM_KIND_YBM: {
        SweepPlan sp{};
        ReducedPlan rp = reduced_plan(*this, transpose, sp);
        sp.tune = tune;
        if (!rp.ft) TM_CHECK(rp.map != nullptr, TM_EINVAL, "reduced operator has no state map");
        launch_reduced(rp, xv, yv, io, comp, wk, scale_in, s);
        break;
      }
      default:
        throw Error(TM_EINVAL, "operator kind not applicable");
    }
    ++matvecs;
  }
}
