// tm_noalloc.hip -- tm_op_apply_dev_noalloc: the stream-ordered apply of
// tm_op_apply_dev for callers that cannot allocate device memory (code under
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every buffer the apply needs beyond the caller's inputs and outputs -- the
// partial sums and per-vector scales of the FP32 input scaling, the
// permutation buffers of the reduced operators' internal order, and the
// operator workspace -- is carved out of one file-scope device array of
// SCRATCH_BYTES.  Each GPU holds its own copy of that array until the process
// ends; its device address is looked up once per device (hipGetSymbolAddress
// on the entry point's first call with that device current) and only that
// address is handed to kernels.  Calls on one GPU share the array, so calls
// that overlap on one GPU race; nothing in it needs to start at zero (every
// kernel writes its part of the scratch before reading it).
#include <cstdio>
#include <mutex>

#include "tm_internal.hpp"

namespace ltm {
namespace {

constexpr size_t SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this GPU's scratch array (looked up on first use).
void *scratch_address()
{
  int dev = 0;
  TM_HIP(hipGetDevice(&dev));
  TM_CHECK(dev >= 0 && dev < MAX_DEVICES, TM_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    TM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TM_CHECK(p != nullptr, TM_EHIP, "scratch array has no device address");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// The body of tm_op_apply_dev_noalloc (tm_api.cpp maps its exceptions to a
// status): the same kernels as tm_op_apply_dev, on scratch memory.
void apply_dev_noalloc(tm_op *op, const void *x, void *y, int64_t ld, int nvec, tm_dtype dtype,
                       tm_dtype compute, int transpose, hipStream_t s)
{
  const bool scaled = compute == TM_F32 && dtype == TM_F64;
  // scratch layout: [partial sums | scales | px | py | workspace]
  const size_t b_part = scaled ? round256(sizeof(double) * kry::NBLK) : 0;
  const size_t b_scale = scaled ? round256(sizeof(double) * (size_t)nvec) : 0;
  const size_t b_perm = round256(op->perm_bytes(dtype));
  const size_t b_work = round256(op->work_bytes(compute));
  const size_t total = b_part + b_scale + 2 * b_perm + b_work;
  if (total > SCRATCH_BYTES) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "this apply needs %zu bytes of scratch; the no-allocation entry point holds at "
                  "most %zu (use tm_op_apply_dev)",
                  total, SCRATCH_BYTES);
    throw Error(TM_EINVAL, msg);
  }
  char *base = static_cast<char *>(scratch_address());
  double *partial = reinterpret_cast<double *>(base);
  double *scale = reinterpret_cast<double *>(base + b_part);
  Workspace ws;
  ws.px = base + b_part + b_scale;
  ws.py = base + b_part + b_scale + b_perm;
  ws.perm_bytes = b_perm;
  ws.work = base + b_part + b_scale + 2 * b_perm;
  ws.work_bytes = b_work;
  const size_t es = dtype == TM_F64 ? 8 : 4;
  if (scaled) {
    // per-vector power-of-two scale so that max|x| is in [1/2, 1)
    for (int v = 0; v < nvec; ++v) {
      kry::pow2_scale(static_cast<const double *>(x) + (size_t)v * ld, op->ncol, partial,
                      scale + v, s);
      op->apply_dev(static_cast<const char *>(x) + (size_t)v * ld * es,
                    static_cast<char *>(y) + (size_t)v * ld * es, ld, 1, dtype, compute,
                    transpose, scale + v, s, &ws);
    }
  } else {
    op->apply_dev(x, y, ld, nvec, dtype, compute, transpose, nullptr, s, &ws);
  }
}

}  // namespace ltm
