// tm_runtime.cpp -- the library's thin layer over the HIP runtime: device
// and pinned buffers, stream-ordered copies, cross-stream ordering, the
// graph-capture / allocation lock and the tile autotuner.  It uses only
// host-side HIP calls, so the CPU test harness (tests/native) links this
// same file against a deferred-execution stand-in of the runtime.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "tm_internal.hpp"

namespace ltm {

size_t g_lds_per_block = 64 * 1024;
int g_num_cu = 256;

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line)
{
  char buf[512];
  std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
                hipGetErrorString(e), expr, file, line);
  throw Error(e == hipErrorOutOfMemory ? TM_ENOMEM : TM_EHIP, buf);
}

namespace {
std::shared_mutex g_capture_mu;
thread_local bool t_capturing = false;
}  // namespace

CaptureGuard::CaptureGuard()
{
  g_capture_mu.lock();
  t_capturing = true;
}
CaptureGuard::~CaptureGuard()
{
  t_capturing = false;
  g_capture_mu.unlock();
}
bool capturing_here() { return t_capturing; }

namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, uint32_t, int, int>, int> g_tune;
}  // namespace

int autotune_tile(TuneFamily fam, uint32_t id, int L, int es, const std::vector<int> &cands,
                  bool tune, hipStream_t s, const std::function<void(int)> &run)
{
  if (!tune || cands.size() < 2) return cands[0];
  const auto key = std::make_tuple((int)fam, id, L, es);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  if (t_capturing) return cands[0];
  hipEvent_t e0, e1;
  TM_HIP(hipEventCreate(&e0));
  TM_HIP(hipEventCreate(&e1));
  int best = cands[0];
  float best_ms = std::numeric_limits<float>::max();
  for (int t : cands) {
    run(t);
    TM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; ++r) run(t);
    TM_HIP(hipEventRecord(e1, s));
    TM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TM_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tune.emplace(key, best).first->second;
}

int autotune_cached(TuneFamily fam, uint32_t id, int L, int es)
{
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(std::make_tuple((int)fam, id, L, es));
  return it == g_tune.end() ? -1 : it->second;
}

void destroy_graph_exec(hipGraphExec_t g)
{
  if (!g) return;
  std::shared_lock<std::shared_mutex> lk(g_capture_mu);
  (void)hipGraphExecDestroy(g);
}

void stream_sync_poll(hipStream_t s)
{
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) TM_HIP(e);
    std::this_thread::yield();
  }
}

void copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
  TM_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
  TM_HIP(hipStreamSynchronize(s));
}

void stream_wait(hipStream_t waiter, hipStream_t producer)
{
  if (waiter == producer) return;
  hipEvent_t e;
  TM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t r = hipEventRecord(e, producer);
  if (r == hipSuccess) r = hipStreamWaitEvent(waiter, e, 0);
  (void)hipEventDestroy(e);  // the wait keeps the recorded point alive
  if (r != hipSuccess) throw_hip(r, "stream_wait", __FILE__, __LINE__);
}

namespace {
// Double-buffered pinned staging for copy_large, one per host thread (the
// batch driver copies from several threads at once).
struct Stage {
  PinnedBuf buf[2];
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Stage()
  {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};
thread_local Stage t_stage;

// host memcpy split over up to four threads (one host thread moves about
// 10 GB/s, below what the DMA engines deliver into pinned memory)
void host_copy(void *d, const void *src, size_t n)
{
  const int nt = n >= (size_t(4) << 20) ? 4 : 1;
  const size_t part = ((n + nt - 1) / nt + 63) & ~size_t(63);
  std::thread th[3];
  for (int t = 1; t < nt; ++t) {
    const size_t o = part * t;
    if (o >= n) break;
    th[t - 1] = std::thread([=] {
      std::memcpy(static_cast<char *>(d) + o, static_cast<const char *>(src) + o,
                  std::min(part, n - o));
    });
  }
  std::memcpy(d, src, std::min(part, n));
  for (auto &x : th)
    if (x.joinable()) x.join();
}
}  // namespace

void copy_large(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
  constexpr size_t CH = size_t(8) << 20;
  if (bytes < (size_t(2) << 20) ||
      (kind != hipMemcpyDeviceToHost && kind != hipMemcpyHostToDevice)) {
    copy_sync(dst, src, bytes, kind, s);
    return;
  }
  Stage &st = t_stage;
  for (int b = 0; b < 2; ++b) {
    st.buf[b].ensure(CH);
    if (!st.ev[b]) TM_HIP(hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming));
  }
  const size_t nch = (bytes + CH - 1) / CH;
  char *d = static_cast<char *>(dst);
  const char *sp = static_cast<const char *>(src);
  if (kind == hipMemcpyDeviceToHost) {
    // DMA of chunk i into one buffer while the host copies chunk i-1 out of
    // the other
    for (size_t i = 0; i <= nch; ++i) {
      if (i < nch) {
        const size_t off = i * CH, n = std::min(CH, bytes - off);
        TM_HIP(hipMemcpyAsync(st.buf[i & 1].as<char>(), sp + off, n, kind, s));
        TM_HIP(hipEventRecord(st.ev[i & 1], s));
      }
      if (i > 0) {
        const size_t j = i - 1, off = j * CH, n = std::min(CH, bytes - off);
        TM_HIP(hipEventSynchronize(st.ev[j & 1]));
        host_copy(d + off, st.buf[j & 1].as<char>(), n);
      }
    }
  } else {
    // the host fills one buffer while the DMA of the previous chunk runs; a
    // buffer is refilled once the DMA that read it has finished
    for (size_t i = 0; i < nch; ++i) {
      const size_t off = i * CH, n = std::min(CH, bytes - off);
      if (i >= 2) TM_HIP(hipEventSynchronize(st.ev[i & 1]));
      host_copy(st.buf[i & 1].as<char>(), sp + off, n);
      TM_HIP(hipMemcpyAsync(d + off, st.buf[i & 1].as<char>(), n, kind, s));
      TM_HIP(hipEventRecord(st.ev[i & 1], s));
    }
    TM_HIP(hipStreamSynchronize(s));
  }
}

// Allocation and release refuse to run on a thread that is capturing a
// graph, before touching the old buffer: a free inside a capture would be
// recorded nowhere and would leave the captured launches holding a dead
// address, so the buffer must stay as it is.  On other threads they hold the
// shared side of the capture lock, so no capture is in progress anywhere
// while memory is allocated or freed (hipFree synchronises the device).
void PinnedBuf::alloc(size_t bytes)
{
  if (t_capturing) throw Error(TM_EHIP, "host allocation during graph capture");
  release();
  if (bytes == 0) return;
  std::shared_lock<std::shared_mutex> lk(g_capture_mu);
  if (hipHostMalloc(&p_, bytes, hipHostMallocDefault) != hipSuccess) {
    p_ = nullptr;
    (void)hipGetLastError();
    throw Error(TM_ENOMEM, "pinned host allocation failed");
  }
  n_ = bytes;
}

void PinnedBuf::release()
{
  if (!p_) return;
  if (t_capturing) throw Error(TM_EHIP, "host free during graph capture");
  {
    std::shared_lock<std::shared_mutex> lk(g_capture_mu);
    (void)hipHostFree(p_);
  }
  p_ = nullptr;
  n_ = 0;
}

void DevBuf::alloc(size_t bytes)
{
  if (t_capturing) throw Error(TM_EHIP, "device allocation during graph capture");
  release();
  if (bytes == 0) return;
  std::shared_lock<std::shared_mutex> lk(g_capture_mu);
  hipError_t e = hipMalloc(&p_, bytes);
  if (e != hipSuccess) {
    p_ = nullptr;
    char buf[128];
    std::snprintf(buf, sizeof buf, "device allocation of %zu bytes failed", bytes);
    (void)hipGetLastError();
    throw Error(TM_ENOMEM, buf);
  }
  n_ = bytes;
}

void DevBuf::release()
{
  if (!p_) return;
  if (t_capturing) throw Error(TM_EHIP, "device free during graph capture");
  {
    std::shared_lock<std::shared_mutex> lk(g_capture_mu);
    (void)hipFree(p_);
  }
  p_ = nullptr;
  n_ = 0;
}

}  // namespace ltm
