// Deferred-execution stand-in of the HIP runtime for the CPU tests (see
// hip/hip_runtime.h for the model).  Test infrastructure only.
#include <cstdio>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

struct FakeGraph {
  std::vector<std::function<void()>> nodes;
};
struct FakeStream {
  std::deque<std::function<void()>> q;
  FakeGraph *cap = nullptr;
  std::thread::id cap_thread;
  bool cap_broken = false;
};
// one recording of an event: fired when the stream reaches it
struct FakeMark {
  FakeStream *s = nullptr;
  bool fired = false;
  long stamp = 0;
};
struct FakeEvent {
  std::shared_ptr<FakeMark> mark;
};

namespace {
std::recursive_mutex g_mu;
FakeStream g_null;
long g_clock = 0, g_launches = 0;
std::set<FakeStream *> g_streams{&g_null};

struct PendingRead {
  const char *src;
  size_t n;
  std::vector<char> snap;
  bool superseded = false;
};
std::list<std::shared_ptr<PendingRead>> g_reads;

[[noreturn]] void die(const char *what)
{
  std::fprintf(stderr, "fake HIP runtime: %s\n", what);
  std::fflush(stderr);
  std::abort();
}

FakeStream *S(hipStream_t s) { return s ? s : &g_null; }

bool capturing_thread()
{
  for (FakeStream *s : g_streams)
    if (s->cap && s->cap_thread == std::this_thread::get_id()) return true;
  return false;
}

void push(FakeStream *s, std::function<void()> fn)
{
  if (s->cap) s->cap->nodes.push_back(std::move(fn));
  else s->q.push_back(std::move(fn));
}

void run_one(FakeStream *s)
{
  auto fn = std::move(s->q.front());
  s->q.pop_front();
  fn();
  ++g_clock;
}

// Run s to empty.  The other streams progress at the same time (one of
// their operations after each of s's), as concurrent device queues would:
// work that is not ordered against s through an event wait interleaves
// with it.
void flush(FakeStream *s)
{
  while (!s->q.empty()) {
    run_one(s);
    for (FakeStream *o : g_streams)
      if (o != s && !o->q.empty()) run_one(o);
  }
}

// run m's stream until the recording m has fired
void flush_until(const std::shared_ptr<FakeMark> &m)
{
  while (m && !m->fired && m->s && !m->s->q.empty()) {
    auto fn = std::move(m->s->q.front());
    m->s->q.pop_front();
    fn();
    ++g_clock;
  }
}

void flush_all()
{
  bool any = true;
  while (any) {
    any = false;
    for (FakeStream *s : g_streams)
      if (!s->q.empty()) {
        flush(s);
        any = true;
      }
  }
}

bool overlaps(const char *a, size_t na, const char *b, size_t nb)
{
  return a < b + nb && b < a + na;
}
}  // namespace

hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "fake HIP error"; }

hipError_t hipMalloc(void **p, size_t n)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (capturing_thread()) return hipErrorStreamCaptureUnsupported;
  *p = std::malloc(n ? n : 1);
  if (!*p) return hipErrorOutOfMemory;
  std::memset(*p, 0xFF, n);  // uninitialised device memory reads as NaN
  return hipSuccess;
}

hipError_t hipFree(void *p)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (capturing_thread()) return hipErrorStreamCaptureUnsupported;
  flush_all();  // hipFree synchronises the device
  std::free(p);
  return hipSuccess;
}

hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (capturing_thread()) return hipErrorStreamCaptureUnsupported;
  *p = std::malloc(n ? n : 1);
  if (!*p) return hipErrorOutOfMemory;
  std::memset(*p, 0xFF, n);
  return hipSuccess;
}

hipError_t hipHostFree(void *p)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (capturing_thread()) return hipErrorStreamCaptureUnsupported;
  flush_all();
  std::free(p);
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  FakeStream *fs = S(st);
  char *dc = static_cast<char *>(d);
  const char *sc = static_cast<const char *>(s);
  if (fs->cap) {  // a copy node reads its source at replay
    push(fs, [=] { std::memmove(dc, sc, n); });
    return hipSuccess;
  }
  if (k == hipMemcpyHostToDevice) {
    auto pr = std::make_shared<PendingRead>();
    pr->src = sc;
    pr->n = n;
    pr->snap.assign(sc, sc + n);
    g_reads.push_back(pr);
    push(fs, [=] {
      if (!pr->superseded && std::memcmp(pr->src, pr->snap.data(), n) != 0)
        die("the host changed the source of a queued host-to-device copy before it ran");
      std::memcpy(dc, pr->snap.data(), n);
      g_reads.remove(pr);
    });
  } else if (k == hipMemcpyDeviceToHost) {
    // a queued read of this host range still runs first (stream order)
    for (auto &pr : g_reads)
      if (overlaps(pr->src, pr->n, dc, n)) pr->superseded = true;
    std::memset(dc, 0xFF, n);  // undefined until the host waits
    push(fs, [=] { std::memmove(dc, sc, n); });
  } else {
    push(fs, [=] { std::memmove(dc, sc, n); });
  }
  return hipSuccess;
}

hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k)
{
  hipError_t e = hipMemcpyAsync(d, s, n, k, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return e;
}

hipError_t hipStreamCreate(hipStream_t *s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  *s = new FakeStream();
  g_streams.insert(*s);
  return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!s) return hipErrorInvalidValue;
  flush(s);
  g_streams.erase(s);
  delete s;
  return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  FakeStream *fs = S(s);
  if (fs->cap) {
    fs->cap_broken = true;
    return hipErrorStreamCaptureUnsupported;
  }
  flush(fs);
  return hipSuccess;
}

// a query is a host wait in this model: the queued work runs, then the
// stream reports complete
hipError_t hipStreamQuery(hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  FakeStream *fs = S(s);
  if (fs->cap) {
    fs->cap_broken = true;
    return hipErrorStreamCaptureUnsupported;
  }
  flush(fs);
  return hipSuccess;
}

hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!e) return hipErrorInvalidValue;
  std::shared_ptr<FakeMark> m = e->mark;  // the recording current now
  if (!m) return hipSuccess;               // never recorded: nothing to wait for
  push(S(s), [m] { flush_until(m); });
  return hipSuccess;
}

hipError_t hipDeviceSynchronize()
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  flush_all();
  return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }

hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  *e = new FakeEvent();
  return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!e) return hipErrorInvalidValue;
  auto m = std::make_shared<FakeMark>();
  m->s = S(s);
  e->mark = m;
  push(S(s), [m] {
    m->fired = true;
    m->stamp = g_clock;
  });
  return hipSuccess;
}

hipError_t hipEventSynchronize(hipEvent_t e)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!e) return hipErrorInvalidValue;
  flush_until(e->mark);
  return hipSuccess;
}

hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!a->mark || !b->mark || !a->mark->fired || !b->mark->fired) return hipErrorNotReady;
  *ms = 0.001f * (float)(b->mark->stamp - a->mark->stamp);
  return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t e)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  delete e;  // queued recordings and waits hold their marks
  return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, int)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  FakeStream *fs = S(s);
  if (fs->cap || fs == &g_null) return hipErrorInvalidValue;
  flush(fs);  // earlier work is not part of the graph
  fs->cap = new FakeGraph();
  fs->cap_thread = std::this_thread::get_id();
  fs->cap_broken = false;
  return hipSuccess;
}

hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  FakeStream *fs = S(s);
  if (!fs->cap) return hipErrorInvalidValue;
  FakeGraph *gr = fs->cap;
  fs->cap = nullptr;
  if (fs->cap_broken) {
    delete gr;
    *g = nullptr;
    return hipErrorStreamCaptureUnsupported;
  }
  *g = gr;
  return hipSuccess;
}

hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!g) return hipErrorInvalidValue;
  *e = new FakeGraph(*g);
  return hipSuccess;
}

hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!e) return hipErrorInvalidValue;
  for (auto &fn : e->nodes) push(S(s), fn);
  ++g_launches;
  return hipSuccess;
}

hipError_t hipGraphDestroy(hipGraph_t g)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  delete g;
  return hipSuccess;
}

hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  delete e;
  return hipSuccess;
}

void fake_enqueue(hipStream_t s, std::function<void()> fn)
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  push(S(s), std::move(fn));
}

size_t fake_pending()
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  size_t n = 0;
  for (FakeStream *s : g_streams) n += s->q.size();
  return n;
}

long fake_graph_launches()
{
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return g_launches;
}
