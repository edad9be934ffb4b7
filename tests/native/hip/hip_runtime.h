// Host-only stand-in for the HIP runtime calls made by the library's host
// code (tm_runtime.cpp, tm_solver.cpp, tm_api.cpp, tm_dense.cpp), so that
// the eigen driver, the buffer and copy layer and the C ABI can be tested
// with g++ (and its sanitizers) on a machine without a GPU.  Test
// infrastructure only: never linked into liblatttmat.so.
//
// The stand-in is a deferred-execution model (fake_runtime.cpp): kernels
// (the host functions of fake_backend.cpp) and asynchronous copies are
// queued per stream and run only when the host waits (stream or event
// synchronise, a free, a device synchronise), in stream order, one stream at
// a time; a stream waits for another only through hipStreamWaitEvent.  A
// device-to-host copy fills its host destination with NaN bytes when it is
// queued, so a host read that comes before the wait sees garbage; a
// host-to-device copy remembers its source when it is queued and aborts if
// the host changed it before the copy ran.  Graph capture records the
// queued closures, and a launch replays them with the pointers they were
// captured with.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

typedef int hipError_t;
struct FakeStream;
struct FakeEvent;
struct FakeGraph;
typedef FakeStream *hipStream_t;
typedef FakeEvent *hipEvent_t;
typedef FakeGraph *hipGraph_t;
typedef FakeGraph *hipGraphExec_t;

enum {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorStreamCaptureUnsupported = 900,
  hipErrorNotReady = 600
};
enum hipMemcpyKind {
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3
};
enum { hipEventDefault = 0, hipEventDisableTiming = 2 };
enum { hipHostMallocDefault = 0 };
enum { hipStreamCaptureModeThreadLocal = 1 };

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamQuery(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipStreamBeginCapture(hipStream_t s, int mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

// ---- model controls (tests only) ----
// Queue a "kernel" on stream s (the null stream if s is null).
void fake_enqueue(hipStream_t s, std::function<void()> fn);
// Number of queued but not yet executed operations over all streams.
size_t fake_pending();
// Graphs launched so far (replays).
long fake_graph_launches();
