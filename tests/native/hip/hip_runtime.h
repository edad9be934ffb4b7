// Host-only stand-in for the handful of HIP runtime calls made by the
// eigen driver (tm_solver.cpp) and the dense eigensolvers, so that the
// driver logic can be unit-tested with g++ on a machine without a GPU.
// Memory is plain host memory; "device" kernels are provided by
// fake_backend.cpp.  Test infrastructure only: never linked into
// liblatttmat.so.
#pragma once
#include <cstddef>
#include <cstring>
#include <cstdlib>
typedef int hipError_t;
typedef void *hipStream_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) { std::memmove(d, s, n); return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipGetLastError() { return 0; }
inline const char *hipGetErrorString(hipError_t) { return "fake"; }
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
// graphs: capture is refused, so the eigensolver stays eager
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum { hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return 1; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) { *g = nullptr; return 1; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t, void *, void *, size_t) { *e = nullptr; return 1; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 1; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return 0; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 0; }
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) { *p = std::malloc(n); return *p ? 0 : 2; }
inline hipError_t hipHostFree(void *p) { std::free(p); return 0; }
