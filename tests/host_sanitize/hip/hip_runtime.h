/*
 * Host-memory stand-in for the few HIP runtime calls of the library's host set-up code
 * (dimspers_amd/csrc/setup.cpp and the buffers of stk_internal.h), so that code can be built with
 * g++ under AddressSanitizer and UndefinedBehaviorSanitizer and run without a GPU.  "Device"
 * buffers are malloc'ed host memory and copies are memcpy; streams and events are empty handles.
 * Only tests/host_sanitize/setup_driver.cpp is built against it.
 */
#pragma once
#include <cstddef>
#include <cstdlib>
#include <cstring>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2 };
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 };

inline const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "out of memory"; }
template <typename T>
inline hipError_t hipMalloc(T **p, size_t n) {
    *p = static_cast<T *>(std::malloc(n));
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = nullptr; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
