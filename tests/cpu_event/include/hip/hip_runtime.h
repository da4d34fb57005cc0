// Host-only stand-in for the HIP runtime header, used to compile the event
// path's device code (gegs_amd/csrc/event.hpp and what it includes) as plain
// C++ for a CPU check of its queue logic.  Every thread runs alone, one after
// another, as a wavefront with a single active lane (lane = global thread
// index mod 64): ballots see only that lane, shuffles from other lanes read
// zero, atomics are plain read-modify-writes.  A wavefront with one active
// lane is a legal GPU execution state, so the kernels must be correct under
// it; what it cannot show are interactions between lanes of one wavefront.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define GEGS_WAVES_PER_EU(n)
#define __shared__

struct dim3 {
    uint32_t x = 1, y = 1, z = 1;
    dim3() {}
    dim3(uint32_t a, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
inline dim3 threadIdx, blockIdx, blockDim, gridDim;
struct float4 {
    float x, y, z, w;
};

inline uint32_t cpu_lane() { return (blockIdx.x * blockDim.x + threadIdx.x) & 63u; }
inline uint32_t __lane_id() { return cpu_lane(); }
inline uint64_t __ballot(int p) { return p ? (1ull << cpu_lane()) : 0ull; }
inline int __any(int p) { return p != 0; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_readlane(x, l) (x)
inline uint32_t __shfl_xor(uint32_t, int, int = 64) { return 0u; }
inline void __syncthreads() {}
inline float __uint_as_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
inline uint32_t __float_as_uint(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
template <class T, class U> inline T atomicAdd(T *p, U v) {
    const T o = *p;
    *p = (T)(o + (T)v);
    return o;
}
inline double unsafeAtomicAdd(double *p, double v) {
    const double o = *p;
    *p = o + v;
    return o;
}

// runtime entry points the host parts of event.hpp name; the CPU driver
// launches the sweep kernels itself and never reaches them
enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorLaunchTimeOut = 702 };
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum {
    hipEventDisableTiming = 2,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamNonBlocking = 1,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipDeviceAttributeMultiprocessorCount = 3
};
template <class T> inline hipError_t hipMalloc(T **p, size_t n) {
    *p = (T *)calloc(1, n);
    return hipSuccess;
}
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n) {
    *p = (T *)calloc(1, n);
    return hipSuccess;
}
inline hipError_t hipFree(void *p) {
    free(p);
    return hipSuccess;
}
inline hipError_t hipHostFree(void *p) {
    free(p);
    return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) {
    *e = nullptr;
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, int) {
    *e = nullptr;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 0.0f;
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, void *, int) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, int) {
    *s = nullptr;
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) {
    memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int *v, int, int) {
    *v = 1;
    return hipSuccess;
}
template <class F> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t) {
    *n = 1;
    return hipSuccess;
}

// a launch runs every thread of the grid to completion, in order
template <class F, class... A> inline void cpu_launch(F f, dim3 g, dim3 b, A... args) {
    gridDim = g;
    blockDim = b;
    for (uint32_t bx = 0; bx < g.x; ++bx)
        for (uint32_t tx = 0; tx < b.x; ++tx) {
            blockIdx.x = bx;
            threadIdx.x = tx;
            f(args...);
        }
}
#define hipLaunchKernelGGL(k, g, b, lds, s, ...) cpu_launch(k, dim3(g), dim3(b), __VA_ARGS__)
