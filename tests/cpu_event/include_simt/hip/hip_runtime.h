// Host-only stand-in for the HIP runtime header with 64-lane wavefronts,
// used to compile the event path's device code (gegs_amd/csrc/event.hpp and
// what it includes) as plain C++ and run it with the cross-lane semantics of
// the GPU.
//
// Every thread of a block is a fiber (ucontext) on one host thread.  A
// wavefront is 64 consecutive threads of a block.  The scheduler runs every
// fiber of the block until it blocks in a wavefront operation (ballot, any,
// readlane, readfirstlane, shfl_xor) or in __syncthreads, then executes one
// operation for a group of lanes, as the hardware's exec mask would:
//
//   * lanes inside a GEGS_WAVE_SUBSET_OP region (an operation shared by the
//     lanes that reach it together, such as a bank append issued from
//     divergent code) form a group of their own and execute it together;
//   * every other wavefront operation needs every lane of the wavefront that
//     has not returned, all at the same operation: lanes at different
//     operations are divergent use of a whole-wavefront operation and stop
//     the run with an error, as do a readlane of a lane outside the group and
//     a __syncthreads reached by part of a wavefront.
//
// Blocks of a launch run one after another (a legal schedule: no kernel
// waits on another block); the waves of a block interleave at every
// wavefront operation.  Atomics are plain read-modify-writes (one host
// thread).  Built with AddressSanitizer, the fiber switches are announced to
// it so that stack checks follow the fibers.
#pragma once
#include <setjmp.h>
#include <ucontext.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define SIMT_ASAN 1
#endif
#endif
#if defined(__SANITIZE_ADDRESS__)
#define SIMT_ASAN 1
#endif
#ifdef SIMT_ASAN
#include <sanitizer/common_interface_defs.h>
#endif

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

namespace simt {

enum Op { OP_NONE = 0, OP_BALLOT, OP_ANY, OP_READLANE, OP_READFIRST, OP_SHFL_XOR, OP_SYNC };
enum State { RUNNABLE, WAITING, DONE };
static const char *op_name(int op) {
    static const char *n[] = {"none", "ballot", "any", "readlane", "readfirstlane", "shfl_xor", "syncthreads"};
    return n[op];
}

constexpr size_t STACK_BYTES = 256 * 1024;

struct Fiber {
    ucontext_t ctx;       // first entry only
    jmp_buf jb;           // where the fiber resumes
    bool started = false;
    char *stack = nullptr;
    State st = RUNNABLE;
    int op = OP_NONE;
    uint64_t arg = 0;     // the lane's operand
    int32_t arg2 = 0;     // source lane (readlane) or lane mask (shfl_xor)
    const void *site = nullptr;   // return address of the operation's call
    uint64_t res = 0;
    int scope = 0;        // GEGS_WAVE_SUBSET_OP region the lane is in (0: none)
    uint32_t tid = 0;
    void *fake_stack = nullptr;
};

struct Sched {
    jmp_buf main_jb;
    void *main_fake = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
    std::vector<Fiber> fibers;
    std::vector<char *> stacks;    // fiber stacks, kept from block to block
    Fiber *cur = nullptr;
    std::function<void()> body;
    uint64_t ops = 0;     // wavefront operations executed
};
inline Sched S;

[[noreturn]] inline void die(const char *what, uint32_t tid, int op) {
    fprintf(stderr, "SIMT error: %s (block %u, thread %u, lane %u, %s)\n", what, blockIdx.x, tid, tid & 63u,
            op_name(op));
    fflush(stderr);
    abort();
}

// Switches: the first entry into a fiber goes through its ucontext, every
// later switch through _setjmp/_longjmp (no signal-mask system call).

// from a fiber back to the scheduler
inline void yield_to_main() {
    Fiber *f = S.cur;
    if (!_setjmp(f->jb)) {
#ifdef SIMT_ASAN
        __sanitizer_start_switch_fiber(&f->fake_stack, S.main_bottom, S.main_size);
#endif
        _longjmp(S.main_jb, 1);
    }
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(f->fake_stack, &S.main_bottom, &S.main_size);
#endif
}

inline void trampoline() {
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &S.main_bottom, &S.main_size);
#endif
    S.body();
    Fiber *f = S.cur;
    f->st = DONE;
#ifdef SIMT_ASAN
    __sanitizer_start_switch_fiber(nullptr, S.main_bottom, S.main_size);
#endif
    _longjmp(S.main_jb, 1);
}

inline void resume(Fiber &f) {
    S.cur = &f;
    threadIdx.x = f.tid;
    if (!_setjmp(S.main_jb)) {
#ifdef SIMT_ASAN
        __sanitizer_start_switch_fiber(&S.main_fake, f.stack, STACK_BYTES);
#endif
        if (!f.started) {
            f.started = true;
            ucontext_t here;
            swapcontext(&here, &f.ctx);   // never returns: the fiber comes back by _longjmp
        }
        _longjmp(f.jb, 1);
    }
#ifdef SIMT_ASAN
    __sanitizer_finish_switch_fiber(S.main_fake, nullptr, nullptr);
#endif
    S.cur = nullptr;
}

// a lane reaches a wavefront operation (or __syncthreads) and waits for its group
inline uint64_t wait(int op, uint64_t arg, int32_t arg2, const void *site) {
    Fiber *f = S.cur;
    if (!f) {
        fprintf(stderr, "SIMT error: wavefront operation outside a kernel\n");
        abort();
    }
    f->op = op;
    f->arg = arg;
    f->arg2 = arg2;
    f->site = site;
    f->st = WAITING;
    yield_to_main();
    return f->res;
}

// one operation for the group g (indices into S.fibers of one wavefront);
// a whole-wavefront group must also be at one call site
inline void execute(const std::vector<int> &g, bool same_site) {
    const int op = S.fibers[g[0]].op;
    const void *site = S.fibers[g[0]].site;
    uint64_t mask = 0;
    for (int i : g) {
        Fiber &f = S.fibers[i];
        if (f.op != op) die("lanes of one group at different wavefront operations", f.tid, f.op);
        if (same_site && f.site != site)
            die("lanes of one wavefront at the same operation from different call sites (divergent use of a "
                "whole-wavefront operation)", f.tid, f.op);
        mask |= 1ull << (f.tid & 63u);
    }
    auto lane_arg = [&](uint32_t lane, const Fiber &asker) -> uint64_t {
        if (!(mask >> lane & 1ull)) die("read of a lane outside the executing group", asker.tid, asker.op);
        for (int i : g)
            if ((S.fibers[i].tid & 63u) == lane) return S.fibers[i].arg;
        die("lane not found", asker.tid, asker.op);
    };
    uint64_t ballot = 0;
    for (int i : g) ballot |= (S.fibers[i].arg ? 1ull : 0ull) << (S.fibers[i].tid & 63u);
    const int first = __builtin_ctzll(mask);
    for (int i : g) {
        Fiber &f = S.fibers[i];
        switch (op) {
        case OP_BALLOT: f.res = ballot; break;
        case OP_ANY: f.res = ballot != 0; break;
        case OP_READLANE:
            if (f.arg2 != S.fibers[g[0]].arg2) die("readlane with a lane index that is not uniform", f.tid, op);
            f.res = lane_arg((uint32_t)f.arg2 & 63u, f);
            break;
        case OP_READFIRST: f.res = lane_arg((uint32_t)first, f); break;
        case OP_SHFL_XOR: f.res = lane_arg(((f.tid & 63u) ^ (uint32_t)f.arg2) & 63u, f); break;
        default: die("unknown operation", f.tid, op);
        }
        f.st = RUNNABLE;
        f.op = OP_NONE;
    }
    S.ops++;
}

// run one block of threads to completion
inline void run_block(uint32_t nthreads) {
    // a block that is not whole wavefronts ends with a partial wavefront
    // whose missing lanes are inactive, as on the GPU
    S.fibers.assign(nthreads, Fiber());
    while (S.stacks.size() < nthreads) S.stacks.push_back((char *)malloc(STACK_BYTES));
    for (uint32_t t = 0; t < nthreads; ++t) {
        Fiber &f = S.fibers[t];
        f.tid = t;
        f.stack = S.stacks[t];
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = STACK_BYTES;
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, (void (*)())trampoline, 0);
    }
    const uint32_t nwaves = (nthreads + 63) / 64;
    for (;;) {
        for (Fiber &f : S.fibers)
            if (f.st == RUNNABLE) resume(f);
        bool all_done = true, progress = false, all_sync = true;
        for (const Fiber &f : S.fibers) {
            if (f.st != DONE) all_done = false;
            if (f.st == WAITING && f.op != OP_SYNC) all_sync = false;
        }
        if (all_done) break;
        for (uint32_t w = 0; w < nwaves; ++w) {
            std::vector<int> sub, rest;
            int scope = 0;
            for (uint32_t l = 0; l < 64 && w * 64 + l < nthreads; ++l) {
                const int i = (int)(w * 64 + l);
                const Fiber &f = S.fibers[i];
                if (f.st != WAITING) continue;
                if (f.scope) {
                    if (!scope) scope = f.scope;
                    if (f.scope == scope) sub.push_back(i);
                } else {
                    rest.push_back(i);
                }
            }
            if (!sub.empty()) {
                execute(sub, false);
                progress = true;
                continue;
            }
            if (rest.empty()) continue;
            bool any_sync = false, any_other = false;
            for (int i : rest) (S.fibers[i].op == OP_SYNC ? any_sync : any_other) = true;
            if (any_sync && any_other) die("__syncthreads reached by part of a wavefront", S.fibers[rest[0]].tid, OP_SYNC);
            if (any_sync) continue;   // the wave waits at the block barrier
            execute(rest, true);
            progress = true;
        }
        if (!progress) {
            if (!all_sync) {
                fprintf(stderr, "SIMT error: no wavefront can proceed\n");
                abort();
            }
            for (Fiber &f : S.fibers)
                if (f.st == WAITING) {
                    f.st = RUNNABLE;
                    f.op = OP_NONE;
                }
        }
    }
    S.fibers.clear();
}

// marks a region whose wavefront operations are shared by the lanes that
// enter it together (id: one per region kind)
struct SubsetScope {
    explicit SubsetScope(int id) {
        if (S.cur) S.cur->scope = id;
    }
    ~SubsetScope() {
        if (S.cur) S.cur->scope = 0;
    }
};

template <class T> inline uint64_t bits(T v) {
    uint64_t u = 0;
    memcpy(&u, &v, sizeof(T));
    return u;
}
template <class T> inline T from_bits(uint64_t u) {
    T v;
    memcpy(&v, &u, sizeof(T));
    return v;
}

}  // namespace simt

#define GEGS_WAVE_SUBSET_OP(id) simt::SubsetScope simt_subset_scope_(id)

inline uint32_t __lane_id() { return threadIdx.x & 63u; }
#define SIMT_SITE __builtin_return_address(0)
__attribute__((noinline)) inline uint64_t __ballot(int p) { return simt::wait(simt::OP_BALLOT, p != 0, 0, SIMT_SITE); }
__attribute__((noinline)) inline int __any(int p) { return (int)simt::wait(simt::OP_ANY, p != 0, 0, SIMT_SITE); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
template <class T> __attribute__((noinline)) T simt_readlane(T x, int lane) {
    return simt::from_bits<T>(simt::wait(simt::OP_READLANE, simt::bits(x), lane, SIMT_SITE));
}
template <class T> __attribute__((noinline)) T simt_readfirstlane(T x) {
    return simt::from_bits<T>(simt::wait(simt::OP_READFIRST, simt::bits(x), 0, SIMT_SITE));
}
#define __builtin_amdgcn_readfirstlane(x) simt_readfirstlane(x)
#define __builtin_amdgcn_readlane(x, l) simt_readlane((x), (int)(l))
__attribute__((noinline)) inline uint32_t __shfl_xor(uint32_t v, int mask, int = 64) {
    return (uint32_t)simt::wait(simt::OP_SHFL_XOR, v, mask, SIMT_SITE);
}
__attribute__((noinline)) inline void __syncthreads() { simt::wait(simt::OP_SYNC, 0, 0, SIMT_SITE); }
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

// a launch runs its blocks one after another, the threads of each block as
// interleaved 64-lane wavefronts
template <class F, class... A> inline void cpu_launch(F f, dim3 g, dim3 b, A... args) {
    gridDim = g;
    blockDim = b;
    simt::S.body = [&]() { f(args...); };
    for (uint32_t bx = 0; bx < g.x; ++bx) {
        blockIdx.x = bx;
        simt::run_block(b.x);
    }
    simt::S.body = nullptr;
}
#define hipLaunchKernelGGL(k, g, b, lds, s, ...) cpu_launch(k, dim3(g), dim3(b), __VA_ARGS__)
