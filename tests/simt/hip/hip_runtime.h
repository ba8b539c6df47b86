// Host stand-in for <hip/hip_runtime.h>, used only to build the library's
// batched-solver sources for the CPU in the test suite (g++ -Itests/simt).
//
// A kernel launch runs each workgroup on the calling thread as blockDim.x
// cooperative fibers (one per lane, own stack).  __syncthreads() is a
// rendezvous of every lane of the workgroup, __shfl_xor a rendezvous of the
// 64 lanes of one wavefront, so the device code runs unchanged with real
// barrier and shuffle semantics.  A barrier that some lanes never reach is
// reported as a deadlock (an exception) instead of hanging.  Lanes run in
// increasing order between rendezvous, or in decreasing order when
// SIMT_REVERSE=1, so two runs expose results that depend on lane order
// (data races between barriers).  Device memory is host memory, filled with
// 0xFF bytes (NaN doubles) on allocation so reads of unwritten memory show.
// Stream capture records launches; a graph launch replays them.
#pragma once

#include <setjmp.h>
#include <ucontext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define HIP_KERNEL_NAME(...) __VA_ARGS__
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)::simt::dynamic_lds();
// the library's 16-byte aligned dynamic LDS (operator new storage: aligned)
#define FSP_DYNAMIC_LDS(var) double *var = (double *)::simt::dynamic_lds()

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace simt {

constexpr unsigned kWaveLanes = 64;
enum LaneState { kRun, kWaitBlock, kWaitWave, kDone };

struct Lane {
    jmp_buf jb;
    ucontext_t uc;
    LaneState state = kRun;
    bool started = false;
    unsigned tid = 0;
    char *stack = nullptr;
};

struct Group {
    dim3 bdim, bid, gdim;
    unsigned nthreads = 1;
    std::vector<Lane> lanes;
    Lane *cur = nullptr;
    jmp_buf sched;
    unsigned b_arrived = 0;
    std::vector<unsigned> w_arrived;
    std::vector<uint64_t> xbuf;
    std::vector<unsigned char> lds;
    std::function<void()> body;
};

inline Group *&group()
{
    static Group *g = nullptr;
    return g;
}

inline void *dynamic_lds() { return group()->lds.data(); }

struct Idx3 {
    unsigned x, y, z;
};
inline Idx3 thread_idx() { return {group()->cur->tid, 0, 0}; }
inline Idx3 block_idx() { return {group()->bid.x, 0, 0}; }
inline Idx3 block_dim() { return {group()->bdim.x, 1, 1}; }
inline Idx3 grid_dim() { return {group()->gdim.x, 1, 1}; }

inline void yield_to_scheduler()
{
    Group *g = group();
    if (!_setjmp(g->cur->jb)) _longjmp(g->sched, 1);
}

inline void syncthreads()
{
    Group *g = group();
    if (++g->b_arrived == g->nthreads) {
        g->b_arrived = 0;
        for (Lane &L : g->lanes)
            if (L.state == kWaitBlock) L.state = kRun;
        return;
    }
    g->cur->state = kWaitBlock;
    yield_to_scheduler();
}

inline void wave_barrier(unsigned w)
{
    Group *g = group();
    const unsigned first = w * kWaveLanes;
    const unsigned count = std::min(kWaveLanes, g->nthreads - first);
    if (++g->w_arrived[w] == count) {
        g->w_arrived[w] = 0;
        for (unsigned t = first; t < first + count; ++t)
            if (g->lanes[t].state == kWaitWave) g->lanes[t].state = kRun;
        return;
    }
    g->cur->state = kWaitWave;
    yield_to_scheduler();
}

template <typename T>
inline T shfl_xor(T v, int mask, int width)
{
    static_assert(sizeof(T) <= 8, "shuffle of more than 8 bytes");
    Group *g = group();
    const unsigned t = g->cur->tid, w = t / kWaveLanes, l = t % kWaveLanes;
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    g->xbuf[t] = bits;
    wave_barrier(w);
    unsigned src = l ^ (unsigned)mask;
    if (src >= ((l + (unsigned)width) & ~((unsigned)width - 1))) src = l;
    const uint64_t got = g->xbuf[w * kWaveLanes + src];
    wave_barrier(w);
    T r;
    std::memcpy(&r, &got, sizeof(T));
    return r;
}

// v of lane `src` of the calling lane's wave (every lane of the wave calls)
inline int readlane(int v, int src)
{
    Group *g = group();
    const unsigned t = g->cur->tid, w = t / kWaveLanes;
    g->xbuf[t] = (uint64_t)(uint32_t)v;
    wave_barrier(w);
    const uint64_t got = g->xbuf[w * kWaveLanes + (unsigned)src];
    wave_barrier(w);
    return (int)(uint32_t)got;
}

inline void lane_entry()
{
    Group *g = group();
    g->body();
    g->cur->state = kDone;
    _longjmp(g->sched, 1);
}

inline bool reverse_order()
{
    const char *e = std::getenv("SIMT_REVERSE");
    return e && e[0] == '1';
}

constexpr size_t kStack = 256 * 1024;

// Runs one workgroup to completion; throws on a barrier deadlock.
inline void run_group(dim3 grid, dim3 block, unsigned bid, size_t lds_bytes,
                      const std::function<void()> &body)
{
    Group g;
    g.bdim = block;
    g.gdim = grid;
    g.bid = dim3(bid);
    g.nthreads = block.x;
    g.lanes.resize(block.x);
    g.w_arrived.assign((block.x + kWaveLanes - 1) / kWaveLanes, 0);
    g.xbuf.assign(block.x, 0);
    g.lds.assign(lds_bytes + 16, 0xFF);
    g.body = body;
    Group *saved = group();
    group() = &g;
    static std::vector<char *> stacks;
    while (stacks.size() < block.x) stacks.push_back((char *)std::malloc(kStack));
    for (unsigned t = 0; t < block.x; ++t) {
        Lane &L = g.lanes[t];
        L.tid = t;
        L.stack = stacks[t];
        getcontext(&L.uc);
        L.uc.uc_stack.ss_sp = L.stack;
        L.uc.uc_stack.ss_size = kStack;
        L.uc.uc_link = nullptr;
        makecontext(&L.uc, (void (*)())lane_entry, 0);
    }
    const bool rev = reverse_order();
    for (;;) {
        bool ran = false, all_done = true;
        for (unsigned k = 0; k < block.x; ++k) {
            Lane &L = g.lanes[rev ? block.x - 1 - k : k];
            if (L.state != kRun) {
                all_done = all_done && L.state == kDone;
                continue;
            }
            all_done = false;
            ran = true;
            g.cur = &L;
            if (!_setjmp(g.sched)) {
                if (!L.started) {
                    L.started = true;
                    setcontext(&L.uc);
                }
                _longjmp(L.jb, 1);
            }
        }
        if (all_done) break;
        if (!ran) {
            unsigned nb = 0, nw = 0, nd = 0;
            for (Lane &L : g.lanes) {
                nb += L.state == kWaitBlock;
                nw += L.state == kWaitWave;
                nd += L.state == kDone;
            }
            group() = saved;
            throw std::runtime_error("simt deadlock in workgroup " + std::to_string(bid) + ": " +
                                     std::to_string(nb) + " lanes at a workgroup barrier, " +
                                     std::to_string(nw) + " at a wavefront shuffle, " +
                                     std::to_string(nd) + " exited");
        }
    }
    group() = saved;
}

// ------------------------------------------------------------ streams/graphs
struct Graph {
    std::vector<std::function<void()>> nodes;
};

inline void *&capturing()
{
    static void *s = nullptr;
    return s;
}
inline Graph *&capture_graph()
{
    static Graph *g = nullptr;
    return g;
}

inline void submit(void *stream, std::function<void()> fn)
{
    if (capturing() && stream == capturing()) {
        capture_graph()->nodes.push_back(std::move(fn));
        return;
    }
    fn();
}

// Sticky launch error (a deadlock), reported by the next synchronising call.
inline std::string &launch_error()
{
    static std::string e;
    return e;
}

template <typename... P, typename... A>
inline void launch(void (*kernel)(P...), dim3 grid, dim3 block, size_t lds, void *stream,
                   A... args)
{
    auto call = [=]() { kernel(args...); };
    submit(stream, [=]() {
        if (!launch_error().empty()) return;
        try {
            for (unsigned b = 0; b < grid.x; ++b) run_group(grid, block, b, lds, call);
        } catch (const std::exception &ex) {
            launch_error() = ex.what();
        }
    });
}

}  // namespace simt

#define threadIdx (::simt::thread_idx())
#define blockIdx (::simt::block_idx())
#define blockDim (::simt::block_dim())
#define gridDim (::simt::grid_dim())

inline void __syncthreads() { ::simt::syncthreads(); }
// wavefront-scope fence and barrier: a rendezvous of the calling lane's wave
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define __builtin_amdgcn_sched_barrier(mask) ((void)0)
#define __builtin_amdgcn_wave_barrier() \
    ::simt::wave_barrier(::simt::group()->cur->tid / ::simt::kWaveLanes)
#define __builtin_amdgcn_readlane(v, lane) ::simt::readlane((v), (lane))
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
    return ::simt::shfl_xor(v, mask, width);
}
template <typename T>
inline T atomicAdd(T *p, T v)
{
    const T old = *p;
    *p = old + v;
    return old;
}
// the device's constant 100 MHz clock: steady_clock in 10 ns ticks
inline unsigned long long wall_clock64()
{
    return (unsigned long long)(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                    std::chrono::steady_clock::now().time_since_epoch())
                                    .count() /
                                10);
}
// clang's non-temporal load hint: a plain load on the CPU
#define __builtin_nontemporal_load(p) (*(p))
using std::isfinite;
using std::isinf;
using std::isnan;
using std::max;
using std::min;

// ------------------------------------------------------------------ runtime
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorNoDevice = 100, hipErrorLaunchFailure = 719 };
typedef void *hipStream_t;
typedef simt::Graph *hipGraph_t;
typedef simt::Graph *hipGraphExec_t;
typedef std::chrono::steady_clock::time_point *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                     hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal,
                            hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1, hipDeviceAttributeWallClockRate = 2 };
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipStreamNonBlocking = 1;

struct hipDeviceProp_t {
    char name[256];
    char gcnArchName[256];
    int multiProcessorCount;
    size_t sharedMemPerBlock;
};

inline int simt_cus()
{
    const char *e = std::getenv("SIMT_CUS");
    return e ? std::atoi(e) : 4;
}

inline const char *hipGetErrorString(hipError_t e)
{
    return e == hipErrorLaunchFailure ? simt::launch_error().c_str() : "simulated HIP error";
}
inline hipError_t simt_status() { return simt::launch_error().empty() ? hipSuccess : hipErrorLaunchFailure; }
inline hipError_t hipGetLastError() { return simt_status(); }
inline hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    *v = a == hipDeviceAttributeMultiprocessorCount ? simt_cus()
         : a == hipDeviceAttributeWallClockRate     ? 100000  // kHz
                                                    : 0;
    return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof p->name, "simt");
    std::snprintf(p->gcnArchName, sizeof p->gcnArchName, "simt-cpu");
    p->multiProcessorCount = simt_cus();
    p->sharedMemPerBlock = 160 * 1024;
    return hipSuccess;
}
template <typename T>
inline hipError_t hipMalloc(T **p, size_t bytes)
{
    const size_t b = ((bytes + 255) / 256) * 256 + 256;
    void *q = std::aligned_alloc(256, b);
    if (!q) return hipErrorInvalidValue;
    std::memset(q, 0xFF, b);
    *p = (T *)q;
    return hipSuccess;
}
inline hipError_t hipFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void **p, size_t bytes, unsigned)
{
    *p = std::malloc(bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipHostFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    if (n) std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st)
{
    simt::submit(st, [=]() { hipMemcpy(d, s, n, k); });
    return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st)
{
    simt::submit(st, [=]() { hipMemset(d, v, n); });
    return hipSuccess;
}
// the calling thread's default stream: a handle that is never captured
#define hipStreamPerThread ((hipStream_t)(uintptr_t)0x2)
inline hipError_t hipDeviceSynchronize() { return simt_status(); }
inline hipError_t hipStreamSynchronize(hipStream_t) { return simt_status(); }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = new int(0);
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s)
{
    delete (int *)s;
    return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    if (!s || simt::capturing()) return hipErrorInvalidValue;
    simt::capturing() = s;
    simt::capture_graph() = new simt::Graph();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    if (s != simt::capturing()) return hipErrorInvalidValue;
    *g = simt::capture_graph();
    simt::capturing() = nullptr;
    simt::capture_graph() = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, char *, size_t)
{
    *e = new simt::Graph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s)
{
    for (auto &fn : e->nodes) simt::submit(s, fn);
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete e;
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new std::chrono::steady_clock::time_point();
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    *e = std::chrono::steady_clock::now();
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// streams run their work at submission, so an event is always complete
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    *ms = std::chrono::duration<float, std::milli>(*b - *a).count();
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
    ::simt::launch(kernel, grid, block, lds, stream, __VA_ARGS__)
