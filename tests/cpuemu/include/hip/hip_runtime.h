// CPU emulation of the HIP subset that libqms uses.  TEST INFRASTRUCTURE ONLY:
// tests/cpuemu/build.py compiles the library's own .hip sources against this
// header with g++ (AddressSanitizer + UBSan) into tests/cpuemu/_build/
// libqms_cpu.so, so that every kernel's indexing can be checked on the CPU
// before it runs on a GPU.  Nothing in the product includes this file.
//
// Execution model:
//  * every launch runs at once, in the calling thread's order (streams are
//    in-order queues, so this is one legal schedule of them);
//  * the blocks of a launch run one after another; the threads of a block
//    are fibers that each run until the next barrier, so __syncthreads()
//    and the barrier-based exchange of __shfl_xor order LDS traffic as on
//    the GPU, and a missing barrier shows up as a wrong result;
//  * __shared__ arrays are function-local statics: one block at a time uses
//    them, as one workgroup owns its LDS;
//  * device memory is host heap memory, so AddressSanitizer reports any access
//    outside an allocation;
//  * stream capture records the work of a stream as closures; launching the
//    instantiated graph replays them.
#pragma once
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <map>
#include <stdio.h>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#endif

#define __global__
#define __device__
#define __align__(n) __attribute__((aligned(n)))
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define HIP_SYMBOL(s) (&(s))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct qemu_uint3 {
    unsigned x, y, z;
};
inline qemu_uint3 threadIdx{0, 0, 0};
inline qemu_uint3 blockIdx{0, 0, 0};
inline dim3 blockDim;
inline dim3 gridDim;

template <class A, class B>
inline auto min(A a, B b) -> typename std::common_type<A, B>::type {
    using C = typename std::common_type<A, B>::type;
    return (C)a < (C)b ? (C)a : (C)b;
}
template <class A, class B>
inline auto max(A a, B b) -> typename std::common_type<A, B>::type {
    using C = typename std::common_type<A, B>::type;
    return (C)a < (C)b ? (C)b : (C)a;
}

// gfx950 builtins: value of the first lane (the operands are wave-uniform in
// every use), reciprocal, scheduling hint.
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_sched_barrier(x) ((void)0)

namespace qemu {

// The threads of one block are fibers on the calling OS thread: each runs
// until it reaches a barrier or returns, and a barrier is released when every
// thread of the block that has not returned is waiting at it.
// Minimal x86-64 stack switch (tests/cpuemu/qemu_switch.S): saves the
// callee-saved registers on the current stack, stores its pointer in *save and
// resumes the stack `to` (glibc's swapcontext also saves the signal mask, a
// system call per switch).
extern "C" void qemu_switch(void **save, void *to);

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;
    bool done = false, waiting = false;
};
constexpr size_t FIBER_STACK = 256 * 1024;

struct Block {
    unsigned nt;
    std::vector<Fiber> *fib;
    std::vector<uint64_t> shfl;
    unsigned bx, by, bz;
    const std::function<void()> *body;
    void *sched_sp = nullptr;
    unsigned cur = 0;
    void *fake = nullptr;       // AddressSanitizer fake-stack handle of the scheduler
    const void *main_bottom = nullptr;
    size_t main_size = 0;
    bool used_barrier = false;
};
inline Block *cur_block = nullptr;

inline void set_ids(Block &b, unsigned t) {
    threadIdx = qemu_uint3{t % blockDim.x, (t / blockDim.x) % blockDim.y, t / (blockDim.x * blockDim.y)};
    blockIdx = qemu_uint3{b.bx, b.by, b.bz};
}

// Switch from the running fiber back to the scheduler.
inline void to_sched(Block &b, bool dying) {
    Fiber &f = (*b.fib)[b.cur];
#if defined(__SANITIZE_ADDRESS__)
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(dying ? nullptr : &fake, b.main_bottom, b.main_size);
    qemu_switch(&f.sp, b.sched_sp);
    __sanitizer_finish_switch_fiber(fake, &b.main_bottom, &b.main_size);
#else
    (void)dying;
    qemu_switch(&f.sp, b.sched_sp);
#endif
}

extern "C" inline void fiber_entry() {
    Block &b = *cur_block;
#if defined(__SANITIZE_ADDRESS__)
    __sanitizer_finish_switch_fiber(nullptr, &b.main_bottom, &b.main_size);
#endif
    (*b.body)();
    (*b.fib)[b.cur].done = true;
    to_sched(b, true);
}

inline void barrier() {
    Block &b = *cur_block;
    if (!b.fib) {
        fprintf(stderr, "qemu: barrier in a kernel first launched without one\n");
        abort();
    }
    b.used_barrier = true;
    (*b.fib)[b.cur].waiting = true;
    to_sched(b, false);
}

inline std::vector<Fiber> &fibers(unsigned n) {
    static std::vector<Fiber> *v = new std::vector<Fiber>();
    if (v->size() < n) {
        v->resize(n);
        for (Fiber &f : *v)
            if (!f.stack) f.stack = (char *)malloc(FIBER_STACK);
    }
    return *v;
}

// Threads of a kernel that has never reached a barrier run one after the
// other on the calling stack (no fibers).
inline void run_block_serial(Block &b) {
    b.fib = nullptr;
    for (unsigned t = 0; t < b.nt; ++t) {
        b.cur = t;
        set_ids(b, t);
        (*b.body)();
    }
}

inline void run_block(Block &b) {
    std::vector<Fiber> &fib = fibers(b.nt);
    b.fib = &fib;
    for (unsigned t = 0; t < b.nt; ++t) {
        Fiber &f = fib[t];
        f.done = false;
        f.waiting = false;
        // initial frame: six zeroed callee-saved registers, then the return
        // address fiber_entry, with the stack 16-byte aligned at its entry
        uintptr_t top = ((uintptr_t)(f.stack + FIBER_STACK)) & ~(uintptr_t)15;
        void **sp = (void **)top;
        *--sp = nullptr;
        *--sp = (void *)&fiber_entry;
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        f.sp = sp;
    }
    for (;;) {
        bool any = false;
        for (unsigned t = 0; t < b.nt; ++t) {
            Fiber &f = fib[t];
            if (f.done || f.waiting) continue;
            any = true;
            b.cur = t;
            set_ids(b, t);
#if defined(__SANITIZE_ADDRESS__)
            __sanitizer_start_switch_fiber(&b.fake, f.stack, FIBER_STACK);
#endif
            qemu_switch(&b.sched_sp, f.sp);
#if defined(__SANITIZE_ADDRESS__)
            __sanitizer_finish_switch_fiber(b.fake, nullptr, nullptr);
#endif
        }
        bool live = false;
        for (unsigned t = 0; t < b.nt; ++t) {
            if (!fib[t].done) live = true;
            fib[t].waiting = false;    // every live thread reached the barrier: release it
        }
        if (!live) break;
        (void)any;
    }
}

// One launch at a time in the whole process.
inline std::recursive_mutex &launch_mutex() {
    static std::recursive_mutex m;
    return m;
}

struct Stream {
    bool capturing = false;
    std::vector<std::function<void()>> rec;
};
struct Graph {
    std::vector<std::function<void()>> ops;
};
inline Stream &legacy_stream() {
    static Stream s;
    return s;
}

}  // namespace qemu

typedef qemu::Stream *hipStream_t;
typedef struct qemu_event {
    double t;
} *hipEvent_t;
typedef qemu::Graph *hipGraph_t;
typedef qemu::Graph *hipGraphExec_t;
typedef struct qemu_graph_node *hipGraphNode_t;

enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
};
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 };
enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
};
#define hipStreamNonBlocking 1u
#define hipStreamDefault 0u
#define hipHostMallocDefault 0u
#define hipEventDefault 0u

inline const char *hipGetErrorString(hipError_t e) {
    switch (e) {
    case hipSuccess: return "no error";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorNotReady: return "not ready";
    default: return "invalid value";
    }
}

namespace qemu {
inline hipStream_t resolve(hipStream_t s) { return s ? s : &legacy_stream(); }
// Runs f now, or records it if the stream is being captured.
inline void submit(hipStream_t s, std::function<void()> f) {
    hipStream_t q = resolve(s);
    if (q->capturing)
        q->rec.push_back(std::move(f));
    else
        f();
}

template <typename... P, typename... A>
void launch(void (*k)(P...), dim3 grid, dim3 block, size_t, hipStream_t s, A &&...args) {
    using Tup = std::tuple<typename std::decay<P>::type...>;
    auto tup = std::make_shared<Tup>(std::forward<A>(args)...);
    submit(s, [k, grid, block, tup]() {
        std::lock_guard<std::recursive_mutex> lk(launch_mutex());
        // kernels seen to run without a barrier take the serial path
        static std::map<const void *, bool> *barrier_free = new std::map<const void *, bool>();
        auto known = barrier_free->find((const void *)k);
        const bool serial = known != barrier_free->end() && known->second;
        bool used = false;
        const unsigned nt = block.x * block.y * block.z;
        if (nt == 0 || nt > 1024 || grid.x == 0 || grid.y == 0 || grid.z == 0) {
            fprintf(stderr, "qemu: invalid launch configuration\n");
            abort();
        }
        gridDim = grid;
        blockDim = block;
        for (unsigned bz = 0; bz < grid.z; ++bz)
            for (unsigned by = 0; by < grid.y; ++by)
                for (unsigned bx = 0; bx < grid.x; ++bx) {
                    const std::function<void()> body = [&]() { std::apply(k, *tup); };
                    Block blk;
                    blk.nt = nt;
                    blk.shfl.assign(nt, 0);
                    blk.bx = bx; blk.by = by; blk.bz = bz;
                    blk.body = &body;
                    Block *prev = cur_block;
                    cur_block = &blk;
                    if (serial) run_block_serial(blk);
                    else run_block(blk);
                    used = used || blk.used_barrier;
                    cur_block = prev;
                }
        if (known == barrier_free->end()) (*barrier_free)[(const void *)k] = !used;
    });
}

inline unsigned flat_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
}  // namespace qemu

#define hipLaunchKernelGGL(K, G, B, SH, S, ...) qemu::launch(K, dim3(G), dim3(B), (size_t)(SH), (S), ##__VA_ARGS__)

inline void __syncthreads() { qemu::barrier(); }
// the workgroup barrier and fences as builtins (LDS-only barriers): every
// fiber stops at the barrier, and fibers share memory, so the fences are void
#define __builtin_amdgcn_s_barrier() qemu::barrier()
#define __builtin_amdgcn_fence(...) ((void)0)

template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    qemu::Block &b = *qemu::cur_block;
    const unsigned t = qemu::flat_tid();
    uint64_t bits = 0;
    memcpy(&bits, &v, sizeof(T));
    b.shfl[t] = bits;
    qemu::barrier();
    const unsigned src = (t & ~(unsigned)(width - 1)) | ((t & (unsigned)(width - 1)) ^ (unsigned)mask);
    T r;
    const uint64_t got = b.shfl[src];
    memcpy(&r, &got, sizeof(T));
    qemu::barrier();
    return r;
}

// ----------------------------------------------------------------------------
// Runtime API
// ----------------------------------------------------------------------------
inline hipError_t hipMalloc(void **p, size_t n) {
    *p = aligned_alloc(256, (n + 255) / 256 * 256);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T **p, size_t n) {
    return hipMalloc((void **)p, n);
}
inline hipError_t hipFree(void *p) {
    free(p);
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned = 0) { return hipMalloc(p, n); }
template <typename T>
inline hipError_t hipHostMalloc(T **p, size_t n, unsigned f = 0) {
    return hipHostMalloc((void **)p, n, f);
}
inline hipError_t hipHostFree(void *p) { return hipFree(p); }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st = nullptr) {
    qemu::submit(st, [d, s, n]() { memcpy(d, s, n); });
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k) {
    return hipMemcpyAsync(d, s, n, k, nullptr);
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr) {
    qemu::submit(st, [d, v, n]() { memset(d, v, n); });
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = new qemu::Stream();
    return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) {
    *e = new qemu_event{0.0};
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline double qemu_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr) {
    qemu::submit(s, [e]() { e->t = qemu_now_ms(); });
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = (float)(b->t - a->t);
    return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
    *v = 256;
    return hipSuccess;
}
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t) {
    *n = 4;
    return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t *fr, size_t *tot) {
    *fr = (size_t)64 << 30;
    *tot = (size_t)64 << 30;
    return hipSuccess;
}
template <typename T>
inline hipError_t hipGetSymbolAddress(void **p, const T *sym) {
    *p = (void *)sym;
    return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    hipStream_t q = qemu::resolve(s);
    if (q->capturing) return hipErrorInvalidValue;
    q->capturing = true;
    q->rec.clear();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    hipStream_t q = qemu::resolve(s);
    if (!q->capturing) return hipErrorInvalidValue;
    q->capturing = false;
    *g = new qemu::Graph{std::move(q->rec)};
    q->rec.clear();
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *e = new qemu::Graph{g->ops};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    for (auto &op : e->ops) qemu::submit(s, op);
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
