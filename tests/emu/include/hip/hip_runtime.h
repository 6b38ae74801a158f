// CPU emulation of the subset of the HIP API that gpuspgemm_amd/csrc uses.
// Test infrastructure: it lets the library's kernels and host driver be
// compiled as plain C++ (one OS thread per GPU lane) and run under
// AddressSanitizer on a machine without a GPU.  Device memory is host heap,
// LDS arrays are static storage (blocks run one after another), and stream
// work is queued per stream and run in an adversarial order (emu::submit_on).
//
// Cross-lane operations use count-based barriers: a wave barrier releases
// when every live lane of the wave has arrived at some wave barrier.  The
// kernels only exchange data inside row groups whose lanes share control
// flow, so every group meets itself at the same barrier occurrence.
#pragma once

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline
#define SPG_GLOBAL  // host threads have one address space
#define __noinline__
#define __restrict__
#define __align__(n) __attribute__((aligned(n)))
#define __launch_bounds__(...)
#define SPG_WAVES_PER_EU(n)  // register budgets mean nothing to host threads

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3() = default;
    dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct int2 {
    int x, y;
};
struct int4 {
    int x, y, z, w;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }

namespace emu {
struct Idx {
    unsigned x = 0, y = 0, z = 0;
};
extern thread_local Idx tl_thread, tl_block;
extern dim3 g_block_dim, g_grid_dim;
void block_barrier();
void wave_barrier();
void exchange_write(std::uint64_t v);      // publish this lane's value
std::uint64_t exchange_read(int lane);     // read lane's published value (after a wave barrier)
void launch(dim3 grid, dim3 block, const std::function<void()>& body);
int lane();
// Streams: work (launches, copies, fills, event records, waits for an
// event) is queued per stream and runs when the
// host synchronises or queries (a stream, an event, the device, a
// synchronous copy or free).  Queued work of different streams runs in the
// most adversarial order the dependencies allow: among the streams whose
// first item may run, the item queued last runs first, so work that lacks
// a stream wait or an event runs before what it should follow.  While the
// calling thread captures a graph, work is recorded instead and runs when
// the graph is launched.
using Work = std::function<void()>;
std::vector<Work>*& capture_target();
struct Event {
    void* stream = nullptr;          // stream of the last record (null: never recorded)
    unsigned long long seq = 0;      // its place in the device's order of queued items
};
void submit_on(void* stream, Work w, const char* label = nullptr);
void record_on(void* stream, Event* ev);
void wait_event_on(void* stream, const Event* ev);
void drain_stream(void* stream);
void drain_event(const Event* ev);
void drain_all();
void drain_blocking();  // the legacy and per-thread default streams
bool event_done(const Event* ev);
void* new_stream();
}  // namespace emu

#define threadIdx (::emu::tl_thread)
#define blockIdx (::emu::tl_block)
#define blockDim (::emu::g_block_dim)
#define gridDim (::emu::g_grid_dim)
#define warpSize 64

using std::max;
using std::min;

inline void __syncthreads() { emu::block_barrier(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_system() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline unsigned __lane_id() { return static_cast<unsigned>(emu::lane()); }

#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_wave_barrier() ::emu::wave_barrier()

template <typename T>
inline T emu_shfl_any(T v, int src) {
    std::uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    emu::exchange_write(bits);
    emu::wave_barrier();
    const std::uint64_t r = emu::exchange_read(src);
    emu::wave_barrier();
    T out;
    std::memcpy(&out, &r, sizeof(T));
    return out;
}
template <typename T>
inline T __shfl(T v, int src_lane, int width = 64) {
    const int self = emu::lane();
    return emu_shfl_any(v, (src_lane & (width - 1)) + (self & ~(width - 1)));
}
template <typename T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
    const int self = emu::lane();
    int idx = self - static_cast<int>(delta);
    idx = idx < (self & ~(width - 1)) ? self : idx;
    return emu_shfl_any(v, idx);
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    const int self = emu::lane();
    int idx = self ^ mask;
    idx = idx >= ((self + width) & ~(width - 1)) ? self : idx;
    return emu_shfl_any(v, idx);
}
// DPP quad permutation (the two patterns the library uses: lane ^ 1 and
// lane ^ 2 within each quad) and the swizzle bit mode (lane ^ xor within
// each group of 32: pattern = xor << 10 | and_mask 0x1f).
inline int __builtin_amdgcn_mov_dpp(int v, int ctrl, int, int, bool) {
    const int self = emu::lane();
    const int sel = (ctrl >> (2 * (self & 3))) & 3;
    return emu_shfl_any(v, (self & ~3) | sel);
}
inline int __builtin_amdgcn_ds_swizzle(int v, int pattern) {
    const int self = emu::lane();
    const int and_mask = pattern & 0x1f, or_mask = (pattern >> 5) & 0x1f, xor_mask = (pattern >> 10) & 0x1f;
    const int j = ((self & and_mask) | or_mask) ^ xor_mask;
    return emu_shfl_any(v, (self & ~31) | j);
}
inline unsigned long long __ballot(int pred) {
    emu::exchange_write(pred ? 1u : 0u);
    emu::wave_barrier();
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
        if (emu::exchange_read(l)) m |= 1ull << l;
    emu::wave_barrier();
    return m;
}
// value of the first active lane of the wave
inline unsigned __builtin_amdgcn_readfirstlane(unsigned v) {
    const unsigned long long active = __ballot(1);
    return static_cast<unsigned>(emu_shfl_any(static_cast<int>(v), __builtin_ctzll(active)));
}
// acc + the set bits of `mask` at lane positions below this lane (low / high 32 lanes)
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned mask, unsigned acc) {
    const int self = emu::lane();
    const unsigned below = self >= 32 ? 0xffffffffu : ((1u << self) - 1u);
    return acc + static_cast<unsigned>(__builtin_popcount(mask & below));
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned mask, unsigned acc) {
    const int self = emu::lane();
    const unsigned below = self < 32 ? 0u : ((1u << (self - 32)) - 1u);
    return acc + static_cast<unsigned>(__builtin_popcount(mask & below));
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }

// ---- atomics (device and LDS alike)
#define __HIP_MEMORY_SCOPE_WAVEFRONT 1
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3
#define __HIP_MEMORY_SCOPE_SYSTEM 4
template <typename T>
inline T __hip_atomic_load(const T* p, int, int) {
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}
template <typename T>
inline void __hip_atomic_store(T* p, T v, int, int) {
    __atomic_store_n(p, v, __ATOMIC_SEQ_CST);
}
template <typename T>
inline T __hip_atomic_fetch_add(T* p, T v, int, int) {
    if constexpr (std::is_integral_v<T>) {
        return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
    } else {
        T old;
        __atomic_load(p, &old, __ATOMIC_SEQ_CST);
        T desired = old + v;
        while (!__atomic_compare_exchange(p, &old, &desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
            desired = old + v;
        return old;
    }
}
inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline int atomicCAS(int* p, int cmp, int val) {
    __atomic_compare_exchange_n(p, &cmp, val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return cmp;
}
inline int atomicMax(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}

// ---- runtime
typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorNotReady = 600;
constexpr hipError_t hipErrorUnknown = 999;
constexpr hipError_t hipErrorInvalidValue = 1;
constexpr hipError_t hipErrorOutOfMemory = 2;
constexpr hipError_t hipErrorMemoryAllocation = 2;
constexpr hipError_t hipErrorNotSupported = 801;
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_pool* hipMemPool_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold };
constexpr unsigned hipStreamNonBlocking = 1;
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;

struct hipDeviceProp_t {
    char gcnArchName[256];
    int multiProcessorCount;
    std::size_t sharedMemPerBlock;
    int clockRate;
    std::size_t totalGlobalMem;
};

inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int);
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = static_cast<hipStream_t>(::emu::new_stream());
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
    *e = reinterpret_cast<hipEvent_t>(new ::emu::Event());
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) { return hipEventCreateWithFlags(e, 0); }
inline ::emu::Event* emu_ev(hipEvent_t e) { return reinterpret_cast<::emu::Event*>(e); }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    ::emu::record_on(s, emu_ev(e));
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    *ms = 0.f;
    return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    ::emu::wait_event_on(s, emu_ev(e));
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t s) {
    ::emu::drain_stream(s);
    return hipSuccess;
}
// the device runs when the host looks: a query completes the queried work
inline hipError_t hipStreamQuery(hipStream_t s) {
    ::emu::drain_stream(s);
    return hipSuccess;
}
inline hipError_t hipEventQuery(hipEvent_t e) { return ::emu::event_done(emu_ev(e)) ? hipSuccess : hipErrorNotReady; }
inline hipError_t hipEventSynchronize(hipEvent_t e) {
    ::emu::drain_event(emu_ev(e));
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
    ::emu::drain_stream(s);
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete emu_ev(e);
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() {
    ::emu::drain_all();
    return hipSuccess;
}
inline hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t*, int) { return hipErrorNotSupported; }
inline hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
inline hipError_t hipMalloc(void** p, std::size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T>
inline hipError_t hipMalloc(T** p, std::size_t n) {
    return hipMalloc(reinterpret_cast<void**>(p), n);
}
inline hipError_t hipFree(void* p) {  // frees wait for the device, as hipFree does
    ::emu::drain_all();
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMallocAsync(void** p, std::size_t n, hipStream_t) { return hipMalloc(p, n); }
inline hipError_t hipFreeAsync(void* p, hipStream_t s) {
    ::emu::submit_on(s, [p] { std::free(p); });
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, std::size_t n, unsigned) { return hipMalloc(p, n); }
// the per-thread default stream (a stand-in handle like the real one)
#define hipStreamPerThread (reinterpret_cast<hipStream_t>(2))
inline hipError_t hipHostFree(void* p) { return hipFree(p); }
// synchronous copies and fills run on the legacy stream: after the work of
// the blocking (default) streams
inline hipError_t hipMemcpy(void* d, const void* s, std::size_t n, hipMemcpyKind) {
    ::emu::drain_blocking();
    if (n) std::memcpy(d, s, n);
    return hipSuccess;
}
// host threads share one address space: a __device__ variable's address is
// the host object's address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
inline hipError_t hipMemset(void* d, int v, std::size_t n) {
    ::emu::drain_blocking();
    std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, std::size_t n, hipMemcpyKind, hipStream_t st) {
    ::emu::submit_on(st, [=] {
        if (n) std::memcpy(d, s, n);
    });
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, std::size_t n, hipStream_t st) {
    ::emu::submit_on(st, [=] { std::memset(d, v, n); });
    return hipSuccess;
}

// graphs: the work recorded between begin and end of a capture, replayed in order
struct emu_graph_exec {
    std::vector<::emu::Work> work;
};
typedef std::vector<::emu::Work>* hipGraph_t;
typedef emu_graph_exec* hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (::emu::capture_target()) return hipErrorInvalidValue;
    ::emu::capture_target() = new std::vector<::emu::Work>;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    *g = ::emu::capture_target();
    ::emu::capture_target() = nullptr;
    return *g ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, std::size_t) {
    *e = new emu_graph_exec{*g};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    ::emu::submit_on(s, [e] {
        for (auto& w : e->work) w();
    });
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
template <typename F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, std::size_t) {
    *n = 1;
    return hipSuccess;
}

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                           \
    ::emu::submit_on(                                                                         \
        stream,                                                                               \
        [=, emu_grid_ = dim3(grid), emu_block_ = dim3(block)]() {                             \
            ::emu::launch(emu_grid_, emu_block_, [&]() { kernel(__VA_ARGS__); });             \
        },                                                                                    \
        #kernel)
