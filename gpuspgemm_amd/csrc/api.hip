// Host driver and C ABI of the gfx950 SpGEMM library.
//
// proposal_inner (src/libproposal/include/proposal/proposal.cuh:26-80)
// creates eight streams, allocates its scratch and synchronises with the
// host at least ten times per multiply.  Here a lane (a main stream, two
// aux streams, a scratch arena that only grows, pinned host words for the
// read-backs) is kept per calling thread or caller stream, outputs come
// from a device block cache, and a multiply is
//
//   nip (with the symbolic binning)
//   -> sym_multi_small on the main stream || sym_max, sym_multi on two aux
//      streams (forked and joined with events)
//   -> scan_count, scan_apply
//   -> one kernel copies the counters to pinned host memory, zeroes them
//      and signals; the host polls the signal (one synchronisation)
//   -> C.col / C.val of the exact size from the block cache
//   -> one symnum launch per non-empty fused bin over the three streams,
//      with no stream waits (the host has seen the scan complete), each
//      stream ending with a kernel that signals the host.
//
// Rows that overflow the largest LDS table add a second round trip
// (global tables, then the scan again), as the reference's fail bin does
// (sym1.cuh:424-454).
#include "spg_kernels.hpp"
#include "../../include/gpuspgemm.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace spg {
namespace {

// Scratch of spg_multiply_noalloc: a fixed 256 MiB device array that every
// GPU the library runs on holds until the process ends.  Host code never
// passes `scratch` itself to a kernel or a copy (in host code the name
// stands for a host-side shadow, not a device address): each device's
// context looks the array's address on that device up once, at set-up,
// with hipGetSymbolAddress, and uses only that address.
constexpr size_t kNoAllocScratch = size_t(256) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratch];

// Products whose mean NIP per row, estimated from the operands' densities,
// is at most this many queue the symbolic pass's small-table part first
// (power-law 10^6 rows: 8.6, measured 0.325 -> 0.305 ms; the 96^3 stencil,
// 704, keeps the other order: 3.9 ms against 4.4 with this one).
constexpr double kSmallFirstNip = 32;

// Tile count (of 2048 rows) from which the scan stage scans the tile sums
// once (k_tile_prefix, one workgroup: ~2 us per 1024 tiles) instead of
// letting every workgroup sum those before it (tiles^2 / 2 reads: 2 M at
// 2048 tiles, 2 * 10^9 at 2^27 rows).
constexpr int kTilePrefixMin = 2048;

thread_local std::string g_last_error;

spg_status fail(spg_status s, const std::string& msg) {
    g_last_error = msg;
    return s;
}

spg_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
    std::string msg = std::string("HIP error (") + file + ":" + std::to_string(line) + "): " +
                      hipGetErrorString(e) + " in " + what;
    return fail(e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? SPG_ERR_OUT_OF_MEMORY : SPG_ERR_HIP,
                msg);
}

#define SPG_HIP(expr)                                                       \
    do {                                                                    \
        const hipError_t e_ = (expr);                                       \
        if (e_ != hipSuccess) return hip_fail(e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define SPG_TRY(expr)                       \
    do {                                    \
        const spg_status s_ = (expr);       \
        if (s_ != SPG_OK) return s_;        \
    } while (0)

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- logging
//
// The reference logs the bin tables, max nnz, nnz(C) and the fail-bin size
// at debug level through spdlog (binning.cuh:181-187, 232-233,
// sym1.cuh:430).  Levels here follow spdlog's order; the default is off.
// The level comes from spg_set_log_level or, before the first call, from
// the SPG_LOG_LEVEL environment variable (a name or a number).
enum LogLevel { kTrace = 0, kDebug = 1, kInfo = 2, kWarn = 3, kError = 4, kCritical = 5, kOff = 6 };
std::atomic<int> g_log_level{-1};

int log_level() {
    int lv = g_log_level.load(std::memory_order_relaxed);
    if (lv >= 0) return lv;
    lv = kOff;
    if (const char* env = std::getenv("SPG_LOG_LEVEL")) {
        static const char* names[] = {"trace", "debug", "info", "warn", "error", "critical", "off"};
        for (int i = 0; i < 7; ++i)
            if (std::strcmp(env, names[i]) == 0) lv = i;
        if (env[0] >= '0' && env[0] <= '6' && env[1] == 0) lv = env[0] - '0';
    }
    int expected = -1;
    g_log_level.compare_exchange_strong(expected, lv);
    return g_log_level.load();
}

bool log_on(int lv) { return log_level() <= lv; }

void log_msg(int lv, const char* fmt, ...) {
    if (!log_on(lv)) return;
    static std::mutex mu;
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> g(mu);
    std::fprintf(stderr, "[gpuspgemm %s] %s\n", lv <= kTrace ? "trace" : lv == kDebug ? "debug" : "info", buf);
}

struct ProfRec {
    std::string name;
    int64_t rows;
    int grid;
    hipEvent_t ev0, ev1;
    double ms;
};

// Instantiated graphs of one phase of a multiply, keyed by everything that
// fixes the phase's launches: sizes, bin counts, tables and the addresses
// of the lane's buffers.  A signature is captured the second time it is
// seen (a product shape run once never pays for a capture); the cache keeps
// the most recently used graphs.
struct GraphCache {
    static constexpr size_t kEntries = 64;
    static constexpr size_t kSeen = 256;
    struct Entry {
        std::vector<int64_t> sig;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    // a fixed array: the handles stay in place for the life of the process
    Entry entries[kEntries];
    size_t n = 0;
    std::deque<std::vector<int64_t>> seen;
    uint64_t clock = 0;

    hipGraphExec_t find(const std::vector<int64_t>& sig) {
        for (size_t i = 0; i < n; ++i)
            if (entries[i].sig == sig) {
                entries[i].used = ++clock;
                return entries[i].exec;
            }
        return nullptr;
    }
    // true when `sig` was seen before; records it otherwise
    bool seen_before(const std::vector<int64_t>& sig) {
        if (std::find(seen.begin(), seen.end(), sig) != seen.end()) return true;
        seen.push_back(sig);
        if (seen.size() > kSeen) seen.pop_front();
        return false;
    }
    void insert(std::vector<int64_t> sig, hipGraphExec_t exec) {
        size_t at = n;
        if (n == kEntries) {  // replace the least recently used
            at = 0;
            for (size_t i = 1; i < n; ++i)
                if (entries[i].used < entries[at].used) at = i;
            (void)hipGraphExecDestroy(entries[at].exec);
        } else {
            ++n;
        }
        seen.erase(std::remove(seen.begin(), seen.end(), sig), seen.end());
        entries[at] = Entry{std::move(sig), exec, ++clock};
    }
    void clear() {
        for (size_t i = 0; i < n; ++i) {
            (void)hipGraphExecDestroy(entries[i].exec);
            entries[i] = Entry{};
        }
        n = 0;
        seen.clear();
    }
};

// ---------------------------------------------------------------- block cache
//
// Device memory of the outputs (C.rpt, C.col, C.val) and of the per-call
// global-table pools.  The stream-ordered pool (hipMallocAsync) took 3 us
// on average and up to 80 us of host time per allocation on the multiply's
// critical path, between the read-back of nnz(C) and the first fused
// launch; here a released block goes back to a cache and is handed out
// again once the GPU work queued before its release has completed (an
// event recorded at the release, queried at reuse).  Blocks are rounded up
// to eight size classes per power of two, and a request takes a ready block
// of at most twice its size.  When hipMalloc fails the cache waits for the
// device, frees its idle blocks and tries again.
struct BlockCache {
    struct Fence {
        hipEvent_t ev = nullptr;
    };
    struct Idle {
        void* p;
        std::shared_ptr<Fence> fence;  // null: no GPU work can still use the block
    };
    std::mutex mu;
    std::multimap<size_t, Idle> idle;              // by size
    std::unordered_map<void*, size_t> live;        // handed-out block -> size
    std::vector<hipEvent_t> spare;                 // events of fences no block refers to any more
    size_t idle_bytes = 0, live_bytes = 0;

    static size_t round_up(size_t b) {
        size_t p2 = 4096;
        while (p2 < b) p2 <<= 1;
        const size_t g = std::max<size_t>(4096, p2 / 8);
        return (b + g - 1) / g * g;
    }

    void drop_fence(std::shared_ptr<Fence>& f) {  // mu held
        if (f && f.use_count() == 1 && f->ev) spare.push_back(f->ev);
        f.reset();
    }

    spg_status alloc(void** out, size_t bytes) {
        *out = nullptr;
        if (bytes == 0) return SPG_OK;
        const size_t want = round_up(bytes);
        std::lock_guard<std::mutex> g(mu);
        for (auto it = idle.lower_bound(want); it != idle.end() && it->first <= 2 * want; ++it) {
            Idle& b = it->second;
            if (b.fence && b.fence->ev) {
                const hipError_t q = hipEventQuery(b.fence->ev);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) continue;  // leave it; a later call finds out
            }
            *out = b.p;
            live[b.p] = it->first;
            live_bytes += it->first;
            idle_bytes -= it->first;
            drop_fence(b.fence);
            idle.erase(it);
            return SPG_OK;
        }
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
            (void)hipGetLastError();
            (void)hipDeviceSynchronize();
            trim_locked();
            e = hipMalloc(&p, want);
        }
        if (e != hipSuccess) return hip_fail(e, "hipMalloc (output block)", __FILE__, __LINE__);
        live[p] = want;
        live_bytes += want;
        *out = p;
        return SPG_OK;
    }

    // Back to the cache.  With a stream, the block is reused only after the
    // work queued on that stream so far has completed.
    spg_status release(void* p, hipStream_t after, bool use_stream) {
        if (!p) return SPG_OK;
        std::lock_guard<std::mutex> g(mu);
        auto it = live.find(p);
        if (it == live.end()) return fail(SPG_ERR_INVALID_ARGUMENT, "pointer was not allocated by the library");
        const size_t bytes = it->second;
        std::shared_ptr<Fence> fence;
        if (use_stream) {
            fence = std::make_shared<Fence>();
            if (!spare.empty()) {
                fence->ev = spare.back();
                spare.pop_back();
            } else {
                SPG_HIP(hipEventCreateWithFlags(&fence->ev, hipEventDisableTiming));
            }
            SPG_HIP(hipEventRecord(fence->ev, after));
        }
        live.erase(it);
        live_bytes -= bytes;
        idle.emplace(bytes, Idle{p, std::move(fence)});
        idle_bytes += bytes;
        return SPG_OK;
    }

    // Several blocks released behind one event.
    spg_status release_all(void* const* ps, int n, hipStream_t after) {
        std::lock_guard<std::mutex> g(mu);
        std::shared_ptr<Fence> fence = std::make_shared<Fence>();
        if (!spare.empty()) {
            fence->ev = spare.back();
            spare.pop_back();
        } else {
            SPG_HIP(hipEventCreateWithFlags(&fence->ev, hipEventDisableTiming));
        }
        SPG_HIP(hipEventRecord(fence->ev, after));
        for (int i = 0; i < n; ++i) {
            if (!ps[i]) continue;
            auto it = live.find(ps[i]);
            if (it == live.end()) return fail(SPG_ERR_INVALID_ARGUMENT, "pointer was not allocated by the library");
            live_bytes -= it->second;
            idle_bytes += it->second;
            idle.emplace(it->second, Idle{ps[i], fence});
            live.erase(it);
        }
        return SPG_OK;
    }

    void trim_locked() {  // device idle
        for (auto& kv : idle) {
            drop_fence(kv.second.fence);
            (void)hipFree(kv.second.p);
        }
        idle.clear();
        idle_bytes = 0;
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
        spare.clear();
    }
};

// ---------------------------------------------------------------- lanes

// Pinned host words a stream's last kernel fills: the fused pass's flags
// word and the completion flag, on a line of their own.
struct alignas(64) DoneWord {
    int flag;
    int flags;
};

// The device resources of one multiply in flight: aux streams that the
// symbolic phase forks onto and the fused-pass launches are spread over,
// events for the forks and joins, a scratch arena that only grows, and
// pinned host memory for the read-backs.  The main stream is the caller's
// (a thread's default stream for the synchronous calls) or the lane's own.
struct Lane {
    static constexpr int kAux = 2;
    static constexpr int kStreams = kAux + 1;
    std::mutex mu;  // one call at a time
    hipStream_t own = nullptr;
    hipStream_t aux[kAux] = {};
    hipEvent_t ev_ready = nullptr;            // fork point on the main stream
    hipEvent_t ev_done[kAux] = {};            // join points on the aux streams
    DoneWord* h_done = nullptr;               // pinned, kStreams words
    char* scratch = nullptr;
    size_t scratch_cap = 0;
    bool meta_clean = false;  // the counters at the start of scratch are zero
    char* gpool = nullptr;    // global-bin tables of the fused pass under graph replay
    size_t gpool_cap = 0;
    DevMeta* h_meta = nullptr;
    char* h_stage = nullptr;
    size_t h_stage_cap = 0;
    GraphCache graphs[2];  // symbolic phase, fused pass
};

spg_status lane_create(Lane& L) {
    SPG_HIP(hipStreamCreateWithFlags(&L.own, hipStreamNonBlocking));
    for (int i = 0; i < Lane::kAux; ++i) {
        SPG_HIP(hipStreamCreateWithFlags(&L.aux[i], hipStreamNonBlocking));
        SPG_HIP(hipEventCreateWithFlags(&L.ev_done[i], hipEventDisableTiming));
    }
    SPG_HIP(hipEventCreateWithFlags(&L.ev_ready, hipEventDisableTiming));
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_done), sizeof(DoneWord) * Lane::kStreams,
                          hipHostMallocDefault));
    std::memset(L.h_done, 0, sizeof(DoneWord) * Lane::kStreams);
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_meta), sizeof(DevMeta), hipHostMallocDefault));
    return SPG_OK;
}

void lane_destroy(Lane& L) {
    if (L.own) (void)hipStreamSynchronize(L.own);
    for (int i = 0; i < Lane::kAux; ++i) {
        if (L.aux[i]) (void)hipStreamSynchronize(L.aux[i]);
        if (L.aux[i]) (void)hipStreamDestroy(L.aux[i]);
        if (L.ev_done[i]) (void)hipEventDestroy(L.ev_done[i]);
    }
    if (L.ev_ready) (void)hipEventDestroy(L.ev_ready);
    L.graphs[0].clear();
    L.graphs[1].clear();
    if (L.h_done) (void)hipHostFree(L.h_done);
    if (L.scratch) (void)hipFree(L.scratch);
    if (L.gpool) (void)hipFree(L.gpool);
    if (L.h_meta) (void)hipHostFree(L.h_meta);
    if (L.h_stage) (void)hipHostFree(L.h_stage);
    if (L.own) (void)hipStreamDestroy(L.own);
    L.own = nullptr;
    for (auto& a : L.aux) a = nullptr;
    L.h_done = nullptr;
    L.scratch = nullptr;
    L.scratch_cap = 0;
    L.gpool = nullptr;
    L.gpool_cap = 0;
    L.h_meta = nullptr;
    L.h_stage = nullptr;
    L.h_stage_cap = 0;
}

// Completion of the work queued on some of a lane's streams.  The host
// clears a word of pinned host memory per stream (lane_arm), a kernel
// queued last on that stream stores 1 into it (k_readback, eagerly or as a
// graph's last node) and the host polls the words (lane_poll): the
// runtime's own completion of an event or stream was measured at ~55 us
// after the GPU finished.  The streams are queried now and then so that a
// failed stream ends the wait with its error.  Every signal is polled for
// before the lane is armed again, so no earlier store can land after the
// clear.
void lane_arm(Lane& L, int q) {
    L.h_done[q].flags = 0;
    __atomic_store_n(&L.h_done[q].flag, 0, __ATOMIC_RELEASE);
}

hipError_t lane_poll(Lane& L, int mask, const hipStream_t* streams) {
    // The streams are queried every 100 us of waiting: a query costs
    // microseconds of host time (up to 11 us in an API trace), and one in
    // progress when the word arrives delays the return.
    constexpr double kQueryEvery = 100e-6;
    double next_query = now_s() + kQueryEvery;
    for (unsigned spin = 1;; ++spin) {
        bool all = true;
        for (int q = 0; q < Lane::kStreams; ++q)
            if ((mask >> q & 1) && __atomic_load_n(&L.h_done[q].flag, __ATOMIC_ACQUIRE) != 1) all = false;
        if (all) return hipSuccess;
        if ((spin & 255) == 0 && now_s() >= next_query) {
            next_query = now_s() + kQueryEvery;
            // a stream that has completed without storing its word failed;
            // the others keep being waited for
            for (int q = 0; q < Lane::kStreams; ++q) {
                if (!(mask >> q & 1) || __atomic_load_n(&L.h_done[q].flag, __ATOMIC_ACQUIRE) == 1) continue;
                const hipError_t e = hipStreamQuery(streams[q]);
                if (e == hipSuccess && __atomic_load_n(&L.h_done[q].flag, __ATOMIC_ACQUIRE) != 1) return hipErrorUnknown;
                if (e != hipSuccess && e != hipErrorNotReady) return e;
            }
        }
    }
}

// ---------------------------------------------------------------- context

struct Context {
    std::atomic<bool> ready{false};
    int device = -1;
    int cus = 0;
    int lds = 0;
    int clock_khz = 0;
    size_t global_mem = 0;
    std::string arch;
    std::mutex mu;  // tables, the lane map, profiling records
    BinTable sym{};
    BinTable num{};
    int occ_sym_multi[kSymParts] = {1, 1, 1}, occ_sym_max = 1;
    int occ_num[2][kNumLdsVariants] = {};
    // tiles from which the scan reads each tile's prefix off one scanned
    // array instead of summing the earlier tiles per workgroup (quadratic
    // in the tile count); SPG_TILE_PREFIX_MIN overrides it
    int tile_prefix_min = kTilePrefixMin;
    Lane lane;  // graph replay and profiling run here, one call at a time
    Lane noalloc_lane;             // spg_multiply_noalloc, one call at a time
    char* noalloc_base = nullptr;  // this device's address of `scratch`
    std::unordered_map<uint64_t, std::unique_ptr<Lane>> lanes;  // per caller stream
    std::vector<std::unique_ptr<Lane>> free_lanes;              // lanes of threads that ended
    BlockCache cache;
    bool graphs = false;  // replay repeated phases from captured graphs
    std::atomic<int64_t> n_captured{0}, n_replayed{0};
    bool profiling = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> events;
    size_t events_used = 0;
};

// Initialisation, graph capture and profiling, and the synchronous helpers
// (which must not run beside a capture) take this lock; plain multiplies
// from different threads or streams do not.
std::recursive_mutex g_mutex;
constexpr int kMaxDevices = 64;
// Contexts live until the process ends and are never destroyed: their
// lanes and cached blocks stay reachable, and no HIP call runs after the
// runtime has shut down.
std::atomic<Context*> g_ctx[kMaxDevices];

// ---------------------------------------------------------------- bin tables

struct TunedBin {
    int variant, nip, nnz;
};
// kTunedSym / kTunedNum: the default bin tables, written by the tuner
// (gpuspgemm_amd/tune.py) from measurements on the target GPU.
#include "tuned_tables.inc"

template <size_t N>
void fill_table(BinTable& t, const TunedBin (&rows)[N], bool sym_phase) {
    t = BinTable{};
    t.n = static_cast<int>(N);
    for (int b = 0; b < t.n; ++b) {
        const int v = rows[b].variant;
        t.variant[b] = v;
        t.rows[b] = sym_phase ? kSymVariants[v].r : (v < kNumGlobalVariant ? kNumVariants[v].r : 1);
        t.nip_cap[b] = rows[b].nip;
        t.nnz_cap[b] = rows[b].nnz;
    }
}

void default_tables(BinTable& sym, BinTable& num) {
    static_assert(sizeof(kTunedSym) / sizeof(kTunedSym[0]) <= kMaxBins, "too many symbolic bins");
    static_assert(sizeof(kTunedNum) / sizeof(kTunedNum[0]) <= kMaxBins, "too many fused bins");
    fill_table(sym, kTunedSym, true);
    fill_table(num, kTunedNum, false);
}

// A symbolic table is a run of multi-bin variants followed by the max
// variant with caps INT_MAX.  Each multi bin's NIP cap must stay below its
// table size, so a row's distinct columns (<= NIP) never fill the table.
spg_status check_sym_table(const BinTable& t) {
    if (t.n < 1 || t.n > kMaxBins) return fail(SPG_ERR_INVALID_ARGUMENT, "symbolic table: bad bin count");
    for (int b = 0; b + 1 < t.n; ++b) {
        const int v = t.variant[b];
        if (v < 0 || v >= kSymMultiVariants)
            return fail(SPG_ERR_INVALID_ARGUMENT, "symbolic table: bins before the last must use the multi-bin variants");
        if (t.nip_cap[b] >= kSymVariants[v].t || t.nip_cap[b] < 1)
            return fail(SPG_ERR_INVALID_ARGUMENT, "symbolic table: NIP cap must be below the variant's table size");
    }
    if (t.variant[t.n - 1] != kSymMaxVariant || t.nip_cap[t.n - 1] != INT_MAX)
        return fail(SPG_ERR_INVALID_ARGUMENT, "symbolic table: last bin must be the largest-table variant with cap INT_MAX");
    return SPG_OK;
}

spg_status check_num_table(const BinTable& t) {
    if (t.n < 1 || t.n > kMaxBins) return fail(SPG_ERR_INVALID_ARGUMENT, "fused table: bad bin count");
    for (int b = 0; b < t.n; ++b) {
        const int v = t.variant[b];
        if (v < 0 || v > kNumGlobalVariant) return fail(SPG_ERR_INVALID_ARGUMENT, "fused table: bad variant");
        if (v < kNumGlobalVariant && (t.nnz_cap[b] > kNumVariants[v].p || t.nnz_cap[b] < 1))
            return fail(SPG_ERR_INVALID_ARGUMENT, "fused table: nnz cap above the variant's capacity");
    }
    if (t.variant[t.n - 1] != kNumGlobalVariant || t.nnz_cap[t.n - 1] != INT_MAX || t.nip_cap[t.n - 1] != INT_MAX)
        return fail(SPG_ERR_INVALID_ARGUMENT, "fused table: last bin must be the global variant with caps INT_MAX");
    return SPG_OK;
}

std::string table_text(const BinTable& t) {
    std::ostringstream out;
    for (int b = 0; b < t.n; ++b)
        out << (b ? ", " : "") << "v" << t.variant[b] << " nip<=" << t.nip_cap[b] << " nnz<=" << t.nnz_cap[b];
    return out.str();
}

spg_status ensure_context(Context*& out) {
    int dev = 0;
    SPG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SPG_ERR_UNSUPPORTED_DEVICE, "device index out of range");
    Context* cp = g_ctx[dev].load(std::memory_order_acquire);
    if (cp && cp->ready.load(std::memory_order_acquire)) {
        out = cp;
        return SPG_OK;
    }
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    cp = g_ctx[dev].load(std::memory_order_acquire);
    if (!cp) {
        cp = new Context();
        g_ctx[dev].store(cp, std::memory_order_release);
    }
    Context& c = *cp;
    out = cp;
    if (c.ready.load(std::memory_order_acquire)) return SPG_OK;
    hipDeviceProp_t prop;
    SPG_HIP(hipGetDeviceProperties(&prop, dev));
    c.arch = prop.gcnArchName;
    if (c.arch.rfind("gfx950", 0) != 0)
        return fail(SPG_ERR_UNSUPPORTED_DEVICE, "unsupported device architecture " + c.arch + " (built for gfx950)");
    c.device = dev;
    c.cus = prop.multiProcessorCount;
    c.lds = static_cast<int>(prop.sharedMemPerBlock);
    c.clock_khz = prop.clockRate;
    c.global_mem = prop.totalGlobalMem;
    if (c.lds < kNumBigLds + 1024 ||
        c.lds < kSymVariants[kSymMaxVariant].t * 4 + 8 * kSymVariants[kSymMaxVariant].ec + 1024)
        return fail(SPG_ERR_UNSUPPORTED_DEVICE, "device offers less LDS per workgroup than the kernels need");
    SPG_TRY(lane_create(c.lane));
    SPG_TRY(lane_create(c.noalloc_lane));
    // the fixed scratch of spg_multiply_noalloc on this device
    SPG_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&c.noalloc_base), HIP_SYMBOL(scratch)));
    default_tables(c.sym, c.num);
    SPG_TRY(check_sym_table(c.sym));  // a stale tuned_tables.inc fails here, loudly
    SPG_TRY(check_num_table(c.num));
    for (int part = 0; part < kSymParts; ++part) c.occ_sym_multi[part] = occupancy_sym_multi<double>(part);
    c.occ_sym_max = occupancy_sym_max<double>();
    if (const char* env = std::getenv("SPG_TILE_PREFIX_MIN")) c.tile_prefix_min = std::max(1, std::atoi(env));
    for (int v = 0; v < kNumLdsVariants; ++v) {
        c.occ_num[0][v] = occupancy_symnum_bin<float>(v);
        c.occ_num[1][v] = occupancy_symnum_bin<double>(v);
    }
    log_msg(kDebug, "context on device %d (%s, %d CUs); symbolic bins: %s; fused bins: %s", dev, c.arch.c_str(),
            c.cus, table_text(c.sym).c_str(), table_text(c.num).c_str());
    c.ready.store(true, std::memory_order_release);
    return SPG_OK;
}

// A thread's lane for its synchronous calls goes back to the context's
// spare list when the thread ends, so that threads that come and go reuse
// lanes instead of adding new ones.
struct ThreadLanes {
    std::vector<std::pair<Context*, uint64_t>> held;
    ~ThreadLanes() {
        for (auto& [c, key] : held) {
            std::lock_guard<std::mutex> g(c->mu);
            auto it = c->lanes.find(key);
            if (it == c->lanes.end()) continue;
            c->free_lanes.push_back(std::move(it->second));
            c->lanes.erase(it);
        }
    }
};

uint64_t thread_key() {
    static std::atomic<uint64_t> next{1};
    thread_local const uint64_t id = next.fetch_add(1);
    return (id << 1) | 1;  // odd: never a stream handle
}

// The lane of a caller stream, or of the calling thread when `stream` is
// null (its default stream).
spg_status lane_for(Context& c, hipStream_t stream, Lane*& out) {
    const uint64_t key = stream ? reinterpret_cast<uint64_t>(stream) : thread_key();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.lanes.find(key);
    if (it != c.lanes.end()) {
        out = it->second.get();
        return SPG_OK;
    }
    std::unique_ptr<Lane> L;
    if (!c.free_lanes.empty()) {
        L = std::move(c.free_lanes.back());
        c.free_lanes.pop_back();
    } else {
        L = std::make_unique<Lane>();
        const spg_status st = lane_create(*L);
        if (st != SPG_OK) {
            lane_destroy(*L);
            return st;
        }
    }
    out = L.get();
    c.lanes.emplace(key, std::move(L));
    if (!stream) {
        thread_local ThreadLanes mine;
        mine.held.emplace_back(&c, key);
    }
    return SPG_OK;
}

spg_status ensure_scratch(Lane& L, size_t bytes, hipStream_t s) {
    if (bytes <= L.scratch_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(s));
    for (hipStream_t a : L.aux) SPG_HIP(hipStreamSynchronize(a));
    if (L.scratch) SPG_HIP(hipFree(L.scratch));
    L.scratch = nullptr;
    const size_t cap = std::max(bytes, L.scratch_cap * 3 / 2);
    L.scratch_cap = 0;
    L.meta_clean = false;
    SPG_HIP(hipMalloc(reinterpret_cast<void**>(&L.scratch), cap));
    L.scratch_cap = cap;
    return SPG_OK;
}

spg_status ensure_gpool(Lane& L, size_t bytes, hipStream_t s) {
    if (bytes <= L.gpool_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(s));
    for (hipStream_t a : L.aux) SPG_HIP(hipStreamSynchronize(a));
    if (L.gpool) SPG_HIP(hipFree(L.gpool));
    L.gpool = nullptr;
    const size_t cap = std::max(bytes, L.gpool_cap * 3 / 2);
    L.gpool_cap = 0;
    SPG_HIP(hipMalloc(reinterpret_cast<void**>(&L.gpool), cap));
    L.gpool_cap = cap;
    return SPG_OK;
}

spg_status ensure_stage(Lane& L, size_t bytes, hipStream_t s) {
    if (bytes <= L.h_stage_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(s));
    if (L.h_stage) SPG_HIP(hipHostFree(L.h_stage));
    L.h_stage = nullptr;
    L.h_stage_cap = 0;
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_stage), bytes, hipHostMallocDefault));
    L.h_stage_cap = bytes;
    return SPG_OK;
}

// ---------------------------------------------------------------- profiling

struct Prof {
    Context& c;
    bool on;
    int open = -1;
    Prof(Context& ctx, bool enabled) : c(ctx), on(enabled) {}
    hipEvent_t next_event() {
        if (c.events_used == c.events.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c.events.push_back(e);
        }
        return c.events[c.events_used++];
    }
    hipStream_t st = nullptr;
    void begin(const std::string& name, int64_t rows, int grid, hipStream_t where) {
        if (log_on(kTrace))
            log_msg(kTrace, "launch %s: %lld rows, grid %d", name.c_str(), static_cast<long long>(rows), grid);
        if (!on) return;
        st = where;
        ProfRec r{name, rows, grid, next_event(), next_event(), 0.0};
        if (r.ev0) (void)hipEventRecord(r.ev0, st);
        c.prof.push_back(r);
        open = static_cast<int>(c.prof.size()) - 1;
    }
    void end() {
        if (!on || open < 0) return;
        if (c.prof[open].ev1) (void)hipEventRecord(c.prof[open].ev1, st);
        open = -1;
    }
};

void prof_reset(Context& c) {
    c.prof.clear();
    c.events_used = 0;
}

void prof_collect(Context& c) {
    for (auto& r : c.prof) {
        float ms = 0.f;
        if (r.ev0 && r.ev1 && hipEventElapsedTime(&ms, r.ev0, r.ev1) == hipSuccess) r.ms = ms;
    }
}

// ---------------------------------------------------------------- multiply

spg_status check_operands(const spg_csr& A, const spg_csr& B) {
    if (A.m < 0 || A.n < 0 || A.nnz < 0 || B.m < 0 || B.n < 0 || B.nnz < 0)
        return fail(SPG_ERR_INVALID_ARGUMENT, "negative matrix size");
    if (A.n != B.m)
        return fail(SPG_ERR_INVALID_ARGUMENT, "Matrix A columns (" + std::to_string(A.n) +
                                                  ") must match matrix B rows (" + std::to_string(B.m) + ")");
    if (!A.rpt || !B.rpt) return fail(SPG_ERR_INVALID_ARGUMENT, "null row pointer array");
    if ((A.nnz > 0 && (!A.col || !A.val)) || (B.nnz > 0 && (!B.col || !B.val)))
        return fail(SPG_ERR_INVALID_ARGUMENT, "null column or value array");
    return SPG_OK;
}

// How a call's work is ordered and waited for.
struct CallMode {
    hipStream_t main = nullptr;  // the stream the multiply is ordered on
    bool own_stream = false;     // `main` is the lane's own stream
    bool events = false;         // fused pass forked and joined with events, one read-back at the end (graph capture)
    bool async = false;          // return once the fused pass is queued; C is ready in `main`'s order
                                 // (single products only: no staged descriptors outlive the call)
    // No allocation at all (spg_multiply_noalloc): scratch and the global
    // tables come from `fixed` (fixed_cap bytes), C's arrays from the
    // caller (out_cap entries of C.col / C.val).
    char* fixed = nullptr;
    size_t fixed_cap = 0;
    long long out_cap = 0;
};

// One batch of products on one lane, in two phases:
//   symbolic()        NIP, binning, exact counts, scan (row pointers) and
//                     the fused-pass bins; ends with one read-back, after
//                     which nnz(C_i) of every product is known on the host
//   launch_numeric()  takes C.col / C.val of the exact sizes from the block
//                     cache and queues the fused pass (no wait)
//   finish()          waits for the streams and checks the consistency flag.
// A plain call runs the three back to back; the server admits a batch by
// output_bytes() between the first two.
struct JobBase {
    virtual ~JobBase() = default;
    virtual spg_status symbolic(int n, const spg_csr* A, const spg_csr* B, spg_csr* out) = 0;
    virtual spg_status launch_numeric() = 0;
    virtual spg_status finish() = 0;
    virtual void abort() = 0;
    virtual size_t output_bytes() const = 0;
};

template <typename V>
struct Job : JobBase {
    Context& c;
    Lane& L;
    Prof prof;
    CallMode mode;
    hipStream_t s = nullptr;
    int count = 0;
    spg_csr* C = nullptr;
    BinTable sym{}, num{};
    std::vector<void*> owned;
    std::vector<Prob<V>> probs;
    std::vector<Tile> tiles;
    ProbSet<V> ps{};
    long long total_rows = 0;
    bool small_first = false;  // the symbolic launches in the order of short products (front())
    int n_tiles = 0;
    bool batch = false;
    bool empty = false;
    DevMeta* meta = nullptr;
    long long* tile_sums = nullptr;
    long long* inst_nnz = nullptr;
    int2* items1 = nullptr;
    RowItem* items2 = nullptr;
    int2* fail_items = nullptr;
    long long* fail_off = nullptr;
    long long* gpool_off = nullptr;
    Prob<V>* d_probs = nullptr;
    Tile* d_tiles = nullptr;
    size_t stage_probs = 0, stage_tiles = 0, stage_inst = 0;
    // counters read back after the symbolic phase (DevMeta itself is
    // over-aligned; a Job is deleted through its base)
    struct {
        int32_t flags = 0;
        int32_t sym_count[kMaxBins] = {};
        int32_t num_count[kMaxBins] = {};
        int32_t fail_count = 0;
        int32_t max_nip = 0, max_nnz = 0;
        unsigned long long total_nip = 0;
        unsigned long long gnum_pool = 0;
        long long nnz_total = 0;
    } hm;
    std::vector<long long> nnz;
    void* num_pool = nullptr;   // from the block cache (eager) or the lane's pool (graph replay)
    bool graphs = false;        // phases may be replayed from the lane's graph caches
    bool signal_queued = false; // the fused pass already queued its completion signals
    int done_mask = 0;          // streams whose completion words the fused pass fills
    size_t scratch_end = 0;     // bytes of the scratch layout
    int32_t* out_col = nullptr; // the caller's C.col / C.val (spg_multiply_noalloc)
    void* out_val = nullptr;
    double t_start = 0, t_sym = 0;

    // Only the context's own lane captures graphs: its calls run under the
    // library mutex, which the library's synchronous helpers (spg_memcpy,
    // spg_malloc, spg_free, spg_device_sync) take too.  HIP rejects a
    // synchronous call on the legacy stream from any thread while some
    // stream of the device is capturing, so lanes of other threads and a
    // server's worker threads never capture.
    Job(Context& ctx, Lane& lane, const BinTable& sym_table, const BinTable& num_table, bool profile, CallMode m)
        : c(ctx), L(lane), prof(ctx, profile), mode(m), s(m.main), sym(sym_table), num(num_table),
          graphs(&lane == &ctx.lane && ctx.graphs && !profile) {
        if (graphs) mode.events = true;
    }

    hipStream_t stream_of(int q) const { return q == 0 ? s : L.aux[q - 1]; }

    void abort() override {
        // kernels queued before the failure may still use the blocks
        if (!owned.empty()) (void)c.cache.release_all(owned.data(), static_cast<int>(owned.size()), s);
        owned.clear();
        if (num_pool && !graphs && !mode.fixed) (void)c.cache.release(num_pool, s, true);
        num_pool = nullptr;
        L.meta_clean = false;
        if (mode.fixed) {
            // the caller's arrays stay with the caller; wait for anything
            // queued on them before returning
            (void)hipStreamSynchronize(s);
            for (hipStream_t a : L.aux) (void)hipStreamSynchronize(a);
            return;
        }
        for (int i = 0; i < count; ++i) C[i] = spg_csr{0, 0, 0, nullptr, nullptr, nullptr};
    }

    spg_status failed(spg_status st) {
        abort();
        return st;
    }

    size_t output_bytes() const override {
        size_t b = 0;
        for (long long n : nnz) b += static_cast<size_t>(n) * (sizeof(int32_t) + sizeof(V));
        return b;
    }

    // The stream `to` waits for the work queued on `from` so far.  (Waits
    // on a value in memory, hipStreamWaitValue64, are served by a spinning
    // blit kernel on this runtime; measured, it slowed the concurrent
    // stencil count pass from 1.36 to 2.2 ms, so joins use events.)
    hipError_t join(hipStream_t from, hipEvent_t ev, hipStream_t to) {
        hipError_t e = hipEventRecord(ev, from);
        return e == hipSuccess ? hipStreamWaitEvent(to, ev, 0) : e;
    }

    spg_status front() {
        // The counters are zeroed by the symbolic phase's read-back; after
        // anything else (first use, grown scratch, an error or an abandoned
        // job) they are cleared here.
        if (graphs || !L.meta_clean) {
            prof.begin("clear", 0, 1, s);
            launch_clear(meta, 0, s);
            prof.end();
        }
        L.meta_clean = false;
        prof.begin("nip", total_rows, n_tiles, s);
        launch_nip<V>(ps, sym, meta, items1, static_cast<int>(total_rows), s);
        prof.end();
        // The launches count disjoint rows, on three streams: the part of the
        // two largest multi-bin tables (rows of a workgroup each, few of
        // them) on the first aux stream; with the small-table part in use,
        // the mid-table part on the second aux stream (its workgroups take
        // the CUs that the small-table grid leaves as its last units drain);
        // the remaining part on the main stream.  The largest-table bin
        // (139 KB of LDS a workgroup) goes where its workgroups find free
        // CUs.  Rows of short products (small_first: a mean NIP per row of
        // at most kSmallFirstNip, estimated from the operands' densities)
        // fill the small-table part, the longest launch: it is queued first,
        // right behind k_nip on the main stream, and the largest-table bin
        // after it there.  Otherwise the largest-table bin is queued first,
        // on the first aux stream, while the CUs are still free (its
        // workgroups without a row exit at once; queued behind the
        // multi-bin grids they were measured to hold back the mid-table
        // launch of the stencil by a third), and its part follows it.
        const int n_multi = sym.n - 1;
        bool part_used[kSymParts] = {};
        for (int b = 0; b < n_multi; ++b) part_used[sym_part(sym.variant[b])] = true;
        const bool split = part_used[0] && part_used[1];
        static const char* const kPartName[kSymParts] = {"sym_multi_small", "sym_multi", "sym_multi_large"};
        auto part_stream = [&](int part) { return part == 2 ? L.aux[0] : (split && part == 1 ? L.aux[1] : s); };
        auto queue_part = [&](int part) {
            const hipStream_t where = part_stream(part);
            int grid = c.cus * c.occ_sym_multi[part];
            grid = (grid + kXcds - 1) / kXcds * kXcds;
            prof.begin(kPartName[part], total_rows, grid, where);
            launch_sym_multi<V>(part, ps, items1, static_cast<int>(total_rows), sym, n_multi, meta, grid, where);
            prof.end();
        };
        auto queue_max = [&](hipStream_t where) {
            const int grid = c.cus * c.occ_sym_max;
            prof.begin("sym_max", 0, grid, where);
            launch_sym_max<V>(ps, items1, static_cast<int>(total_rows), sym, sym.n - 1, meta, fail_items, fail_off,
                              grid, where);
            prof.end();
        };
        SPG_HIP(hipEventRecord(L.ev_ready, s));
        int main_part = -1;  // the part queued on the main stream
        for (int part : {0, 1})
            if (main_part < 0 && part_used[part] && part_stream(part) == s) main_part = part;
        if (small_first && main_part >= 0) queue_part(main_part);
        SPG_HIP(hipStreamWaitEvent(L.aux[0], L.ev_ready, 0));
        if (!small_first) queue_max(L.aux[0]);
        if (split) SPG_HIP(hipStreamWaitEvent(L.aux[1], L.ev_ready, 0));
        for (int part : {2, 1, 0})
            if (part_used[part] && !(small_first && part == main_part)) queue_part(part);
        if (small_first) queue_max(s);
        SPG_HIP(join(L.aux[0], L.ev_done[0], s));
        if (split) SPG_HIP(join(L.aux[1], L.ev_done[1], s));
        return SPG_OK;
    }

    // scan stage and the read-back of the counters (zeroing them for the
    // next multiply), ending with the completion signal of the main stream
    // (the caller armed it and polls it)
    spg_status scan(int after_fail) {
        prof.begin("scan_count", total_rows, n_tiles, s);
        launch_scan_count<V>(ps, num, meta, tile_sums, after_fail, s);
        prof.end();
        prof.begin("scan_apply", total_rows, n_tiles, s);
        launch_scan_apply<V>(ps, num, meta, tile_sums, items2, gpool_off, inst_nnz, after_fail, c.tile_prefix_min,
                             s);
        prof.end();
        // counters (and nnz of every product of a batch) to pinned host memory, then the signal
        launch_readback(reinterpret_cast<const int*>(meta), reinterpret_cast<int*>(L.h_meta),
                        static_cast<int>(sizeof(DevMeta) / sizeof(int)), inst_nnz,
                        reinterpret_cast<long long*>(L.h_stage + stage_inst), batch ? count : 0, meta,
                        &L.h_done[0].flag, s);
        SPG_HIP(hipGetLastError());
        return SPG_OK;
    }

    // Signature of a phase: every value its launches depend on.  Buffer
    // contents (operand descriptors, tiles) are copied from pinned host
    // memory by the phase itself and so are not part of it.
    std::vector<int64_t> signature(int phase) const {
        std::vector<int64_t> k{phase, static_cast<int64_t>(sizeof(V)), count, n_tiles, total_rows, small_first,
                               reinterpret_cast<int64_t>(L.scratch), reinterpret_cast<int64_t>(L.h_stage)};
        const BinTable& t = phase == 0 ? sym : num;
        k.push_back(t.n);
        for (int b = 0; b < t.n; ++b) {
            k.push_back(t.variant[b]);
            k.push_back(t.nip_cap[b]);
            k.push_back(t.nnz_cap[b]);
            if (phase == 1) k.push_back(hm.num_count[b]);
        }
        if (phase == 1) k.push_back(reinterpret_cast<int64_t>(num_pool));
        return k;
    }

    // Queue one phase.  `record` queues its work on the lane's streams,
    // ending with the completion signal.  With graph replay on, a phase
    // whose signature was seen before is captured into a graph once and
    // replayed from then on: one launch in place of a dozen.
    template <typename F>
    spg_status queue_phase(int phase, F&& record) {
        if (!graphs) return record();
        std::vector<int64_t> sig = signature(phase);
        GraphCache& gc = L.graphs[phase];
        hipGraphExec_t exec = gc.find(sig);
        if (exec) {
            ++c.n_replayed;
        } else {
            if (!gc.seen_before(sig)) return record();
            SPG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            spg_status st = record();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (st == SPG_OK && e != hipSuccess) st = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
            if (st == SPG_OK) {
                const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                if (ei != hipSuccess) st = hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
            }
            if (g) (void)hipGraphDestroy(g);
            SPG_TRY(st);
            gc.insert(std::move(sig), exec);
            ++c.n_captured;
        }
        SPG_HIP(hipGraphLaunch(exec, s));
        return SPG_OK;
    }

    spg_status symbolic(int n, const spg_csr* A, const spg_csr* B, spg_csr* out) override {
        t_start = now_s();
        count = n;
        C = out;
        batch = count > 1;
        if (mode.fixed) {  // the caller's output arrays (a single product)
            out_col = C[0].col;
            out_val = C[0].val;
        }
        for (int i = 0; i < count; ++i) {
            SPG_TRY(check_operands(A[i], B[i]));
            C[i] = mode.fixed ? spg_csr{0, A[i].m, B[i].n, C[i].rpt, C[i].col, C[i].val}
                              : spg_csr{0, A[i].m, B[i].n, nullptr, nullptr, nullptr};
        }
        // output row pointers
        for (int i = 0; i < count; ++i) {
            if (!mode.fixed) {
                void* p = nullptr;
                const spg_status st = c.cache.alloc(&p, sizeof(int32_t) * (static_cast<size_t>(A[i].m) + 1));
                if (st != SPG_OK) return failed(st);
                owned.push_back(p);
                C[i].rpt = static_cast<int32_t*>(p);
            }
            total_rows += A[i].m;
            n_tiles += (A[i].m + kTileRows - 1) / kTileRows;
            if (A[i].m == 0) {
                const hipError_t e = hipMemsetAsync(C[i].rpt, 0, sizeof(int32_t), s);
                if (e != hipSuccess) return failed(hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__));
            }
        }
        nnz.assign(count, 0);
        {
            double est_nip = 0;  // NIP of all rows if B's entries were spread evenly over its rows
            for (int i = 0; i < count; ++i)
                if (B[i].m > 0) est_nip += static_cast<double>(A[i].nnz) * B[i].nnz / B[i].m;
            small_first = est_nip <= kSmallFirstNip * static_cast<double>(total_rows);
        }
        // graph replay reads the descriptors from device memory in every call
        batch = count > 1 || graphs;
        if (total_rows >= INT_MAX) return failed(fail(SPG_ERR_OVERFLOW, "total row count exceeds int32"));
        if (n_tiles == 0) {  // every product has zero rows
            empty = true;
            return SPG_OK;
        }
        const size_t M = static_cast<size_t>(total_rows);

        // scratch layout
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t at = off;
            off = align256(off + bytes);
            return at;
        };
        const size_t o_meta = take(sizeof(DevMeta));
        const size_t o_tsum = take(sizeof(long long) * n_tiles);
        const size_t o_inst = take(sizeof(long long) * count);
        const size_t o_nip = take(sizeof(int32_t) * M);
        size_t nnz_a = 0;
        for (int i = 0; i < count; ++i) nnz_a += static_cast<size_t>(A[i].nnz);
        const size_t o_bref = take(sizeof(long long) * nnz_a);
        const size_t o_it1 = take(sizeof(int2) * M * sym.n);  // a region of M items per symbolic bin
        const size_t o_it2 = take(sizeof(RowItem) * M);
        const size_t o_fit = take(sizeof(int2) * M);
        const size_t o_foff = take(sizeof(long long) * M);
        const size_t o_goff = take(sizeof(long long) * M);
        // descriptors: a table per stream of the fused pass, each copied on
        // its own stream (the eager fused pass has no stream waits)
        const size_t o_probs = take(batch ? sizeof(Prob<V>) * count * Lane::kStreams : 0);
        const size_t o_tiles = take(batch ? sizeof(Tile) * n_tiles : 0);
        scratch_end = off;
        char* base = nullptr;
        if (mode.fixed) {
            if (off > mode.fixed_cap)
                return failed(fail(SPG_ERR_OUT_OF_MEMORY,
                                   "the product needs " + std::to_string(off) + " bytes of scratch; the fixed scratch "
                                   "of spg_multiply_noalloc holds " + std::to_string(mode.fixed_cap)));
            base = mode.fixed;
        } else {
            const spg_status st = ensure_scratch(L, off, s);
            if (st != SPG_OK) return failed(st);
            base = L.scratch;
        }
        meta = reinterpret_cast<DevMeta*>(base + o_meta);
        tile_sums = reinterpret_cast<long long*>(base + o_tsum);
        inst_nnz = reinterpret_cast<long long*>(base + o_inst);
        int32_t* nip = reinterpret_cast<int32_t*>(base + o_nip);
        long long* bref = reinterpret_cast<long long*>(base + o_bref);
        items1 = reinterpret_cast<int2*>(base + o_it1);
        items2 = reinterpret_cast<RowItem*>(base + o_it2);
        fail_items = reinterpret_cast<int2*>(base + o_fit);
        fail_off = reinterpret_cast<long long*>(base + o_foff);
        gpool_off = reinterpret_cast<long long*>(base + o_goff);
        d_probs = reinterpret_cast<Prob<V>*>(base + o_probs);
        d_tiles = reinterpret_cast<Tile*>(base + o_tiles);

        // descriptors
        probs.assign(count, Prob<V>{});
        long long row_base = 0;
        size_t entry_base = 0;
        for (int i = 0; i < count; ++i) {
            Prob<V>& p = probs[i];
            p.a_rpt = A[i].rpt;
            p.a_col = A[i].col;
            p.a_val = static_cast<const V*>(A[i].val);
            p.b_rpt = B[i].rpt;
            p.b_col = B[i].col;
            p.b_val = static_cast<const V*>(B[i].val);
            p.c_rpt = C[i].rpt;
            p.c_col = nullptr;
            p.c_val = nullptr;
            p.nip = nip + row_base;
            p.bref = bref + entry_base;
            p.m = A[i].m;
            p.n = B[i].n;
            if (batch) {
                const int first = static_cast<int>(tiles.size());
                for (int r0 = 0; r0 < A[i].m; r0 += kTileRows)
                    tiles.push_back(Tile{i, r0, std::min(kTileRows, A[i].m - r0), first});
            }
            row_base += A[i].m;
            entry_base += static_cast<size_t>(A[i].nnz);
        }
        ps.p0 = probs[0];
        ps.n_inst = count;
        ps.n_tiles = n_tiles;
        ps.many = batch ? d_probs : nullptr;
        ps.tiles = batch ? d_tiles : nullptr;
        if (batch) {
            stage_probs = 0;
            stage_tiles = align256(sizeof(Prob<V>) * count);
            stage_inst = stage_tiles + align256(sizeof(Tile) * n_tiles);
            const spg_status st = ensure_stage(L, stage_inst + align256(sizeof(long long) * count), s);
            if (st != SPG_OK) return failed(st);
            std::memcpy(L.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);
            std::memcpy(L.h_stage + stage_tiles, tiles.data(), sizeof(Tile) * n_tiles);
        }

        lane_arm(L, 0);
        spg_status st = queue_phase(0, [&]() -> spg_status {
            if (batch) {
                SPG_HIP(hipMemcpyAsync(d_probs, L.h_stage + stage_probs, sizeof(Prob<V>) * count,
                                       hipMemcpyHostToDevice, s));
                SPG_HIP(hipMemcpyAsync(d_tiles, L.h_stage + stage_tiles, sizeof(Tile) * n_tiles,
                                       hipMemcpyHostToDevice, s));
            }
            SPG_TRY(front());
            return scan(0);
        });
        hipStream_t main_only[Lane::kStreams];
        for (hipStream_t& m : main_only) m = s;
        if (st == SPG_OK) {
            const hipError_t e = lane_poll(L, 1, main_only);
            if (e != hipSuccess) st = hip_fail(e, "symbolic phase", __FILE__, __LINE__);
        }
        if (st == SPG_OK && (L.h_meta->flags.v & kFlagMismatch)) {
            // left by the fused pass of an asynchronous multiply queued earlier on this stream
            st = fail(SPG_ERR_INTERNAL, "the fused pass of an earlier multiply on this stream found a different "
                                        "column count than its count pass (are B's rows free of duplicate "
                                        "columns and indices in range?)");
        }
        if (st == SPG_OK && L.h_meta->fail_count.v > 0) {
            // rows that overflowed the largest LDS table: count them again in
            // global memory, then redo the scan stage
            const int fails = L.h_meta->fail_count.v;
            hm.fail_count = fails;
            void* sym_pool = nullptr;
            const size_t pool_bytes = sizeof(int32_t) * L.h_meta->fail_pool.v;
            if (mode.fixed) {  // after the scratch layout, in the fixed array
                if (scratch_end + pool_bytes > mode.fixed_cap)
                    st = fail(SPG_ERR_OUT_OF_MEMORY, "the global count tables of the heaviest rows do not fit in the "
                                                     "fixed scratch of spg_multiply_noalloc");
                else
                    sym_pool = mode.fixed + scratch_end;
            } else {
                st = c.cache.alloc(&sym_pool, pool_bytes);
            }
            if (st == SPG_OK) {
                prof.begin("sym_global", fails, fails, s);
                launch_sym_global<V>(ps, fail_items, fail_off, static_cast<int*>(sym_pool), fails, s);
                prof.end();
                launch_clear(meta, kMetaScanWord, s);
                lane_arm(L, 0);
                st = scan(1);
                if (st == SPG_OK) {
                    const hipError_t e = lane_poll(L, 1, main_only);
                    if (e != hipSuccess) st = hip_fail(e, "symbolic phase", __FILE__, __LINE__);
                }
            }
            // the main stream has completed (or failed): the pool is idle, or
            // waits for the stream
            if (sym_pool && !mode.fixed) (void)c.cache.release(sym_pool, s, st != SPG_OK);
        }
        if (st != SPG_OK) return failed(st);
        hm.flags = L.h_meta->flags.v;
        for (int b = 0; b < kMaxBins; ++b) {
            hm.num_count[b] = L.h_meta->num_count[b].v;
            hm.sym_count[b] = L.h_meta->sym_count[b].v;
        }
        hm.gnum_pool = L.h_meta->gnum_pool.v;
        hm.nnz_total = L.h_meta->nnz_total.v;
        hm.max_nip = L.h_meta->max_nip.v;
        hm.max_nnz = L.h_meta->max_nnz.v;
        hm.total_nip = L.h_meta->total_nip.v;
        if (hm.flags & kFlagNipOverflow)
            return failed(fail(SPG_ERR_OVERFLOW, "a row has 2^31 or more intermediate products"));
        if (hm.flags & kFlagNnzOverflow) return failed(fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32"));
        long long all_nnz = 0;
        for (int i = 0; i < count; ++i) {
            if (A[i].m == 0) continue;
            nnz[i] = batch ? reinterpret_cast<long long*>(L.h_stage + stage_inst)[i] : hm.nnz_total;
            if (nnz[i] >= INT_MAX) return failed(fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32"));
            all_nnz += nnz[i];
        }
        t_sym = now_s();
        if (log_on(kDebug)) {
            std::ostringstream sb, nb;
            for (int b = 0; b < sym.n; ++b) sb << (b ? " " : "") << hm.sym_count[b];
            for (int b = 0; b < num.n; ++b) nb << (b ? " " : "") << hm.num_count[b];
            log_msg(kDebug,
                    "symbolic phase: %d product(s), %lld rows, NIP %llu (max %d per row), nnz(C) %lld (max %d per "
                    "row), rows per symbolic bin [%s], fail bin %d, rows per fused bin [%s], host %.1f us",
                    count, total_rows, hm.total_nip, hm.max_nip, all_nnz, hm.max_nnz, sb.str().c_str(), hm.fail_count,
                    nb.str().c_str(), 1e6 * (t_sym - t_start));
        }
        return SPG_OK;
    }

    spg_status launch_numeric() override {
        if (empty) return SPG_OK;
        // exact-size outputs
        if (mode.fixed) {  // the caller's arrays (a single product)
            C[0].nnz = static_cast<int32_t>(nnz[0]);
            if (nnz[0] > mode.out_cap)
                return failed(fail(SPG_ERR_OUT_OF_MEMORY, "nnz(C) = " + std::to_string(nnz[0]) +
                                                              " exceeds the output capacity " +
                                                              std::to_string(mode.out_cap)));
            C[0].col = out_col;
            C[0].val = out_val;
            probs[0].c_col = out_col;
            probs[0].c_val = static_cast<V*>(out_val);
        }
        for (int i = 0; i < count && !mode.fixed; ++i) {
            if (C[i].m == 0) continue;
            C[i].nnz = static_cast<int32_t>(nnz[i]);
            void *pc = nullptr, *pv = nullptr;
            spg_status st = c.cache.alloc(&pc, sizeof(int32_t) * nnz[i]);
            if (st == SPG_OK) owned.push_back(pc);
            if (st == SPG_OK) st = c.cache.alloc(&pv, sizeof(V) * nnz[i]);
            if (st == SPG_OK) owned.push_back(pv);
            if (st != SPG_OK) return failed(st);
            C[i].col = static_cast<int32_t*>(pc);
            C[i].val = pv;
            probs[i].c_col = C[i].col;
            probs[i].c_val = static_cast<V*>(pv);
        }
        ps.p0 = probs[0];
        if (batch) std::memcpy(L.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);

        // fused pass: one launch per non-empty bin, or per merged set of
        // bins, heaviest first, spread over the lane's streams so that
        // launches overlap
        struct Launch {
            int variant, off, count;
        };
        std::vector<Launch> launches;
        int glob_bin = -1, glob_off = 0;
        NumMultiBins mb[2] = {};  // bins of the two merged sets
        int set_units[2] = {0, 0};
        {
            int item_off = 0;
            for (int b = 0; b < num.n; ++b) {
                const int cnt = hm.num_count[b];
                const int v = num.variant[b];
                const int set = num_multi_set(v);
                if (cnt > 0) {
                    if (v == kNumGlobalVariant) {
                        glob_bin = b;
                        glob_off = item_off;
                    } else if (set >= 0) {
                        NumMultiBins& m = mb[set];
                        const int units = (cnt + kNumVariants[v].r - 1) / kNumVariants[v].r;
                        m.variant[m.n] = v;
                        m.off[m.n] = item_off;
                        m.count[m.n] = cnt;
                        m.slice0[m.n + 1] = m.slice0[m.n] + (units + kXcds - 1) / kXcds;
                        ++m.n;
                        set_units[set] += units;
                    } else {
                        launches.push_back(Launch{v, item_off, cnt});
                    }
                }
                item_off += cnt;
            }
            for (int set = 0; set < 2; ++set) {
                NumMultiBins& m = mb[set];
                // one bin, or set 1 holding a large bin: launches of their own
                if (m.n == 1 || (set == 1 && set_units[1] > kNumMultiMaxUnits1)) {
                    for (int i = 0; i < m.n; ++i) launches.push_back(Launch{m.variant[i], m.off[i], m.count[i]});
                    m.n = 0;
                }
            }
        }
        // Stream plan: the single launches in order of decreasing row
        // capacity times lanes per row, dealt round robin over the lane's
        // streams but one per merged launch.  The heavy-row launches (few
        // workgroups, each a long chain) are queued first, so their
        // workgroups reach the CUs before the merged launches fill them, and
        // then run beside them; set 1 (longer rows) is queued before set 0.
        // (Measured against a longest-first plan: this one was faster.)
        std::stable_sort(launches.begin(), launches.end(), [](const Launch& x, const Launch& y) {
            return kNumVariants[x.variant].p * kNumVariants[x.variant].wg / kNumVariants[x.variant].r >
                   kNumVariants[y.variant].p * kNumVariants[y.variant].wg / kNumVariants[y.variant].r;
        });
        const int merged_sets[2] = {1, 0};  // queue order
        int n_merged = 0;
        for (int set : merged_sets) n_merged += mb[set].n > 0 ? 1 : 0;
        std::vector<int> stream_idx(launches.size());
        int n_streams = 1;
        int merged_q[2] = {-1, -1};
        size_t merged_at = 0;  // the merged launches are queued before single launch merged_at
        {
            const int n_deal = Lane::kStreams - n_merged;
            int k = glob_bin >= 0 ? 1 : 0;  // the global bin, when present, is queued first on the main stream
            for (size_t i = 0; i < launches.size(); ++i, ++k) stream_idx[i] = k % n_deal;
            int q = std::min(k, n_deal);
            for (int set : merged_sets)
                if (mb[set].n > 0) merged_q[set] = q++;
            merged_at = std::min(launches.size(), static_cast<size_t>(n_deal));
            n_streams = std::max(q, 1);
        }
        if (glob_bin >= 0) {
            const size_t pool_bytes = sizeof(int32_t) * hm.gnum_pool;
            spg_status st = SPG_OK;
            if (graphs) {  // a buffer of the lane, so that its address recurs
                st = ensure_gpool(L, pool_bytes, s);
                num_pool = L.gpool;
            } else if (mode.fixed) {  // after the scratch layout, in the fixed array
                if (scratch_end + pool_bytes > mode.fixed_cap)
                    st = fail(SPG_ERR_OUT_OF_MEMORY, "the global tables of the heaviest rows do not fit in the fixed "
                                                     "scratch of spg_multiply_noalloc");
                else
                    num_pool = mode.fixed + scratch_end;
            } else {
                st = c.cache.alloc(&num_pool, pool_bytes);
            }
            if (st != SPG_OK) return failed(st);
        }
        // Eagerly, no stream waits at all: the host has seen the scan stage
        // complete (the read-back it polled was queued after it), and C and
        // the pool come from the block cache, not from stream-ordered
        // allocations.  Each stream ends with a kernel that copies the
        // flags word and signals the host; an asynchronous call instead
        // joins the aux streams back to the main stream.  Under graph
        // capture the aux streams fork from and join the main stream with
        // events, and one read-back ends the phase.
        done_mask = 0;
        if (!mode.events && !mode.async)
            for (int q = 0; q < n_streams; ++q) {
                lane_arm(L, q);
                done_mask |= 1 << q;
            }
        if (mode.events) lane_arm(L, 0);
        const spg_status st = queue_phase(1, [&]() -> spg_status {
            hipError_t e = hipSuccess;
            // A batch reads the descriptors (with C's addresses) from device
            // memory.  Under graph capture one table is copied on the main
            // stream and the aux streams fork after it; eagerly every stream
            // copies its own table before its first launch.
            ProbSet<V> ps_q[Lane::kStreams];
            for (int q = 0; q < Lane::kStreams; ++q) {
                ps_q[q] = ps;
                if (q > 0 && batch && !mode.events) ps_q[q].many = d_probs + static_cast<size_t>(q) * count;
            }
            if (batch)
                e = hipMemcpyAsync(d_probs, L.h_stage + stage_probs, sizeof(Prob<V>) * count, hipMemcpyHostToDevice, s);
            if (e == hipSuccess && mode.events && n_streams > 1) e = hipEventRecord(L.ev_ready, s);
            bool aux_ready[Lane::kStreams] = {true};  // the others false
            auto stream_for = [&](int q) -> hipStream_t {
                if (!aux_ready[q]) {
                    if (e == hipSuccess && mode.events) e = hipStreamWaitEvent(L.aux[q - 1], L.ev_ready, 0);
                    if (e == hipSuccess && !mode.events && batch)
                        e = hipMemcpyAsync(d_probs + static_cast<size_t>(q) * count, L.h_stage + stage_probs,
                                           sizeof(Prob<V>) * count, hipMemcpyHostToDevice, L.aux[q - 1]);
                    aux_ready[q] = true;
                }
                return stream_of(q);
            };
            if (e == hipSuccess && glob_bin >= 0) {
                const hipStream_t on = stream_of(0);
                prof.begin("symnum_global", hm.num_count[glob_bin], hm.num_count[glob_bin], on);
                launch_symnum_global<V>(ps, items2, gpool_off, glob_off, hm.num_count[glob_bin],
                                        static_cast<int*>(num_pool), meta, on);
                prof.end();
            }
            auto queue_merged = [&]() {
                for (int set : merged_sets) {
                    if (e != hipSuccess || mb[set].n == 0) continue;
                    const hipStream_t on = stream_for(merged_q[set]);
                    if (e != hipSuccess) return;
                    int rows = 0;
                    for (int b = 0; b < mb[set].n; ++b) rows += mb[set].count[b];
                    prof.begin("symnum_multi" + std::to_string(set), rows, kXcds * mb[set].slice0[mb[set].n], on);
                    e = launch_symnum_multi<V>(set, ps_q[merged_q[set]], items2, mb[set], meta, on);
                    prof.end();
                }
            };
            for (size_t li = 0; li < launches.size(); ++li) {
                if (li == merged_at) queue_merged();
                if (e != hipSuccess) break;
                const Launch& Lh = launches[li];
                const NumVariant& nv = kNumVariants[Lh.variant];
                const int units = (Lh.count + nv.r - 1) / nv.r;
                const int grid = (units + kXcds - 1) / kXcds * kXcds;  // one unit per workgroup
                const hipStream_t on = stream_for(stream_idx[li]);
                if (e != hipSuccess) break;
                prof.begin("symnum_v" + std::to_string(Lh.variant), Lh.count, grid, on);
                e = launch_symnum_bin<V>(Lh.variant, ps_q[stream_idx[li]], items2, Lh.off, Lh.count, meta, grid, on);
                prof.end();
            }
            if (merged_at == launches.size()) queue_merged();
            if (e == hipSuccess) e = hipGetLastError();
            if (e == hipSuccess && mode.events) {
                for (int i = 1; i < n_streams && e == hipSuccess; ++i) e = join(L.aux[i - 1], L.ev_done[i - 1], s);
                // the flags word back and the signal
                if (e == hipSuccess)
                    launch_readback(&meta->flags.v, &L.h_done[0].flags, 1, nullptr, nullptr, 0, nullptr,
                                    &L.h_done[0].flag, s);
            } else if (e == hipSuccess && mode.async) {
                for (int i = 1; i < n_streams && e == hipSuccess; ++i) e = join(L.aux[i - 1], L.ev_done[i - 1], s);
            } else if (e == hipSuccess) {
                for (int q = 0; q < n_streams; ++q)
                    launch_readback(&meta->flags.v, &L.h_done[q].flags, 1, nullptr, nullptr, 0, nullptr,
                                    &L.h_done[q].flag, stream_of(q));
            }
            if (e == hipSuccess) e = hipGetLastError();
            return e == hipSuccess ? SPG_OK : hip_fail(e, "fused pass", __FILE__, __LINE__);
        });
        if (st != SPG_OK) return failed(st);
        if (mode.events) done_mask = 1;
        signal_queued = true;
        return SPG_OK;
    }

    spg_status finish() override {
        if (empty) {
            owned.clear();
            return SPG_OK;
        }
        hipStream_t streams[Lane::kStreams];
        for (int q = 0; q < Lane::kStreams; ++q) streams[q] = stream_of(q);
        if (mode.async) {
            // C is complete once the main stream reaches this point; the pool
            // waits for it too
            if (num_pool && !graphs && !mode.fixed) (void)c.cache.release(num_pool, s, true);
            num_pool = nullptr;
            L.meta_clean = true;  // zeroed by the read-back; a mismatch flag is found by the next call
            owned.clear();
            return SPG_OK;
        }
        hipError_t e = hipSuccess;
        if (!signal_queued) {
            lane_arm(L, 0);
            launch_signal(&L.h_done[0].flag, s);
            e = hipGetLastError();
            done_mask = 1;
        }
        if (e == hipSuccess) e = lane_poll(L, done_mask, streams);
        if (e != hipSuccess) return failed(hip_fail(e, "fused pass", __FILE__, __LINE__));
        if (num_pool && !graphs && !mode.fixed) (void)c.cache.release(num_pool, nullptr, false);
        num_pool = nullptr;
        int flags = 0;
        for (int q = 0; q < Lane::kStreams; ++q)
            if (done_mask >> q & 1) flags |= L.h_done[q].flags;
        if (flags & kFlagMismatch)
            return failed(fail(SPG_ERR_INTERNAL, "fused pass found a different column count than the count pass "
                                                 "(are B's rows free of duplicate columns and indices in range?)"));
        L.meta_clean = true;  // zeroed by the symbolic read-back, untouched since
        owned.clear();        // the outputs now belong to the caller
        if (log_on(kDebug))
            log_msg(kDebug, "fused pass done: host %.1f us after the read-back, %.1f us in all",
                    1e6 * (now_s() - t_sym), 1e6 * (now_s() - t_start));
        return SPG_OK;
    }
};

template <typename V>
spg_status multiply_impl(Context& c, Lane& L, const BinTable& sym, const BinTable& num, bool profile, CallMode mode,
                         int count, const spg_csr* A, const spg_csr* B, spg_csr* C) {
    if (profile) prof_reset(c);
    Job<V> job(c, L, sym, num, profile, mode);
    SPG_TRY(job.symbolic(count, A, B, C));
    SPG_TRY(job.launch_numeric());
    SPG_TRY(job.finish());
    if (profile) prof_collect(c);
    return SPG_OK;
}

// A multiply from a host thread: on its own lane (per caller stream, or per
// thread for the synchronous calls, whose stream is the thread's default
// stream) without the library lock; with graph replay or profiling on, on
// the context's lane under the lock.
spg_status multiply_entry(int count, const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype,
                          hipStream_t stream, bool async) {
    if (count < 1 || !A || !B || !C) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    if (dtype != SPG_F32 && dtype != SPG_F64) return fail(SPG_ERR_INVALID_ARGUMENT, "unknown dtype");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    g_last_error.clear();
    auto run = [&](Lane& L, const BinTable& sym, const BinTable& num, bool profile, CallMode mode) {
        return dtype == SPG_F64 ? multiply_impl<double>(*c, L, sym, num, profile, mode, count, A, B, C)
                                : multiply_impl<float>(*c, L, sym, num, profile, mode, count, A, B, C);
    };
    {
        std::unique_lock<std::recursive_mutex> lock(g_mutex);
        if (c->graphs || c->profiling) {
            // the context lane's own stream, ordered after the caller's stream
            CallMode mode{c->lane.own, true, false, false};
            hipStream_t caller = stream ? stream : hipStreamPerThread;
            SPG_HIP(hipEventRecord(c->lane.ev_ready, caller));
            SPG_HIP(hipStreamWaitEvent(c->lane.own, c->lane.ev_ready, 0));
            const spg_status st = run(c->lane, c->sym, c->num, c->profiling, mode);
            if (st == SPG_OK && async) SPG_HIP(hipStreamSynchronize(c->lane.own));
            return st;
        }
    }
    Lane* L = nullptr;
    SPG_TRY(lane_for(*c, stream, L));
    BinTable sym, num;
    {
        std::lock_guard<std::mutex> g(c->mu);
        sym = c->sym;
        num = c->num;
    }
    std::lock_guard<std::mutex> lk(L->mu);
    CallMode mode{stream ? stream : hipStreamPerThread, false, false, async};
    return run(*L, sym, num, false, mode);
}

// A single product into caller-provided C arrays with the device's fixed
// scratch array: no device or pinned allocation inside the call.  Calls on
// one device take turns on the no-alloc lane; the array keeps the last
// call's contents, and the job clears the counters that must start at zero.
spg_status multiply_noalloc(const spg_csr* A, const spg_csr* B, spg_csr* C, int64_t capacity, spg_dtype dtype) {
    if (!A || !B || !C) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    if (dtype != SPG_F32 && dtype != SPG_F64) return fail(SPG_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (capacity < 0) return fail(SPG_ERR_INVALID_ARGUMENT, "negative output capacity");
    if (!C->rpt) return fail(SPG_ERR_INVALID_ARGUMENT, "C->rpt must hold A->m + 1 row pointers");
    if (capacity > 0 && (!C->col || !C->val))
        return fail(SPG_ERR_INVALID_ARGUMENT, "C->col and C->val must hold `capacity` entries");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    g_last_error.clear();
    if (!c->noalloc_base) return fail(SPG_ERR_INTERNAL, "the fixed scratch of this device has no address");
    BinTable sym, num;
    {
        std::lock_guard<std::mutex> g(c->mu);
        sym = c->sym;
        num = c->num;
    }
    Lane& L = c->noalloc_lane;
    std::lock_guard<std::mutex> lk(L.mu);
    CallMode mode{hipStreamPerThread, false, false, false};
    mode.fixed = c->noalloc_base;
    mode.fixed_cap = kNoAllocScratch;
    mode.out_cap = capacity;
    return dtype == SPG_F64 ? multiply_impl<double>(*c, L, sym, num, false, mode, 1, A, B, C)
                            : multiply_impl<float>(*c, L, sym, num, false, mode, 1, A, B, C);
}

// ---------------------------------------------------------------- server
struct Request {
    int64_t ticket = 0;
    spg_csr A{}, B{};
    spg_dtype dtype = SPG_F64;
    spg_csr C{};
    spg_status st = SPG_OK;
    std::string err;
    double t_submit = 0, t_done = 0;
    bool done = false;
};

// One admitted batch on one lane.  C stays in place while the job runs.
struct Flight {
    int lane = -1;
    std::unique_ptr<JobBase> job;
    std::vector<std::shared_ptr<Request>> reqs;
    std::vector<spg_csr> A, B, C;
    size_t bytes = 0;
};

struct Server {
    static constexpr int kLanes = 3;
    // A batch gathers requests up to this many nonzeros of A in total; a
    // larger request runs alone, so it never holds small ones back.
    static constexpr long long kBatchNnz = 4 << 20;
    Context* c = nullptr;
    size_t budget = 0;
    int max_batch = 1;
    Lane lanes[kLanes];
    std::mutex mu;
    std::condition_variable cv_work;  // requests queued / stop
    std::condition_variable cv_done;  // a request finished
    std::condition_variable cv_mem;   // an admitted batch released its output budget
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> requests;
    int64_t next_ticket = 1;
    bool stop = false;
    size_t in_flight = 0;  // output bytes of admitted batches still running
    int admitted = 0;      // such batches
    std::thread workers[kLanes];
    int64_t n_batches = 0, n_served = 0, max_seen = 0, n_deferred = 0;
    size_t peak_bytes = 0;  // largest in_flight seen

    void publish(Flight& f, spg_status st, const std::string& err) {
        const double t = now_s();
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < f.reqs.size(); ++i) {
            Request& r = *f.reqs[i];
            r.st = st;
            r.err = err;
            if (st == SPG_OK) r.C = f.C[i];
            r.t_done = t;
            r.done = true;
        }
        n_served += static_cast<int64_t>(f.reqs.size());
        cv_done.notify_all();
    }

    // One worker per lane: take a batch, run its symbolic phase, wait until
    // its exact output fits the budget beside the batches the other lanes
    // are running, run the fused pass, publish.  Lanes work concurrently, so
    // a small batch never waits behind the symbolic phase of a large one.
    void run(int lane) {
        (void)hipSetDevice(c->device);
        Lane& L = lanes[lane];
        int cap = max_batch;
        while (true) {
            std::vector<std::shared_ptr<Request>> take;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
                if (pending.empty()) return;  // stop, and the queue is drained
                // the oldest request and the next ones of its value type, up
                // to `cap` requests and kBatchNnz nonzeros of A
                const spg_dtype dt = pending.front()->dtype;
                long long nz = 0;
                for (auto it = pending.begin(); it != pending.end() && static_cast<int>(take.size()) < cap;) {
                    const long long w = (*it)->A.nnz;
                    if ((*it)->dtype == dt && (take.empty() || nz + w <= kBatchNnz)) {
                        nz += w;
                        take.push_back(*it);
                        it = pending.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            Flight f;
            f.lane = lane;
            f.reqs = take;
            const int n = static_cast<int>(take.size());
            for (auto& r : take) {
                f.A.push_back(r->A);
                f.B.push_back(r->B);
            }
            f.C.assign(n, spg_csr{});
            BinTable sym, num;
            {
                std::lock_guard<std::mutex> g(c->mu);
                sym = c->sym;
                num = c->num;
            }
            const CallMode mode{L.own, true, false, false};
            if (take.front()->dtype == SPG_F64)
                f.job = std::make_unique<Job<double>>(*c, L, sym, num, false, mode);
            else
                f.job = std::make_unique<Job<float>>(*c, L, sym, num, false, mode);
            spg_status st = f.job->symbolic(n, f.A.data(), f.B.data(), f.C.data());
            if (st != SPG_OK) {
                publish(f, st, g_last_error);
                continue;
            }
            f.bytes = f.job->output_bytes();
            bool admit = false;
            {
                // admission by exact output memory
                std::unique_lock<std::mutex> lk(mu);
                cv_mem.wait(lk, [&] { return in_flight + f.bytes <= budget || admitted == 0; });
                if (in_flight + f.bytes <= budget) {
                    admit = true;
                    in_flight += f.bytes;
                    peak_bytes = std::max(peak_bytes, in_flight);
                    ++admitted;
                    ++n_batches;
                    max_seen = std::max<int64_t>(max_seen, n);
                } else if (n > 1) {  // over budget even alone: split, back to the front of the queue
                    for (int i = n - 1; i >= 0; --i) pending.push_front(take[i]);
                    n_deferred += n;
                }
            }
            if (!admit) {
                f.job->abort();
                if (n > 1) {
                    cap = std::max(1, n / 2);
                    cv_work.notify_all();
                } else {
                    f.C.assign(1, spg_csr{});
                    publish(f, SPG_ERR_OUT_OF_MEMORY, "the request's output exceeds the server's memory budget");
                }
                continue;
            }
            cap = max_batch;
            st = f.job->launch_numeric();
            if (st == SPG_OK) st = f.job->finish();
            publish(f, st, st == SPG_OK ? std::string() : g_last_error);
            {
                std::lock_guard<std::mutex> lk(mu);
                in_flight -= f.bytes;
                --admitted;
            }
            cv_mem.notify_all();
        }
    }
};

}  // namespace
}  // namespace spg

// ================================================================== C ABI

using namespace spg;

extern "C" {

spg_status spg_init(void) {
    Context* c = nullptr;
    return ensure_context(c);
}

spg_status spg_multiply(const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype) {
    return multiply_entry(1, A, B, C, dtype, nullptr, false);
}

spg_status spg_multiply_batched(int32_t count, const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype) {
    return multiply_entry(count, A, B, C, dtype, nullptr, false);
}

spg_status spg_multiply_async(const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype, void* stream) {
    return multiply_entry(1, A, B, C, dtype, static_cast<hipStream_t>(stream), true);
}

spg_status spg_multiply_noalloc(const spg_csr* A, const spg_csr* B, spg_csr* C, int64_t capacity, spg_dtype dtype) {
    return spg::multiply_noalloc(A, B, C, capacity, dtype);
}

static spg_status csr_free_after(spg_csr* C, hipStream_t after) {
    if (!C) return fail(SPG_ERR_INVALID_ARGUMENT, "null matrix");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    void* blocks[3] = {C->rpt, C->col, C->val};
    SPG_TRY(c->cache.release_all(blocks, 3, after));
    *C = spg_csr{0, 0, 0, nullptr, nullptr, nullptr};
    return SPG_OK;
}

spg_status spg_csr_free(spg_csr* C) { return csr_free_after(C, hipStreamPerThread); }

spg_status spg_csr_free_async(spg_csr* C, void* stream) {
    return csr_free_after(C, stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread);
}

spg_status spg_get_nip(const spg_csr* A, const spg_csr* B, int64_t* nip) {
    if (!A || !B || !nip) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    SPG_TRY(check_operands(*A, *B));
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    const hipStream_t s = hipStreamPerThread;
    void* d = nullptr;
    SPG_HIP(hipMalloc(&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), s);
    if (e == hipSuccess) {
        launch_total_nip(A->rpt, A->col, B->rpt, A->m, static_cast<unsigned long long*>(d), s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "spg_get_nip", __FILE__, __LINE__);
    *nip = static_cast<int64_t>(h);
    return SPG_OK;
}

spg_status spg_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    *ptr = nullptr;
    if (bytes == 0) return SPG_OK;
    std::lock_guard<std::recursive_mutex> lock(g_mutex);  // never during a graph capture
    SPG_HIP(hipMalloc(ptr, bytes));
    return SPG_OK;
}

spg_status spg_free(void* ptr) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    if (ptr) SPG_HIP(hipFree(ptr));
    return SPG_OK;
}

spg_status spg_memcpy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPG_OK;
    if (!dst || !src) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    SPG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return SPG_OK;
}

spg_status spg_device_sync(void) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    SPG_HIP(hipDeviceSynchronize());
    return SPG_OK;
}

const char* spg_status_str(spg_status status) {
    switch (status) {
        case SPG_OK: return "success";
        case SPG_ERR_INVALID_ARGUMENT: return "invalid argument";
        case SPG_ERR_OUT_OF_MEMORY: return "out of device memory";
        case SPG_ERR_HIP: return "HIP runtime error";
        case SPG_ERR_OVERFLOW: return "index overflow";
        case SPG_ERR_UNSUPPORTED_DEVICE: return "unsupported device";
        case SPG_ERR_INTERNAL: return "internal consistency check failed";
    }
    return "unknown status";
}

const char* spg_last_error(void) { return g_last_error.c_str(); }

spg_status spg_set_log_level(int32_t level) {
    if (level < kTrace || level > kOff) return fail(SPG_ERR_INVALID_ARGUMENT, "log level must be 0 (trace) .. 6 (off)");
    g_log_level.store(level);
    return SPG_OK;
}

spg_status spg_device_info_get(spg_device_info* info) {
    if (!info) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->arch, sizeof(info->arch), "%s", c->arch.c_str());
    info->compute_units = c->cus;
    info->lds_per_block = c->lds;
    info->clock_khz = c->clock_khz;
    info->global_mem_bytes = static_cast<int64_t>(c->global_mem);
    return SPG_OK;
}

static spg_status parse_table(int32_t phase, const char* text, BinTable& t) {
    t = BinTable{};
    std::istringstream in(text ? text : "");
    std::string line;
    while (std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos || line[0] == '#') continue;
        std::istringstream ls(line);
        long long v, nip, nnz;
        if (!(ls >> v >> nip >> nnz)) return fail(SPG_ERR_INVALID_ARGUMENT, "bad tuning line: " + line);
        if (t.n == kMaxBins) return fail(SPG_ERR_INVALID_ARGUMENT, "too many bins");
        if (nip < 1 || nnz < 1 || nip > INT_MAX || nnz > INT_MAX)
            return fail(SPG_ERR_INVALID_ARGUMENT, "caps out of range: " + line);
        const int b = t.n++;
        t.variant[b] = static_cast<int>(v);
        t.nip_cap[b] = static_cast<int>(nip);
        t.nnz_cap[b] = static_cast<int>(nnz);
        if (phase == 0)
            t.rows[b] = (v >= 0 && v <= kSymMaxVariant) ? kSymVariants[v].r : 1;
        else
            t.rows[b] = (v >= 0 && v < kNumGlobalVariant) ? kNumVariants[v].r : 1;
    }
    return phase == 0 ? check_sym_table(t) : check_num_table(t);
}

spg_status spg_set_tuning(int32_t phase, const char* table) {
    if (phase != 0 && phase != 1) return fail(SPG_ERR_INVALID_ARGUMENT, "phase must be 0 or 1");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    BinTable t;
    if (!table || !*table) {  // reset to the defaults
        BinTable sym, num;
        default_tables(sym, num);
        t = phase == 0 ? sym : num;
    } else {
        SPG_TRY(parse_table(phase, table, t));
    }
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    std::lock_guard<std::mutex> g(c->mu);
    (phase == 0 ? c->sym : c->num) = t;
    log_msg(kDebug, "%s bins set: %s", phase == 0 ? "symbolic" : "fused", table_text(t).c_str());
    return SPG_OK;
}

static spg_status write_text(const std::string& s, char* buf, size_t len) {
    if (!buf || len == 0) return fail(SPG_ERR_INVALID_ARGUMENT, "null buffer");
    if (s.size() + 1 > len) return fail(SPG_ERR_INVALID_ARGUMENT, "buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SPG_OK;
}

spg_status spg_get_tuning(int32_t phase, char* buf, size_t buf_len) {
    if (phase != 0 && phase != 1) return fail(SPG_ERR_INVALID_ARGUMENT, "phase must be 0 or 1");
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    BinTable t;
    {
        std::lock_guard<std::mutex> g(c->mu);
        t = phase == 0 ? c->sym : c->num;
    }
    std::ostringstream out;
    for (int b = 0; b < t.n; ++b) out << t.variant[b] << ' ' << t.nip_cap[b] << ' ' << t.nnz_cap[b] << '\n';
    return write_text(out.str(), buf, buf_len);
}

spg_status spg_get_variants(int32_t phase, char* buf, size_t buf_len) {
    std::ostringstream out;
    if (phase == 0) {
        for (int v = 0; v <= kSymMaxVariant; ++v)
            out << v << " wg=" << kSymVariants[v].wg << " rows=" << kSymVariants[v].r << " table=" << kSymVariants[v].t
                << " chunk=" << kSymVariants[v].ec << " slots=" << kSymVariants[v].pmax
                << (v == kSymMaxVariant ? " max" : "") << '\n';
    } else if (phase == 1) {
        for (int v = 0; v < kNumGlobalVariant; ++v)
            out << v << " wg=" << kNumVariants[v].wg << " rows=" << kNumVariants[v].r << " capacity="
                << kNumVariants[v].p << " chunk=" << kNumVariants[v].ec << " slots=" << kNumVariants[v].pmax
                << (num_multi_set(v) >= 0 ? " merged=" + std::to_string(num_multi_set(v)) : std::string()) << '\n';
        out << kNumGlobalVariant << " global\n";
    } else {
        return fail(SPG_ERR_INVALID_ARGUMENT, "phase must be 0 or 1");
    }
    return write_text(out.str(), buf, buf_len);
}

spg_status spg_set_graphs(int32_t enabled) {
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    c->graphs = enabled != 0;
    return SPG_OK;
}

spg_status spg_graph_stats(int64_t* captured, int64_t* replayed) {
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    if (captured) *captured = c->n_captured.load();
    if (replayed) *replayed = c->n_replayed.load();
    return SPG_OK;
}

spg_status spg_set_profiling(int32_t enabled) {
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    c->profiling = enabled != 0;
    prof_reset(*c);
    return SPG_OK;
}

int32_t spg_profile_count(void) {
    Context* c = nullptr;
    if (ensure_context(c) != SPG_OK) return 0;
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    return static_cast<int32_t>(c->prof.size());
}

spg_status spg_profile_get(int32_t index, char* name, size_t name_len, int64_t* rows, int32_t* grid, double* ms) {
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    if (index < 0 || index >= static_cast<int32_t>(c->prof.size()))
        return fail(SPG_ERR_INVALID_ARGUMENT, "profile index out of range");
    const ProfRec& r = c->prof[index];
    if (name) SPG_TRY(write_text(r.name, name, name_len));
    if (rows) *rows = r.rows;
    if (grid) *grid = r.grid;
    if (ms) *ms = r.ms;
    return SPG_OK;
}

spg_status spg_server_create(size_t memory_budget_bytes, int32_t max_batch, spg_server** out) {
    if (!out || max_batch < 1 || memory_budget_bytes == 0) return fail(SPG_ERR_INVALID_ARGUMENT, "bad server arguments");
    *out = nullptr;
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    auto* srv = new Server();
    srv->c = c;
    srv->budget = memory_budget_bytes;
    srv->max_batch = max_batch;
    for (auto& L : srv->lanes) {
        const spg_status st = lane_create(L);
        if (st != SPG_OK) {
            for (auto& M : srv->lanes) lane_destroy(M);
            delete srv;
            return st;
        }
    }
    for (int i = 0; i < Server::kLanes; ++i) srv->workers[i] = std::thread([srv, i] { srv->run(i); });
    *out = reinterpret_cast<spg_server*>(srv);
    return SPG_OK;
}
spg_status spg_server_submit(spg_server* server, const spg_csr* A, const spg_csr* B, spg_dtype dtype,
                             int64_t* ticket) {
    if (!server || !A || !B || !ticket) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    if (dtype != SPG_F32 && dtype != SPG_F64) return fail(SPG_ERR_INVALID_ARGUMENT, "unknown dtype");
    SPG_TRY(check_operands(*A, *B));
    auto* srv = reinterpret_cast<Server*>(server);
    auto r = std::make_shared<Request>();
    r->A = *A;
    r->B = *B;
    r->dtype = dtype;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return fail(SPG_ERR_INVALID_ARGUMENT, "server is shutting down");
        r->ticket = srv->next_ticket++;
        srv->requests[r->ticket] = r;
        srv->pending.push_back(r);
    }
    srv->cv_work.notify_one();
    *ticket = r->ticket;
    return SPG_OK;
}

spg_status spg_server_wait(spg_server* server, int64_t ticket, spg_csr* C, double* latency_ms) {
    if (!server || !C) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    auto* srv = reinterpret_cast<Server*>(server);
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->requests.find(ticket);
        if (it == srv->requests.end()) return fail(SPG_ERR_INVALID_ARGUMENT, "unknown ticket");
        r = it->second;
        srv->cv_done.wait(lk, [&] { return r->done; });
        srv->requests.erase(ticket);  // by key: submits during the wait may have rehashed the map
    }
    *C = r->C;
    if (latency_ms) *latency_ms = 1e3 * (r->t_done - r->t_submit);
    if (r->st != SPG_OK) return fail(r->st, r->err);
    return SPG_OK;
}

spg_status spg_server_stats(spg_server* server, int64_t* batches, int64_t* served, int64_t* max_batch_seen,
                            int64_t* deferred, int64_t* peak_bytes) {
    if (!server) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    auto* srv = reinterpret_cast<Server*>(server);
    std::lock_guard<std::mutex> lk(srv->mu);
    if (batches) *batches = srv->n_batches;
    if (served) *served = srv->n_served;
    if (max_batch_seen) *max_batch_seen = srv->max_seen;
    if (deferred) *deferred = srv->n_deferred;
    if (peak_bytes) *peak_bytes = static_cast<int64_t>(srv->peak_bytes);
    return SPG_OK;
}

spg_status spg_server_destroy(spg_server* server) {
    if (!server) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    auto* srv = reinterpret_cast<Server*>(server);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& w : srv->workers)
        if (w.joinable()) w.join();
    for (auto& kv : srv->requests) {  // results never claimed
        spg_csr& C = kv.second->C;
        if (kv.second->st == SPG_OK) (void)spg_csr_free(&C);
    }
    for (auto& L : srv->lanes) lane_destroy(L);
    delete srv;
    return SPG_OK;
}

}  // extern "C"
