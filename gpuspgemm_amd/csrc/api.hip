// Host driver and C ABI of the gfx950 SpGEMM library.
//
// proposal_inner (src/libproposal/include/proposal/proposal.cuh:26-80)
// creates eight streams, allocates its scratch and synchronises with the
// host at least ten times per multiply.  Here a per-device context keeps a
// lane (four streams, a scratch arena that only grows, pinned host memory
// for the read-backs) and the bin tables; a multiply is
//
//   nip (with the symbolic binning), sym_multi_small || sym_multi ||
//   sym_max (three streams), scan_count, scan_apply
//   -> one kernel copies the counters to pinned host memory and signals;
//      the host polls the signal (one synchronisation)
//   -> allocate C.col / C.val of the exact size
//   -> one symnum launch per non-empty fused bin over the four streams,
//      then a kernel that copies the flags back, zeroes the counters and
//      signals the host
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
#include <memory>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

namespace spg {
namespace {

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

// The device resources of one multiply in flight: a main stream plus aux
// streams that the fused-pass launches are spread over, events, a scratch
// arena that only grows, and pinned host memory for the read-backs.  The
// context owns one lane for the plain calls; a server owns lanes of its own
// so that the symbolic phase of one batch overlaps the fused pass of the
// batch before it.
struct Lane {
    static constexpr int kStreams = 4;
    hipStream_t s = nullptr;
    hipStream_t aux[kStreams - 1] = {};
    hipEvent_t ev_ready = nullptr;
    hipEvent_t ev_done[kStreams - 1] = {};
    hipEvent_t ev_host = nullptr;  // error check while the host polls
    int* h_flag = nullptr;          // pinned: cleared by the host, set to 1 by k_signal
    char* scratch = nullptr;
    size_t scratch_cap = 0;
    bool meta_clean = false;  // the counters at the start of scratch are zero
    char* gpool = nullptr;  // global-bin tables of the fused pass under graph replay
    size_t gpool_cap = 0;
    DevMeta* h_meta = nullptr;
    char* h_stage = nullptr;
    size_t h_stage_cap = 0;
    GraphCache graphs[2];  // symbolic phase, fused pass
};

struct Context {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;  // the context lane's main stream
    int cus = 0;
    int lds = 0;
    int clock_khz = 0;
    size_t global_mem = 0;
    std::string arch;
    bool async_alloc = true;
    BinTable sym{};
    BinTable num{};
    int occ_sym_multi[2] = {1, 1}, occ_sym_max = 1;
    int occ_num[2][kNumLdsVariants] = {};
    Lane lane;
    bool graphs = false;  // replay repeated phases from captured graphs
    std::atomic<int64_t> n_captured{0}, n_replayed{0};
    bool profiling = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> events;
    size_t events_used = 0;
};

std::recursive_mutex g_mutex;
constexpr int kMaxDevices = 64;
Context g_ctx[kMaxDevices];

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

// ---------------------------------------------------------------- context

spg_status lane_create(Lane& L) {
    SPG_HIP(hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking));
    for (int i = 0; i < Lane::kStreams - 1; ++i) {
        SPG_HIP(hipStreamCreateWithFlags(&L.aux[i], hipStreamNonBlocking));
        SPG_HIP(hipEventCreateWithFlags(&L.ev_done[i], hipEventDisableTiming));
    }
    SPG_HIP(hipEventCreateWithFlags(&L.ev_ready, hipEventDisableTiming));
    SPG_HIP(hipEventCreateWithFlags(&L.ev_host, hipEventDisableTiming));
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_flag), sizeof(int), hipHostMallocDefault));
    *L.h_flag = 0;
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&L.h_meta), sizeof(DevMeta), hipHostMallocDefault));
    return SPG_OK;
}

void lane_destroy(Lane& L) {
    if (L.s) (void)hipStreamSynchronize(L.s);
    for (int i = 0; i < Lane::kStreams - 1; ++i) {
        if (L.aux[i]) (void)hipStreamSynchronize(L.aux[i]);
        if (L.aux[i]) (void)hipStreamDestroy(L.aux[i]);
        if (L.ev_done[i]) (void)hipEventDestroy(L.ev_done[i]);
    }
    if (L.ev_ready) (void)hipEventDestroy(L.ev_ready);
    if (L.ev_host) (void)hipEventDestroy(L.ev_host);
    L.graphs[0].clear();
    L.graphs[1].clear();
    if (L.h_flag) (void)hipHostFree(L.h_flag);
    if (L.scratch) (void)hipFree(L.scratch);
    if (L.gpool) (void)hipFree(L.gpool);
    if (L.h_meta) (void)hipHostFree(L.h_meta);
    if (L.h_stage) (void)hipHostFree(L.h_stage);
    if (L.s) (void)hipStreamDestroy(L.s);
    L = Lane{};
}

// Completion of the work queued on a lane's main stream.  The host clears a
// word of pinned host memory (lane_arm), a kernel queued last stores 1 into
// it (launch_signal, eagerly or as a graph's last node) and the host polls
// the word (lane_poll): the runtime's own completion of an event or stream
// was measured at ~55 us after the GPU finished, at each of the multiply's
// two read-backs.  The event is queried now and then so that a failed stream
// ends the wait with its error.  Every signal is polled for before the lane
// is armed again, so no earlier store can land after the clear.
void lane_arm(Lane& L) { __atomic_store_n(L.h_flag, 0, __ATOMIC_RELEASE); }

hipError_t lane_poll(Lane& L) {
    hipError_t e = hipEventRecord(L.ev_host, L.s);
    if (e != hipSuccess) return e;
    for (unsigned spin = 1;; ++spin) {
        if (__atomic_load_n(L.h_flag, __ATOMIC_ACQUIRE) == 1) return hipSuccess;
        if ((spin & 1023) == 0) {
            e = hipEventQuery(L.ev_host);
            if (e == hipSuccess) return __atomic_load_n(L.h_flag, __ATOMIC_ACQUIRE) == 1 ? hipSuccess : hipErrorUnknown;
            if (e != hipErrorNotReady) return e;
        }
    }
}

hipError_t lane_wait(Lane& L) {
    lane_arm(L);
    launch_signal(L.h_flag, L.s);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? lane_poll(L) : e;
}

spg_status ensure_context(Context*& out) {
    int dev = 0;
    SPG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SPG_ERR_UNSUPPORTED_DEVICE, "device index out of range");
    Context& c = g_ctx[dev];
    out = &c;
    if (c.ready) return SPG_OK;
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
    // Keep freed blocks in the stream-ordered pool: C is allocated per call.
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t threshold = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
    } else {
        c.async_alloc = false;
    }
    SPG_TRY(lane_create(c.lane));
    c.stream = c.lane.s;
    default_tables(c.sym, c.num);
    SPG_TRY(check_sym_table(c.sym));  // a stale tuned_tables.inc fails here, loudly
    SPG_TRY(check_num_table(c.num));
    c.occ_sym_multi[0] = occupancy_sym_multi<double>(0);
    c.occ_sym_multi[1] = occupancy_sym_multi<double>(1);
    c.occ_sym_max = occupancy_sym_max<double>();
    for (int v = 0; v < kNumLdsVariants; ++v) {
        c.occ_num[0][v] = occupancy_symnum_bin<float>(v);
        c.occ_num[1][v] = occupancy_symnum_bin<double>(v);
    }
    c.ready = true;
    return SPG_OK;
}

spg_status dev_alloc(Context& c, void** p, size_t bytes, hipStream_t s) {
    *p = nullptr;
    if (bytes == 0) return SPG_OK;
    if (c.async_alloc) {
        const hipError_t e = hipMallocAsync(p, bytes, s);
        if (e == hipSuccess) return SPG_OK;
        if (e != hipErrorNotSupported) return hip_fail(e, "hipMallocAsync", __FILE__, __LINE__);
        c.async_alloc = false;
    }
    SPG_HIP(hipMalloc(p, bytes));
    return SPG_OK;
}

spg_status dev_release(Context& c, void* p, hipStream_t s) {
    if (!p) return SPG_OK;
    if (c.async_alloc) SPG_HIP(hipFreeAsync(p, s));
    else SPG_HIP(hipFree(p));
    return SPG_OK;
}

spg_status ensure_scratch(Lane& L, size_t bytes) {
    if (bytes <= L.scratch_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(L.s));
    if (L.scratch) SPG_HIP(hipFree(L.scratch));
    L.scratch = nullptr;
    const size_t cap = std::max(bytes, L.scratch_cap * 3 / 2);
    L.scratch_cap = 0;
    L.meta_clean = false;
    SPG_HIP(hipMalloc(reinterpret_cast<void**>(&L.scratch), cap));
    L.scratch_cap = cap;
    return SPG_OK;
}

spg_status ensure_gpool(Lane& L, size_t bytes) {
    if (bytes <= L.gpool_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(L.s));
    if (L.gpool) SPG_HIP(hipFree(L.gpool));
    L.gpool = nullptr;
    const size_t cap = std::max(bytes, L.gpool_cap * 3 / 2);
    L.gpool_cap = 0;
    SPG_HIP(hipMalloc(reinterpret_cast<void**>(&L.gpool), cap));
    L.gpool_cap = cap;
    return SPG_OK;
}

spg_status ensure_stage(Lane& L, size_t bytes) {
    if (bytes <= L.h_stage_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(L.s));
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

// One batch of products on one lane, in two phases:
//   symbolic()        NIP, binning, exact counts, scan (row pointers) and
//                     the fused-pass bins; ends with one read-back, after
//                     which nnz(C_i) of every product is known on the host
//   launch_numeric()  allocates C.col / C.val of the exact sizes and queues
//                     the fused pass on the lane's streams (no wait)
//   finish()          waits for the lane and checks the consistency flag.
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
    int count = 0;
    spg_csr* C = nullptr;
    BinTable sym{}, num{};
    std::vector<void*> owned;
    std::vector<Prob<V>> probs;
    std::vector<Tile> tiles;
    ProbSet<V> ps{};
    long long total_rows = 0;
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
        int32_t num_count[kMaxBins] = {};
        unsigned long long gnum_pool = 0;
        long long nnz_total = 0;
    } hm;
    std::vector<long long> nnz;
    void* num_pool = nullptr;   // allocated per call (eager) or the lane's pool (graph replay)
    bool graphs = false;        // phases may be replayed from the lane's graph caches
    bool signal_queued = false; // the fused pass already queued its completion signal

    // Only the context's own lane captures graphs: its calls run under the
    // library mutex, which the library's synchronous helpers (spg_memcpy,
    // spg_malloc, spg_free, spg_device_sync) take too.  HIP rejects a
    // synchronous call on the legacy stream from any thread while some
    // stream of the device is capturing, so a server's worker threads,
    // which run beside other threads, never capture.
    Job(Context& ctx, Lane& lane, const BinTable& sym_table, const BinTable& num_table, bool profile)
        : c(ctx), L(lane), prof(ctx, profile), sym(sym_table), num(num_table),
          graphs(&lane == &ctx.lane && ctx.graphs && !profile) {}

    void abort() override {
        for (void* p : owned) (void)dev_release(c, p, L.s);
        owned.clear();
        if (num_pool && !graphs) (void)dev_release(c, num_pool, L.s);
        num_pool = nullptr;
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

    spg_status front() {
        const hipStream_t s = L.s;
        if (batch) {
            SPG_HIP(hipMemcpyAsync(d_probs, L.h_stage + stage_probs, sizeof(Prob<V>) * count, hipMemcpyHostToDevice,
                                   s));
            SPG_HIP(hipMemcpyAsync(d_tiles, L.h_stage + stage_tiles, sizeof(Tile) * n_tiles, hipMemcpyHostToDevice,
                                   s));
        }
        // The counters are cleared by the last kernel of a multiply that
        // completed; after anything else (first use, grown scratch, an error
        // or an abandoned job) they are cleared here.
        if (graphs || !L.meta_clean) {
            prof.begin("clear", 0, 1, s);
            launch_clear(meta, 0, s);
            prof.end();
        }
        L.meta_clean = false;
        prof.begin("nip", total_rows, n_tiles, s);
        launch_nip<V>(ps, sym, meta, items1, static_cast<int>(total_rows), s);
        prof.end();
        // The largest-table bin runs on an aux stream beside the multi-bin
        // launch (it is queued first: its workgroups without a row exit at
        // once and leave the CUs to the multi-bin grid).
        SPG_HIP(hipEventRecord(L.ev_ready, s));
        SPG_HIP(hipStreamWaitEvent(L.aux[0], L.ev_ready, 0));
        {
            const int grid = c.cus * c.occ_sym_max;
            prof.begin("sym_max", 0, grid, L.aux[0]);
            launch_sym_max<V>(ps, items1, static_cast<int>(total_rows), sym, sym.n - 1, meta, fail_items, fail_off,
                              grid, L.aux[0]);
            prof.end();
        }
        const int n_multi = sym.n - 1;
        bool part_used[2] = {false, false};
        for (int b = 0; b < n_multi; ++b) part_used[sym.variant[b] < kSymSmallVariants ? 0 : 1] = true;
        // With both parts in use the larger-table part runs on a second aux
        // stream: its workgroups take the CUs that the small-table grid
        // leaves as its last units drain.  The parts count disjoint rows.
        const bool split = part_used[0] && part_used[1];
        if (split) SPG_HIP(hipStreamWaitEvent(L.aux[1], L.ev_ready, 0));
        for (int part = 0; part < 2; ++part) {
            if (!part_used[part]) continue;
            const hipStream_t where = split && part == 1 ? L.aux[1] : s;
            int grid = c.cus * c.occ_sym_multi[part];
            grid = (grid + kXcds - 1) / kXcds * kXcds;
            prof.begin(part == 0 ? "sym_multi_small" : "sym_multi", total_rows, grid, where);
            launch_sym_multi<V>(part, ps, items1, static_cast<int>(total_rows), sym, n_multi, meta, grid, where);
            prof.end();
        }
        SPG_HIP(hipEventRecord(L.ev_done[0], L.aux[0]));
        SPG_HIP(hipStreamWaitEvent(s, L.ev_done[0], 0));
        if (split) {
            SPG_HIP(hipEventRecord(L.ev_done[1], L.aux[1]));
            SPG_HIP(hipStreamWaitEvent(s, L.ev_done[1], 0));
        }
        return SPG_OK;
    }

    // scan stage and the read-back of the counters, ending with the
    // completion signal (the caller armed the lane and polls it)
    spg_status scan(int after_fail) {
        const hipStream_t s = L.s;
        prof.begin("scan_count", total_rows, n_tiles, s);
        launch_scan_count<V>(ps, num, meta, tile_sums, after_fail, s);
        prof.end();
        prof.begin("scan_apply", total_rows, n_tiles, s);
        launch_scan_apply<V>(ps, num, meta, tile_sums, items2, gpool_off, inst_nnz, after_fail, s);
        prof.end();
        // counters (and nnz of every product of a batch) to pinned host memory, then the signal
        launch_readback(reinterpret_cast<const int*>(meta), reinterpret_cast<int*>(L.h_meta),
                        static_cast<int>(sizeof(DevMeta) / sizeof(int)), inst_nnz,
                        reinterpret_cast<long long*>(L.h_stage + stage_inst), batch ? count : 0, nullptr, L.h_flag, s);
        SPG_HIP(hipGetLastError());
        return SPG_OK;
    }

    // Signature of a phase: every value its launches depend on.  Buffer
    // contents (operand descriptors, tiles) are copied from pinned host
    // memory by the phase itself and so are not part of it.
    std::vector<int64_t> signature(int phase) const {
        std::vector<int64_t> k{phase, static_cast<int64_t>(sizeof(V)), count, n_tiles, total_rows,
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
        lane_arm(L);
        if (!graphs) return record();
        std::vector<int64_t> sig = signature(phase);
        GraphCache& gc = L.graphs[phase];
        hipGraphExec_t exec = gc.find(sig);
        if (exec) {
            ++c.n_replayed;
        } else {
            if (!gc.seen_before(sig)) return record();
            SPG_HIP(hipStreamBeginCapture(L.s, hipStreamCaptureModeRelaxed));
            spg_status st = record();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(L.s, &g);
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
        SPG_HIP(hipGraphLaunch(exec, L.s));
        return SPG_OK;
    }

    spg_status symbolic(int n, const spg_csr* A, const spg_csr* B, spg_csr* out) override {
        count = n;
        C = out;
        batch = count > 1;
        const hipStream_t s = L.s;
        for (int i = 0; i < count; ++i) {
            SPG_TRY(check_operands(A[i], B[i]));
            C[i] = spg_csr{0, A[i].m, B[i].n, nullptr, nullptr, nullptr};
        }
        // output row pointers
        for (int i = 0; i < count; ++i) {
            void* p = nullptr;
            const spg_status st = dev_alloc(c, &p, sizeof(int32_t) * (static_cast<size_t>(A[i].m) + 1), s);
            if (st != SPG_OK) return failed(st);
            owned.push_back(p);
            C[i].rpt = static_cast<int32_t*>(p);
            total_rows += A[i].m;
            n_tiles += (A[i].m + kTileRows - 1) / kTileRows;
            if (A[i].m == 0) {
                const hipError_t e = hipMemsetAsync(C[i].rpt, 0, sizeof(int32_t), s);
                if (e != hipSuccess) return failed(hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__));
            }
        }
        nnz.assign(count, 0);
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
        const size_t o_it1 = take(sizeof(int2) * M * sym.n);  // a region of M items per symbolic bin
        const size_t o_it2 = take(sizeof(RowItem) * M);
        const size_t o_fit = take(sizeof(int2) * M);
        const size_t o_foff = take(sizeof(long long) * M);
        const size_t o_goff = take(sizeof(long long) * M);
        const size_t o_probs = take(batch ? sizeof(Prob<V>) * count : 0);
        const size_t o_tiles = take(batch ? sizeof(Tile) * n_tiles : 0);
        {
            const spg_status st = ensure_scratch(L, off);
            if (st != SPG_OK) return failed(st);
        }
        char* base = L.scratch;
        meta = reinterpret_cast<DevMeta*>(base + o_meta);
        tile_sums = reinterpret_cast<long long*>(base + o_tsum);
        inst_nnz = reinterpret_cast<long long*>(base + o_inst);
        int32_t* nip = reinterpret_cast<int32_t*>(base + o_nip);
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
            p.m = A[i].m;
            p.n = B[i].n;
            if (batch) {
                const int first = static_cast<int>(tiles.size());
                for (int r0 = 0; r0 < A[i].m; r0 += kTileRows)
                    tiles.push_back(Tile{i, r0, std::min(kTileRows, A[i].m - r0), first});
            }
            row_base += A[i].m;
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
            const spg_status st = ensure_stage(L, stage_inst + align256(sizeof(long long) * count));
            if (st != SPG_OK) return failed(st);
            std::memcpy(L.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);
            std::memcpy(L.h_stage + stage_tiles, tiles.data(), sizeof(Tile) * n_tiles);
        }

        spg_status st = queue_phase(0, [&]() -> spg_status {
            SPG_TRY(front());
            return scan(0);
        });
        if (st == SPG_OK) {
            const hipError_t e = lane_poll(L);
            if (e != hipSuccess) st = hip_fail(e, "symbolic phase", __FILE__, __LINE__);
        }
        if (st == SPG_OK && L.h_meta->fail_count.v > 0) {
            // rows that overflowed the largest LDS table: count them again in
            // global memory, then redo the scan stage
            const int fails = L.h_meta->fail_count.v;
            void* sym_pool = nullptr;
            st = dev_alloc(c, &sym_pool, sizeof(int32_t) * L.h_meta->fail_pool.v, s);
            if (st == SPG_OK) {
                prof.begin("sym_global", fails, fails, s);
                launch_sym_global<V>(ps, fail_items, fail_off, static_cast<int*>(sym_pool), fails, s);
                prof.end();
                launch_clear(meta, kMetaScanWord, s);
                lane_arm(L);
                st = scan(1);
                if (st == SPG_OK) {
                    const hipError_t e = lane_poll(L);
                    if (e != hipSuccess) st = hip_fail(e, "symbolic phase", __FILE__, __LINE__);
                }
            }
            if (sym_pool) (void)dev_release(c, sym_pool, s);
        }
        if (st != SPG_OK) return failed(st);
        hm.flags = L.h_meta->flags.v;
        for (int b = 0; b < kMaxBins; ++b) hm.num_count[b] = L.h_meta->num_count[b].v;
        hm.gnum_pool = L.h_meta->gnum_pool.v;
        hm.nnz_total = L.h_meta->nnz_total.v;
        if (hm.flags & kFlagNipOverflow)
            return failed(fail(SPG_ERR_OVERFLOW, "a row has 2^31 or more intermediate products"));
        if (hm.flags & kFlagNnzOverflow) return failed(fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32"));
        for (int i = 0; i < count; ++i) {
            if (A[i].m == 0) continue;
            nnz[i] = batch ? reinterpret_cast<long long*>(L.h_stage + stage_inst)[i] : hm.nnz_total;
            if (nnz[i] >= INT_MAX) return failed(fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32"));
        }
        return SPG_OK;
    }

    spg_status launch_numeric() override {
        if (empty) return SPG_OK;
        const hipStream_t s = L.s;
        // exact-size outputs
        for (int i = 0; i < count; ++i) {
            if (C[i].m == 0) continue;
            C[i].nnz = static_cast<int32_t>(nnz[i]);
            void *pc = nullptr, *pv = nullptr;
            spg_status st = dev_alloc(c, &pc, sizeof(int32_t) * nnz[i], s);
            if (st == SPG_OK) owned.push_back(pc);
            if (st == SPG_OK) st = dev_alloc(c, &pv, sizeof(V) * nnz[i], s);
            if (st == SPG_OK) owned.push_back(pv);
            if (st != SPG_OK) return failed(st);
            C[i].col = static_cast<int32_t*>(pc);
            C[i].val = pv;
            probs[i].c_col = C[i].col;
            probs[i].c_val = static_cast<V*>(pv);
        }
        ps.p0 = probs[0];
        if (batch) std::memcpy(L.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);

        // fused pass: one launch per non-empty bin, heaviest bins first,
        // spread over the lane's streams so that small bins overlap
        struct Launch {
            int variant, off, count;
        };
        std::vector<Launch> launches;
        int glob_bin = -1, glob_off = 0;
        {
            int item_off = 0;
            for (int b = 0; b < num.n; ++b) {
                const int cnt = hm.num_count[b];
                const int v = num.variant[b];
                if (cnt > 0) {
                    if (v == kNumGlobalVariant) {
                        glob_bin = b;
                        glob_off = item_off;
                    } else {
                        launches.push_back(Launch{v, item_off, cnt});
                    }
                }
                item_off += cnt;
            }
        }
        // Stream plan: launches in order of decreasing row capacity times
        // lanes per row, dealt round robin over the lane's streams.  The
        // heavy-row launches (few workgroups, each a long chain) are queued
        // first, so their workgroups reach the CUs before the launches of
        // many small rows fill them, and then run beside those.  (Measured
        // against a longest-first plan and a plan that keeps the
        // GPU-filling launches on one stream: this one was fastest.)
        std::stable_sort(launches.begin(), launches.end(), [](const Launch& x, const Launch& y) {
            return kNumVariants[x.variant].p * kNumVariants[x.variant].wg / kNumVariants[x.variant].r >
                   kNumVariants[y.variant].p * kNumVariants[y.variant].wg / kNumVariants[y.variant].r;
        });
        std::vector<int> stream_idx(launches.size());
        int n_streams = 1;
        {
            int k = glob_bin >= 0 ? 1 : 0;  // the global bin, when present, is queued first on the main stream
            for (size_t i = 0; i < launches.size(); ++i, ++k) stream_idx[i] = k % Lane::kStreams;
            n_streams = std::min(std::max(k, 1), static_cast<int>(Lane::kStreams));
        }
        auto stream_of = [&](int q) { return q == 0 ? s : L.aux[q - 1]; };
        if (glob_bin >= 0) {
            const size_t pool_bytes = sizeof(int32_t) * hm.gnum_pool;
            spg_status st = SPG_OK;
            if (graphs) {  // a buffer of the lane, so that its address recurs
                st = ensure_gpool(L, pool_bytes);
                num_pool = L.gpool;
            } else {
                st = dev_alloc(c, &num_pool, pool_bytes, s);
            }
            if (st != SPG_OK) return failed(st);
        }
        const spg_status st = queue_phase(1, [&]() -> spg_status {
            hipError_t e = hipSuccess;
            if (batch)
                e = hipMemcpyAsync(d_probs, L.h_stage + stage_probs, sizeof(Prob<V>) * count, hipMemcpyHostToDevice, s);
            // C.col, C.val and the pool were allocated on the main stream: an
            // aux stream waits for that point before its first launch.  The
            // waits are queued just before those launches, so that the host
            // time each one takes (several us) is not spent before the first
            // launch of the main stream, while the GPU idles.
            if (e == hipSuccess && n_streams > 1) e = hipEventRecord(L.ev_ready, s);
            bool aux_ready[Lane::kStreams] = {true, false, false, false};
            auto stream_for = [&](int q) -> hipStream_t {
                if (!aux_ready[q]) {
                    if (e == hipSuccess) e = hipStreamWaitEvent(L.aux[q - 1], L.ev_ready, 0);
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
            for (size_t li = 0; li < launches.size(); ++li) {
                if (e != hipSuccess) break;
                const Launch& Lh = launches[li];
                const NumVariant& nv = kNumVariants[Lh.variant];
                const int units = (Lh.count + nv.r - 1) / nv.r;
                const int grid = (units + kXcds - 1) / kXcds * kXcds;  // one unit per workgroup
                const hipStream_t on = stream_for(stream_idx[li]);
                if (e != hipSuccess) break;
                prof.begin("symnum_v" + std::to_string(Lh.variant), Lh.count, grid, on);
                e = launch_symnum_bin<V>(Lh.variant, ps, items2, Lh.off, Lh.count, meta, grid, on);
                prof.end();
            }
            for (int i = 1; i < n_streams && e == hipSuccess; ++i) {
                e = hipEventRecord(L.ev_done[i - 1], L.aux[i - 1]);
                if (e == hipSuccess) e = hipStreamWaitEvent(s, L.ev_done[i - 1], 0);
            }
            if (e == hipSuccess) e = hipGetLastError();
            if (e == hipSuccess) {
                // the flags word back, the counters zeroed for the next multiply, the signal
                launch_readback(&meta->flags.v, &L.h_meta->flags.v, 1, nullptr, nullptr, 0, meta, L.h_flag, s);
                e = hipGetLastError();
            }
            return e == hipSuccess ? SPG_OK : hip_fail(e, "fused pass", __FILE__, __LINE__);
        });
        if (st != SPG_OK) return failed(st);
        signal_queued = true;
        if (num_pool && !graphs) {
            (void)dev_release(c, num_pool, s);
            num_pool = nullptr;
        }
        return SPG_OK;
    }

    spg_status finish() override {
        const hipError_t e = signal_queued ? lane_poll(L) : lane_wait(L);
        if (e != hipSuccess) return failed(hip_fail(e, "fused pass", __FILE__, __LINE__));
        if (!empty && (L.h_meta->flags.v & kFlagMismatch))
            return failed(fail(SPG_ERR_INTERNAL, "fused pass found a different column count than the count pass "
                                                 "(are B's rows free of duplicate columns and indices in range?)"));
        if (signal_queued) L.meta_clean = true;  // the last read-back zeroed the counters
        owned.clear();  // the outputs now belong to the caller
        return SPG_OK;
    }
};

template <typename V>
spg_status multiply_impl(Context& c, int count, const spg_csr* A, const spg_csr* B, spg_csr* C) {
    prof_reset(c);
    Job<V> job(c, c.lane, c.sym, c.num, c.profiling);
    SPG_TRY(job.symbolic(count, A, B, C));
    SPG_TRY(job.launch_numeric());
    SPG_TRY(job.finish());
    if (c.profiling) prof_collect(c);
    return SPG_OK;
}

// ---------------------------------------------------------------- server

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

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
                std::lock_guard<std::recursive_mutex> g(g_mutex);
                sym = c->sym;
                num = c->num;
            }
            if (take.front()->dtype == SPG_F64)
                f.job = std::make_unique<Job<double>>(*c, L, sym, num, false);
            else
                f.job = std::make_unique<Job<float>>(*c, L, sym, num, false);
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
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    return ensure_context(c);
}

spg_status spg_multiply(const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype) {
    return spg_multiply_batched(1, A, B, C, dtype);
}

spg_status spg_multiply_batched(int32_t count, const spg_csr* A, const spg_csr* B, spg_csr* C, spg_dtype dtype) {
    if (count < 1 || !A || !B || !C) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    if (dtype != SPG_F32 && dtype != SPG_F64) return fail(SPG_ERR_INVALID_ARGUMENT, "unknown dtype");
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    g_last_error.clear();
    return dtype == SPG_F64 ? multiply_impl<double>(*c, count, A, B, C) : multiply_impl<float>(*c, count, A, B, C);
}

spg_status spg_csr_free(spg_csr* C) {
    if (!C) return fail(SPG_ERR_INVALID_ARGUMENT, "null matrix");
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    SPG_TRY(dev_release(*c, C->rpt, c->stream));
    SPG_TRY(dev_release(*c, C->col, c->stream));
    SPG_TRY(dev_release(*c, C->val, c->stream));
    *C = spg_csr{0, 0, 0, nullptr, nullptr, nullptr};
    return SPG_OK;
}

spg_status spg_get_nip(const spg_csr* A, const spg_csr* B, int64_t* nip) {
    if (!A || !B || !nip) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    SPG_TRY(check_operands(*A, *B));
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    void* d = nullptr;
    SPG_HIP(hipMalloc(&d, sizeof(unsigned long long)));
    unsigned long long h = 0;
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), c->stream);
    if (e == hipSuccess) {
        launch_total_nip(A->rpt, A->col, B->rpt, A->m, static_cast<unsigned long long*>(d), c->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
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

spg_status spg_device_info_get(spg_device_info* info) {
    if (!info) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
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
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    if (!table || !*table) {  // reset to the defaults
        BinTable sym, num;
        default_tables(sym, num);
        (phase == 0 ? c->sym : c->num) = phase == 0 ? sym : num;
        return SPG_OK;
    }
    BinTable t;
    SPG_TRY(parse_table(phase, table, t));
    (phase == 0 ? c->sym : c->num) = t;
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
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    const BinTable& t = phase == 0 ? c->sym : c->num;
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
                << kNumVariants[v].p << " chunk=" << kNumVariants[v].ec << " slots=" << kNumVariants[v].pmax << '\n';
        out << kNumGlobalVariant << " global\n";
    } else {
        return fail(SPG_ERR_INVALID_ARGUMENT, "phase must be 0 or 1");
    }
    return write_text(out.str(), buf, buf_len);
}

spg_status spg_set_graphs(int32_t enabled) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    c->graphs = enabled != 0;
    return SPG_OK;
}

spg_status spg_graph_stats(int64_t* captured, int64_t* replayed) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    if (captured) *captured = c->n_captured.load();
    if (replayed) *replayed = c->n_replayed.load();
    return SPG_OK;
}

spg_status spg_set_profiling(int32_t enabled) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
    c->profiling = enabled != 0;
    prof_reset(*c);
    return SPG_OK;
}

int32_t spg_profile_count(void) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    if (ensure_context(c) != SPG_OK) return 0;
    return static_cast<int32_t>(c->prof.size());
}

spg_status spg_profile_get(int32_t index, char* name, size_t name_len, int64_t* rows, int32_t* grid, double* ms) {
    std::lock_guard<std::recursive_mutex> lock(g_mutex);
    Context* c = nullptr;
    SPG_TRY(ensure_context(c));
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
    {
        std::lock_guard<std::recursive_mutex> lock(g_mutex);
        SPG_TRY(ensure_context(c));
    }
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
