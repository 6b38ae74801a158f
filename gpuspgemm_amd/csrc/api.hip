// Host driver and C ABI of the gfx950 SpGEMM library.
//
// proposal_inner (src/libproposal/include/proposal/proposal.cuh:26-80)
// creates eight streams, allocates its scratch and synchronises with the
// host at least ten times per multiply.  Here a per-device context keeps
// one stream, a scratch arena that only grows, pinned host memory for the
// one read-back, and the bin tables; a multiply is
//
//   clear, nip, scatter, sym_multi, sym_max, scan_count, scan_apply
//   -> one device-to-host copy of the counters + one synchronisation
//   -> allocate C.col / C.val of the exact size
//   -> symnum_multi (+ symnum_big / symnum_global when those bins are used)
//
// Rows that overflow the largest LDS table add a second round trip
// (global tables, then the scan again), as the reference's fail bin does
// (sym1.cuh:424-454).
#include "spg_kernels.hpp"
#include "../../include/gpuspgemm.h"

#include <algorithm>
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

struct Context {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    int cus = 0;
    int lds = 0;
    int clock_khz = 0;
    size_t global_mem = 0;
    std::string arch;
    bool async_alloc = true;
    BinTable sym{};
    BinTable num{};
    int occ_sym_multi = 1, occ_sym_max = 1;
    int occ_num[2][kNumLdsVariants] = {};
    static constexpr int kStreams = 4;  // fused-pass launches are spread over these
    hipStream_t aux[kStreams - 1] = {};
    hipEvent_t ev_ready = nullptr;
    hipEvent_t ev_done[kStreams - 1] = {};
    char* scratch = nullptr;
    size_t scratch_cap = 0;
    DevMeta* h_meta = nullptr;
    char* h_stage = nullptr;
    size_t h_stage_cap = 0;
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
    SPG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    // Keep freed blocks in the stream-ordered pool: C is allocated per call.
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t threshold = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
    } else {
        c.async_alloc = false;
    }
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.h_meta), sizeof(DevMeta), hipHostMallocDefault));
    default_tables(c.sym, c.num);
    SPG_TRY(check_sym_table(c.sym));  // a stale tuned_tables.inc fails here, loudly
    SPG_TRY(check_num_table(c.num));
    c.occ_sym_multi = occupancy_sym_multi<double>();
    c.occ_sym_max = occupancy_sym_max<double>();
    for (int v = 0; v < kNumLdsVariants; ++v) {
        c.occ_num[0][v] = occupancy_symnum_bin<float>(v);
        c.occ_num[1][v] = occupancy_symnum_bin<double>(v);
    }
    for (int i = 0; i < Context::kStreams - 1; ++i) {
        SPG_HIP(hipStreamCreateWithFlags(&c.aux[i], hipStreamNonBlocking));
        SPG_HIP(hipEventCreateWithFlags(&c.ev_done[i], hipEventDisableTiming));
    }
    SPG_HIP(hipEventCreateWithFlags(&c.ev_ready, hipEventDisableTiming));
    c.ready = true;
    return SPG_OK;
}

spg_status dev_alloc(Context& c, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return SPG_OK;
    if (c.async_alloc) {
        const hipError_t e = hipMallocAsync(p, bytes, c.stream);
        if (e == hipSuccess) return SPG_OK;
        if (e != hipErrorNotSupported) return hip_fail(e, "hipMallocAsync", __FILE__, __LINE__);
        c.async_alloc = false;
    }
    SPG_HIP(hipMalloc(p, bytes));
    return SPG_OK;
}

spg_status dev_release(Context& c, void* p) {
    if (!p) return SPG_OK;
    if (c.async_alloc) SPG_HIP(hipFreeAsync(p, c.stream));
    else SPG_HIP(hipFree(p));
    return SPG_OK;
}

spg_status ensure_scratch(Context& c, size_t bytes) {
    if (bytes <= c.scratch_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(c.stream));
    if (c.scratch) SPG_HIP(hipFree(c.scratch));
    c.scratch = nullptr;
    c.scratch_cap = 0;
    const size_t cap = std::max(bytes, c.scratch_cap * 3 / 2);
    SPG_HIP(hipMalloc(reinterpret_cast<void**>(&c.scratch), cap));
    c.scratch_cap = cap;
    return SPG_OK;
}

spg_status ensure_stage(Context& c, size_t bytes) {
    if (bytes <= c.h_stage_cap) return SPG_OK;
    SPG_HIP(hipStreamSynchronize(c.stream));
    if (c.h_stage) SPG_HIP(hipHostFree(c.h_stage));
    c.h_stage = nullptr;
    SPG_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.h_stage), bytes, hipHostMallocDefault));
    c.h_stage_cap = bytes;
    return SPG_OK;
}

// ---------------------------------------------------------------- profiling

struct Prof {
    Context& c;
    int open = -1;
    explicit Prof(Context& ctx) : c(ctx) {}
    hipEvent_t next_event() {
        if (c.events_used == c.events.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c.events.push_back(e);
        }
        return c.events[c.events_used++];
    }
    hipStream_t st = nullptr;
    void begin(const std::string& name, int64_t rows, int grid, hipStream_t on = nullptr) {
        if (!c.profiling) return;
        st = on ? on : c.stream;
        ProfRec r{name, rows, grid, next_event(), next_event(), 0.0};
        if (r.ev0) (void)hipEventRecord(r.ev0, st);
        c.prof.push_back(r);
        open = static_cast<int>(c.prof.size()) - 1;
    }
    void end() {
        if (!c.profiling || open < 0) return;
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

template <typename V>
spg_status multiply_impl(Context& c, int count, const spg_csr* A, const spg_csr* B, spg_csr* C) {
    const int di = sizeof(V) == 8 ? 1 : 0;
    hipStream_t s = c.stream;
    prof_reset(c);
    Prof prof(c);
    for (int i = 0; i < count; ++i) {
        SPG_TRY(check_operands(A[i], B[i]));
        C[i] = spg_csr{0, A[i].m, B[i].n, nullptr, nullptr, nullptr};
    }
    // output row pointers
    std::vector<void*> owned;
    auto cleanup = [&]() {
        for (void* p : owned) (void)dev_release(c, p);
        for (int i = 0; i < count; ++i) C[i] = spg_csr{0, 0, 0, nullptr, nullptr, nullptr};
    };
    long long total_rows = 0;
    int n_tiles = 0;
    for (int i = 0; i < count; ++i) {
        void* p = nullptr;
        const spg_status st = dev_alloc(c, &p, sizeof(int32_t) * (static_cast<size_t>(A[i].m) + 1));
        if (st != SPG_OK) {
            cleanup();
            return st;
        }
        owned.push_back(p);
        C[i].rpt = static_cast<int32_t*>(p);
        total_rows += A[i].m;
        n_tiles += (A[i].m + kTileRows - 1) / kTileRows;
        if (A[i].m == 0) {
            const hipError_t e = hipMemsetAsync(C[i].rpt, 0, sizeof(int32_t), s);
            if (e != hipSuccess) {
                cleanup();
                return hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__);
            }
        }
    }
    if (total_rows >= INT_MAX) {
        cleanup();
        return fail(SPG_ERR_OVERFLOW, "total row count exceeds int32");
    }
    if (n_tiles == 0) {  // every product has zero rows
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
        }
        return SPG_OK;
    }
    const size_t M = static_cast<size_t>(total_rows);
    const bool batch = count > 1;

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
    const size_t o_it1 = take(sizeof(int2) * M);
    const size_t o_it2 = take(sizeof(int2) * M);
    const size_t o_fit = take(sizeof(int2) * M);
    const size_t o_foff = take(sizeof(long long) * M);
    const size_t o_goff = take(sizeof(long long) * M);
    const size_t o_probs = take(batch ? sizeof(Prob<V>) * count : 0);
    const size_t o_tiles = take(batch ? sizeof(Tile) * n_tiles : 0);
    {
        const spg_status st = ensure_scratch(c, off);
        if (st != SPG_OK) {
            cleanup();
            return st;
        }
    }
    char* base = c.scratch;
    DevMeta* meta = reinterpret_cast<DevMeta*>(base + o_meta);
    long long* tile_sums = reinterpret_cast<long long*>(base + o_tsum);
    long long* inst_nnz = reinterpret_cast<long long*>(base + o_inst);
    int32_t* nip = reinterpret_cast<int32_t*>(base + o_nip);
    int2* items1 = reinterpret_cast<int2*>(base + o_it1);
    int2* items2 = reinterpret_cast<int2*>(base + o_it2);
    int2* fail_items = reinterpret_cast<int2*>(base + o_fit);
    long long* fail_off = reinterpret_cast<long long*>(base + o_foff);
    long long* gpool_off = reinterpret_cast<long long*>(base + o_goff);
    Prob<V>* d_probs = reinterpret_cast<Prob<V>*>(base + o_probs);
    Tile* d_tiles = reinterpret_cast<Tile*>(base + o_tiles);

    // descriptors
    std::vector<Prob<V>> probs(count);
    std::vector<Tile> tiles;
    {
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
    }
    ProbSet<V> ps{};
    ps.p0 = probs[0];
    ps.n_inst = count;
    ps.n_tiles = n_tiles;
    ps.many = batch ? d_probs : nullptr;
    ps.tiles = batch ? d_tiles : nullptr;
    size_t stage_probs = 0, stage_tiles = 0, stage_inst = 0;
    if (batch) {
        stage_probs = 0;
        stage_tiles = align256(sizeof(Prob<V>) * count);
        stage_inst = stage_tiles + align256(sizeof(Tile) * n_tiles);
        const spg_status st = ensure_stage(c, stage_inst + align256(sizeof(long long) * count));
        if (st != SPG_OK) {
            cleanup();
            return st;
        }
        std::memcpy(c.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);
        std::memcpy(c.h_stage + stage_tiles, tiles.data(), sizeof(Tile) * n_tiles);
    }

    auto run_front = [&]() -> spg_status {
        if (batch) {
            SPG_HIP(hipMemcpyAsync(d_probs, c.h_stage + stage_probs, sizeof(Prob<V>) * count, hipMemcpyHostToDevice,
                                   s));
            SPG_HIP(hipMemcpyAsync(d_tiles, c.h_stage + stage_tiles, sizeof(Tile) * n_tiles, hipMemcpyHostToDevice,
                                   s));
        }
        prof.begin("clear", 0, 1);
        launch_clear(meta, 0, s);
        prof.end();
        prof.begin("nip", total_rows, n_tiles);
        launch_nip<V>(ps, c.sym, meta, s);
        prof.end();
        prof.begin("scatter_sym", total_rows, n_tiles);
        launch_scatter<V>(ps, c.sym, meta, items1, s);
        prof.end();
        // The largest-table bin runs on an aux stream beside the multi-bin
        // launch (it is queued first: its workgroups without a row exit at
        // once and leave the CUs to the multi-bin grid).
        SPG_HIP(hipEventRecord(c.ev_ready, s));
        SPG_HIP(hipStreamWaitEvent(c.aux[0], c.ev_ready, 0));
        {
            const int grid = c.cus * c.occ_sym_max;
            prof.begin("sym_max", 0, grid, c.aux[0]);
            launch_sym_max<V>(ps, items1, c.sym, c.sym.n - 1, meta, fail_items, fail_off, grid, c.aux[0]);
            prof.end();
        }
        const int n_multi = c.sym.n - 1;
        if (n_multi > 0) {
            int grid = c.cus * c.occ_sym_multi;
            grid = (grid + kXcds - 1) / kXcds * kXcds;
            prof.begin("sym_multi", total_rows, grid);
            launch_sym_multi<V>(ps, items1, c.sym, n_multi, meta, grid, s);
            prof.end();
        }
        SPG_HIP(hipEventRecord(c.ev_done[0], c.aux[0]));
        SPG_HIP(hipStreamWaitEvent(s, c.ev_done[0], 0));
        return SPG_OK;
    };
    auto run_scan = [&](int after_fail) -> spg_status {
        prof.begin("scan_count", total_rows, n_tiles);
        launch_scan_count<V>(ps, c.num, meta, tile_sums, after_fail, s);
        prof.end();
        prof.begin("scan_apply", total_rows, n_tiles);
        launch_scan_apply<V>(ps, c.num, meta, tile_sums, items2, gpool_off, inst_nnz, after_fail, s);
        prof.end();
        SPG_HIP(hipMemcpyAsync(c.h_meta, meta, sizeof(DevMeta), hipMemcpyDeviceToHost, s));
        if (batch)
            SPG_HIP(hipMemcpyAsync(c.h_stage + stage_inst, inst_nnz, sizeof(long long) * count,
                                   hipMemcpyDeviceToHost, s));
        SPG_HIP(hipStreamSynchronize(s));
        return SPG_OK;
    };

    spg_status st = run_front();
    if (st == SPG_OK) st = run_scan(0);
    void* sym_pool = nullptr;
    if (st == SPG_OK && c.h_meta->fail_count.v > 0) {
        // rows that overflowed the largest LDS table: count them again in
        // global memory, then redo the scan stage
        const int fails = c.h_meta->fail_count.v;
        st = dev_alloc(c, &sym_pool, sizeof(int32_t) * c.h_meta->fail_pool.v);
        if (st == SPG_OK) {
            prof.begin("sym_global", fails, fails);
            launch_sym_global<V>(ps, fail_items, fail_off, static_cast<int*>(sym_pool), fails, s);
            prof.end();
            launch_clear(meta, kMetaScanWord, s);
            st = run_scan(1);
        }
        if (sym_pool) (void)dev_release(c, sym_pool);
    }
    if (st != SPG_OK) {
        cleanup();
        return st;
    }
    const DevMeta hm = *c.h_meta;
    if (hm.flags.v & kFlagNipOverflow) {
        cleanup();
        return fail(SPG_ERR_OVERFLOW, "a row has 2^31 or more intermediate products");
    }
    if (hm.flags.v & kFlagNnzOverflow) {
        cleanup();
        return fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32");
    }

    // exact-size outputs
    for (int i = 0; i < count; ++i) {
        const long long nnz = batch ? reinterpret_cast<long long*>(c.h_stage + stage_inst)[i]
                                    : (A[i].m > 0 ? hm.nnz_total.v : 0);
        if (A[i].m == 0) continue;
        if (nnz >= INT_MAX) {
            cleanup();
            return fail(SPG_ERR_OVERFLOW, "nnz(C) does not fit int32");
        }
        C[i].nnz = static_cast<int32_t>(nnz);
        void *pc = nullptr, *pv = nullptr;
        st = dev_alloc(c, &pc, sizeof(int32_t) * nnz);
        if (st == SPG_OK) owned.push_back(pc);
        if (st == SPG_OK) st = dev_alloc(c, &pv, sizeof(V) * nnz);
        if (st == SPG_OK) owned.push_back(pv);
        if (st != SPG_OK) {
            cleanup();
            return st;
        }
        C[i].col = static_cast<int32_t*>(pc);
        C[i].val = pv;
        probs[i].c_col = C[i].col;
        probs[i].c_val = static_cast<V*>(pv);
    }
    ps.p0 = probs[0];
    if (batch) {
        std::memcpy(c.h_stage + stage_probs, probs.data(), sizeof(Prob<V>) * count);
        const hipError_t e =
            hipMemcpyAsync(d_probs, c.h_stage + stage_probs, sizeof(Prob<V>) * count, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(e, "hipMemcpyAsync", __FILE__, __LINE__);
        }
    }

    // fused pass: one launch per non-empty bin, heaviest bins first, spread
    // over the context's streams so that small bins overlap
    struct Launch {
        int variant, off, count;
    };
    std::vector<Launch> launches;
    int glob_bin = -1, glob_off = 0;
    {
        int item_off = 0;
        for (int b = 0; b < c.num.n; ++b) {
            const int cnt = hm.num_count[b].v;
            const int v = c.num.variant[b];
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
    std::stable_sort(launches.begin(), launches.end(), [](const Launch& x, const Launch& y) {
        return kNumVariants[x.variant].p * kNumVariants[x.variant].wg / kNumVariants[x.variant].r >
               kNumVariants[y.variant].p * kNumVariants[y.variant].wg / kNumVariants[y.variant].r;
    });
    const int n_launch = static_cast<int>(launches.size()) + (glob_bin >= 0 ? 1 : 0);
    const int n_streams = std::min(n_launch, static_cast<int>(Context::kStreams));
    auto stream_of = [&](int k) { return k % Context::kStreams == 0 ? s : c.aux[k % Context::kStreams - 1]; };
    void* num_pool = nullptr;
    if (glob_bin >= 0) {
        st = dev_alloc(c, &num_pool, sizeof(int32_t) * hm.gnum_pool.v);
        if (st != SPG_OK) {
            cleanup();
            return st;
        }
    }
    hipError_t e = hipSuccess;
    if (n_streams > 1) {  // C.col, C.val and the pool were allocated on the main stream
        e = hipEventRecord(c.ev_ready, s);
        for (int i = 1; i < n_streams && e == hipSuccess; ++i) e = hipStreamWaitEvent(c.aux[i - 1], c.ev_ready, 0);
    }
    int k = 0;
    if (e == hipSuccess && glob_bin >= 0) {
        const hipStream_t on = stream_of(k++);
        prof.begin("symnum_global", hm.num_count[glob_bin].v, hm.num_count[glob_bin].v, on);
        launch_symnum_global<V>(ps, items2, gpool_off, glob_off, hm.num_count[glob_bin].v,
                                static_cast<int*>(num_pool), meta, on);
        prof.end();
    }
    for (const Launch& L : launches) {
        if (e != hipSuccess) break;
        const NumVariant& nv = kNumVariants[L.variant];
        const int units = (L.count + nv.r - 1) / nv.r;
        int grid = std::min(units, c.cus * c.occ_num[di][L.variant]);
        grid = (grid + kXcds - 1) / kXcds * kXcds;
        const hipStream_t on = stream_of(k++);
        prof.begin("symnum_v" + std::to_string(L.variant), L.count, grid, on);
        e = launch_symnum_bin<V>(L.variant, ps, items2, L.off, L.count, meta, grid, on);
        prof.end();
    }
    for (int i = 1; i < n_streams && e == hipSuccess; ++i) {
        e = hipEventRecord(c.ev_done[i - 1], c.aux[i - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, c.ev_done[i - 1], 0);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(&c.h_meta->flags.v, &meta->flags.v, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (num_pool) (void)dev_release(c, num_pool);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "fused pass", __FILE__, __LINE__);
    }
    if (c.h_meta->flags.v & kFlagMismatch) {
        cleanup();
        return fail(SPG_ERR_INTERNAL, "fused pass found a different column count than the count pass "
                                      "(are B's rows free of duplicate columns and indices in range?)");
    }
    if (c.profiling) prof_collect(c);
    return SPG_OK;
}

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
    SPG_TRY(dev_release(*c, C->rpt));
    SPG_TRY(dev_release(*c, C->col));
    SPG_TRY(dev_release(*c, C->val));
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
    SPG_HIP(hipMalloc(ptr, bytes));
    return SPG_OK;
}

spg_status spg_free(void* ptr) {
    if (ptr) SPG_HIP(hipFree(ptr));
    return SPG_OK;
}

spg_status spg_memcpy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPG_OK;
    if (!dst || !src) return fail(SPG_ERR_INVALID_ARGUMENT, "null argument");
    SPG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return SPG_OK;
}

spg_status spg_device_sync(void) {
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

}  // extern "C"
