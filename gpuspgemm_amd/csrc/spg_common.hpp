// Shared definitions of the gfx950 SpGEMM kernels and their host driver.
//
// Pipeline of one multiply (reference phases in brackets, paths relative to
// src/libproposal/include/proposal/):
//   k_clear      zero the device counters (first use and   [setup.cuh:146-150]
//                after errors; a completed multiply leaves them zeroed)
//   k_nip        NIP per row, max NIP, rows -> SYM bin     [setup.cuh:29-95, binning.cuh:39-120]
//                lists (one region per bin)
//   k_sym_multi  exact nnz per row by LDS hash tables      [sym1.cuh:28-167]
//   k_sym_max    largest LDS table with overflow detection [sym1.cuh:169-245]
//   k_sym_global global-memory tables for overflowed rows  [sym1.cuh:247-306]
//   k_scan_count / k_scan_apply
//                exclusive scan nnz -> row_ptr, NUM bins   [binning.cuh:208-243]
//   k_symnum_*   sorted columns and values of each row in one pass: the
//                hash fill and in-LDS sort of sym2.cuh:27-288 followed, on
//                the same LDS-resident sorted column list, by the
//                binary-search accumulation of num.cuh:28-298.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace spg {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kEmpty = -1;         // empty hash slot (columns are >= 0)
constexpr int kMaxBins = 16;
constexpr int kTileRows = 2048;    // rows per tile of the row-parallel passes
constexpr int kRowThreads = 1024;  // threads of the row-parallel passes (few workgroups: few atomics per counter)
constexpr int kXcds = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMultiThreads = 256; // workgroup size of the multi-bin kernels

// One product C = A . B.  `nip` is per-row scratch; c_rpt first holds nnz
// per row and then the row pointers (the reference reuses C.rpt the same
// way, proposal.cuh:40-56).  `bref` is per-entry scratch of A: the start
// and length of the row of B that entry e of A selects, recorded by the NIP
// pass (which reads those row pointers anyway), so that the later passes go
// from A's entries straight to B's columns, one dependent load fewer.
template <typename V>
struct Prob {
    const int32_t* a_rpt;
    const int32_t* a_col;
    const V* a_val;
    const int32_t* b_rpt;
    const int32_t* b_col;
    const V* b_val;
    int32_t* c_rpt;
    int32_t* c_col;
    V* c_val;
    int32_t* nip;
    long long* bref;
    int32_t m;
    int32_t n;
};

// An entry's row of B packed in one word: start in the low half, length in
// the high half.
__host__ __device__ __forceinline__ long long pack_bref(int start, int len) {
    return static_cast<long long>(static_cast<unsigned int>(start)) | (static_cast<long long>(len) << 32);
}
__host__ __device__ __forceinline__ int bref_start(long long r) { return static_cast<int>(r); }
__host__ __device__ __forceinline__ int bref_len(long long r) { return static_cast<int>(r >> 32); }

// Global address space.  Data pointers reach the kernels inside structs, so
// the compiler sees them as generic and would emit flat instructions, which
// also count against the LDS wait counter: every wait for an LDS result
// would then wait for all outstanding memory loads too.  Device code reads
// and writes through pointers cast to this address space instead.
#ifndef SPG_GLOBAL
#define SPG_GLOBAL __attribute__((address_space(1)))
#endif

// Minimum waves per SIMD a kernel must keep: caps its VGPRs at 512 / n.
#ifndef SPG_WAVES_PER_EU
#define SPG_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

template <typename T>
__device__ __forceinline__ T SPG_GLOBAL* to_global(T* p) {
    return (T SPG_GLOBAL*)p;
}

// Device view of one product: the pointers of Prob in the global address space.
template <typename V>
struct GProb {
    const int32_t SPG_GLOBAL* a_rpt;
    const int32_t SPG_GLOBAL* a_col;
    const V SPG_GLOBAL* a_val;
    const int32_t SPG_GLOBAL* b_rpt;
    const int32_t SPG_GLOBAL* b_col;
    const V SPG_GLOBAL* b_val;
    int32_t SPG_GLOBAL* c_rpt;
    int32_t SPG_GLOBAL* c_col;
    V SPG_GLOBAL* c_val;
    int32_t SPG_GLOBAL* nip;
    long long SPG_GLOBAL* bref;
    int32_t m;
    int32_t n;
};

// (R: a Prob lvalue in any address space)
template <typename V, typename R>
__device__ __forceinline__ GProb<V> global_view(R&& p) {
    return GProb<V>{to_global(p.a_rpt), to_global(p.a_col), to_global(p.a_val), to_global(p.b_rpt),
                    to_global(p.b_col), to_global(p.b_val), to_global(p.c_rpt), to_global(p.c_col),
                    to_global(p.c_val), to_global(p.nip),   to_global(p.bref),  p.m,
                    p.n};
}

// Tile of a row-parallel pass: instance, first row, row count, index of the
// instance's first tile (read on the device as one int4).
struct alignas(16) Tile {
    int32_t inst, row0, rows, first_tile;
};

// All products of one pass.  A single product travels in kernel arguments
// (`many` null); a batch, or any call replayed from a graph, is read from
// device memory, so that the kernel arguments stay the same across calls.
template <typename V>
struct ProbSet {
    Prob<V> p0;
    const Prob<V>* many;
    const Tile* tiles;
    int32_t n_inst;
    int32_t n_tiles;

    __device__ __forceinline__ GProb<V> get(int i) const {
        if (many == nullptr) return global_view<V>(p0);
        return global_view<V>(to_global(many)[i]);
    }
    __device__ __forceinline__ Tile tile(int t) const {
        if (tiles == nullptr) {
            const int r0 = t * kTileRows;
            return Tile{0, r0, min(kTileRows, p0.m - r0), 0};
        }
        typedef int32_t Words __attribute__((vector_size(16)));
        const Words v = to_global(reinterpret_cast<const Words*>(tiles))[t];
        return Tile{v[0], v[1], v[2], v[3]};
    }
};

// A row of the fused pass as the scan stage hands it over: its place and
// the values a kernel reads first, so that the kernel's first dependent
// load is the row's entries of A.
struct alignas(16) RowItem {
    int32_t inst, row;
    int32_t a0, a1;  // the row's entries of A
    int32_t c0, cn;  // its first position in C and its nnz
    int32_t nip, pad;
};

__device__ __forceinline__ RowItem load_item(const RowItem* items, long long i) {
    typedef int32_t Words __attribute__((vector_size(16)));
    const auto w = to_global(reinterpret_cast<const Words*>(items + i));
    const Words x = w[0], y = w[1];
    return RowItem{x[0], x[1], x[2], x[3], y[0], y[1], y[2], y[3]};
}

// Runtime bin table of one phase: row -> first bin whose caps admit it.
struct BinTable {
    int32_t n;
    int32_t variant[kMaxBins];
    int32_t rows[kMaxBins];  // rows per workgroup unit of the bin's variant
    int32_t nip_cap[kMaxBins];
    int32_t nnz_cap[kMaxBins];
};

__host__ __device__ __forceinline__ int find_bin(const BinTable& t, int nip, int nnz) {
    for (int b = 0; b < t.n; ++b)
        if (nip <= t.nip_cap[b] && nnz <= t.nnz_cap[b]) return b;
    return t.n - 1;
}

enum MetaFlags : int32_t {
    kFlagNipOverflow = 1,   // a row's NIP >= 2^31
    kFlagNnzOverflow = 2,   // nnz(C) >= 2^31
    kFlagMismatch = 4,      // fused pass found a different nnz than the count pass
};

// A device counter on a 128-byte line of its own.  Counters are updated
// atomically by many workgroups; atomics from all XCDs to one line are
// serialised (about 10 ns each), so no two counters share a line.
template <typename T>
struct alignas(128) Counter {
    T v;
};

// Device counters of one multiply (zeroed by k_clear, or by the read-back
// that ends the multiply before).  The second group
// belongs to the scan / fused-bin stage and is cleared again on its own
// when overflowed rows force that stage to run twice.
struct DevMeta {
    Counter<int32_t> sym_count[kMaxBins];
    Counter<int32_t> fail_count;
    Counter<int32_t> flags;
    Counter<int32_t> max_nip;
    Counter<unsigned long long> total_nip;
    Counter<unsigned long long> fail_pool;  // int32 slots of the global tables of overflowed rows
    // ---- scan stage
    Counter<int32_t> num_count[kMaxBins];
    Counter<int32_t> num_cursor[kMaxBins];
    Counter<int32_t> max_nnz;
    Counter<unsigned long long> gnum_pool;    // int32 slots needed by the global fused bin
    Counter<unsigned long long> gnum_cursor;  // slots handed out so far
    Counter<long long> nnz_total;             // nnz(C) of a single product
};
constexpr int kMetaScanWord = static_cast<int>(offsetof(DevMeta, num_count) / sizeof(int32_t));

// ---------------------------------------------------------------- device helpers

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}



__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Compiler-level ordering of LDS accesses among the lanes of one wave (a
// wave's LDS instructions execute in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int T>
__device__ __forceinline__ uint32_t hash_slot(int key) {
    static_assert((T & (T - 1)) == 0 && T >= 2, "table size must be a power of two");
    constexpr int kBits = __builtin_ctz(T);
    return (static_cast<uint32_t>(key) * 2654435761u) >> (32 - kBits);
}

__device__ __forceinline__ uint32_t hash_slot_rt(int key, int bits) {
    return (static_cast<uint32_t>(key) * 2654435761u) >> (32 - bits);
}

__device__ __forceinline__ int lds_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Open-addressing insert with linear probing into an LDS table of T slots,
// continuing from slot h whose content `cur` has already been read.
// Returns 1 when the key was not present.  The table never fills: callers
// size T above the number of distinct keys of the row.
template <int T>
__device__ __forceinline__ int table_insert_from(int* tab, int key, uint32_t h, int cur) {
    while (true) {
        if (cur == key) return 0;
        if (cur == kEmpty) {
            const int old = atomicCAS(tab + h, kEmpty, key);
            if (old == kEmpty) return 1;
            if (old == key) return 0;
        }
        h = (h + 1) & (T - 1);
        cur = lds_load(tab + h);
    }
}

template <int T>
__device__ __forceinline__ int table_insert(int* tab, int key) {
    const uint32_t h = hash_slot<T>(key);
    return table_insert_from<T>(tab, key, h, lds_load(tab + h));
}

// Insert up to N staged keys: the first probes of all keys are read
// together (independent LDS loads), then each key finishes its probe.
template <int T, int N>
__device__ __forceinline__ int table_insert_all(int* tab, const int (&key)[N], const bool (&ok)[N]) {
    uint32_t h[N];
    int cur[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        h[j] = hash_slot<T>(key[j]);
        cur[j] = ok[j] ? lds_load(tab + h[j]) : key[j];
    }
    int added = 0;
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (ok[j] && cur[j] != key[j]) added += table_insert_from<T>(tab, key[j], h[j], cur[j]);
    return added;
}

// Sum over the L consecutive lanes of a row group (L a power of two <= 64).
template <int L>
__device__ __forceinline__ int group_sum(int v) {
    if constexpr (L > 1) {
#pragma unroll
        for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    }
    return v;
}

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// Wave-aggregated LDS counters: every lane of the wave calls with its bin
// (negative: none); one atomic per distinct bin of the wave instead of one
// per lane on the same address.  wave_rank_add also returns each lane's
// rank among the lanes counted before it.
__device__ __forceinline__ void wave_count_add(int* cnt, int bin) {
    unsigned long long todo = __ballot(bin >= 0);
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int b = __shfl(bin, leader);
        const unsigned long long m = __ballot(bin == b);
        if (lane_id() == leader) atomicAdd(&cnt[b], __popcll(m));
        todo &= ~m;
    }
}

__device__ __forceinline__ int wave_rank_add(int* cnt, int bin) {
    int rank = 0;
    unsigned long long todo = __ballot(bin >= 0);
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const int b = __shfl(bin, leader);
        const unsigned long long m = __ballot(bin == b);
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&cnt[b], __popcll(m));
        base = __shfl(base, leader);
        if (bin == b) rank = base + __popcll(m & ((1ull << lane_id()) - 1));
        todo &= ~m;
    }
    return rank;
}

// Lower bounds of keys[J0, J0 + N), searched in lockstep: each halving step
// reads the probes of all N keys at once, so a lane waits log2(P) LDS round
// trips for the N keys instead of N * log2(P).
template <int P, int N, int J0 = 0, int K>
__device__ __forceinline__ void lower_bound_all(const int* cols, const int (&key)[K], int (&pos)[N]) {
    static_assert(J0 + N <= K, "key range out of bounds");
#pragma unroll
    for (int j = 0; j < N; ++j) pos[j] = 0;
#pragma unroll
    for (int step = P >> 1; step > 0; step >>= 1) {
        int probe[N];
#pragma unroll
        for (int j = 0; j < N; ++j) probe[j] = cols[pos[j] + step - 1];
#pragma unroll
        for (int j = 0; j < N; ++j) pos[j] += probe[j] < key[J0 + j] ? step : 0;
    }
}

template <typename V>
__device__ __forceinline__ void lds_add(V* p, V v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// XCD-aware share of a unit range [0, total) for this workgroup: each XCD
// takes one contiguous slice so that neighbouring rows (which read the same
// rows of B) meet in the same L2.  gridDim.x is a multiple of kXcds.
struct UnitRange {
    int first, end, stride;
};
__device__ __forceinline__ UnitRange xcd_units(int total) {
    const int xcd = blockIdx.x % kXcds;
    const int local = blockIdx.x / kXcds;
    const int per_xcd = gridDim.x / kXcds;
    const int slice = (total + kXcds - 1) / kXcds;
    const int lo = xcd * slice;
    const int hi = min(total, lo + slice);
    return UnitRange{lo + local, hi, per_xcd};
}

}  // namespace spg
