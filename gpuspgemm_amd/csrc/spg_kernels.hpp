// Compiled kernel variants and host-side launchers.
//
// A variant fixes how a workgroup works on rows: R rows per workgroup (so
// L = WG / R lanes per row), A entries per chunk, products per lane per
// round, and the LDS table size.  The runtime bin tables (BinTable) map
// rows to variants; they replace the reference's hand-set per-arch tables
// (src/libproposal/include/proposal/parameters.cuh:22-413) and can be
// rewritten at run time by the tuner.
#pragma once

#include "spg_common.hpp"

namespace spg {

struct SymVariant {
    int wg, r, t, ec, pmax;  // workgroup, rows per workgroup, table slots, A entries per chunk, products per lane
};
struct NumVariant {
    int wg, r, p, ec, pmax;  // ... p = sorted-list capacity (the hash table has 2p slots)
};

// Symbolic (count) variants.  0..8 run inside one multi-bin launch of 256
// threads; 9 is the largest LDS table with overflow detection
// (sym1.cuh:169-245).  A bin may take rows of NIP below the table size.
constexpr SymVariant kSymVariants[] = {
    {256, 256, 16, 4, 8},  {256, 128, 32, 4, 8},  {256, 64, 64, 8, 8},    {256, 32, 128, 16, 8},
    {256, 16, 256, 32, 8}, {256, 4, 512, 64, 8},  {256, 4, 1024, 64, 12}, {256, 1, 4096, 256, 8},
    {256, 1, 8192, 256, 8}, {1024, 1, 32768, 1024, 4},
};
constexpr int kSymMultiVariants = 9;
constexpr int kSymMaxVariant = 9;
constexpr int kSymMaxThreshold = 26214;  // distinct keys admitted in the 32768-slot table (80 %)
constexpr int kSymMultiLds = 36 * 1024;  // bytes: max over variants 0..8 of r * sym_row_ints * 4
// The multi-bin variants run in three launches ("parts"), each with the
// register and LDS budget of its own variants: small tables (0..3), mid
// tables (4..6) and the two largest tables (7, 8; a row per workgroup).
// A launch's LDS is its largest variant's, so keeping the 8192-slot table
// apart lets the mid tables (the 96^3 stencil's rows) run at 6 waves per
// SIMD instead of 4 (80 VGPRs, 8 bytes spilled per lane; measured faster
// than 5 waves without a spill).  kSymMultiMinWaves is each part's
// residency floor, as for kNumMinWaves.
constexpr int kSymParts = 3;
constexpr int kSymPartLo[kSymParts] = {0, 4, 7};
constexpr int kSymPartHi[kSymParts] = {4, 7, 9};
constexpr int kSymMultiMinWaves[kSymParts] = {7, 6, 4};
constexpr int sym_part(int variant) { return variant < kSymPartHi[0] ? 0 : (variant < kSymPartHi[1] ? 1 : 2); }

// Fused sorted-columns + numeric variants, one kernel each: all use 256
// threads except 9 (a 1024-thread workgroup per row); rows above every
// capacity use global-memory tables (variant 19).  A row whose products fit
// in one chunk and in lanes * pmax keeps them in registers between the two
// halves.  Variants 10..18 cover the small capacities with 2 to 32 lanes
// per row; the tuner picks among them per bin.
constexpr NumVariant kNumVariants[] = {
    {256, 256, 8, 4, 16},   {256, 128, 16, 4, 16}, {256, 64, 32, 8, 16},  {256, 32, 64, 16, 16},
    {256, 16, 128, 32, 16}, {256, 4, 128, 64, 12}, {256, 4, 512, 64, 16}, {256, 1, 1024, 256, 8},
    {256, 1, 2048, 256, 8}, {1024, 1, 8192, 1024, 4}, {256, 32, 8, 8, 2}, {256, 16, 16, 16, 2},
    {256, 16, 32, 16, 4},   {256, 8, 64, 32, 4},   {256, 64, 8, 8, 4},    {256, 128, 8, 4, 8},
    {256, 32, 16, 8, 4},    {256, 64, 16, 8, 8},   {256, 32, 32, 16, 8},
};
constexpr int kNumLdsVariants = 19;  // variants with an LDS kernel (0..18)
// Waves per SIMD that each fused variant's kernel must keep (its register
// budget is 512 / waves VGPRs).  Fetching a lane's products as one batch
// of loads (every load in flight before the first multiply) needs more
// registers than fetching them one by one; the floor keeps the residency
// the variant had with the one-by-one fetch, and the compiler schedules
// the batch within it.  Measured on gfx950: the 96^3 stencil's fused pass
// 3.23 -> 2.81 ms, the power-law's 198k-row bin 70 -> 53 us.
constexpr int kNumMinWaves[kNumLdsVariants] = {3, 3, 3, 3, 3, 4, 3, 4, 3, 4, 8, 8, 7, 6, 7, 5, 7, 5, 5};
constexpr int kNumGlobalVariant = 19;

// Fused variants whose bins share one launch.  Set 0, small rows (256
// threads, up to 32 lanes per row, about 64 VGPRs): their bins hold most
// rows of skewed matrices, and one launch keeps the GPU filled across them
// where separate launches would queue behind each other on the lane's few
// hardware queues.  Set 1, the 256-thread variants of longer rows: merged
// only while their bins are small (a few thousand units, as the heavy tail
// of a skewed matrix), so that those long-latency units all start at once
// beside set 0 instead of queueing one launch after another; a large bin
// of set 1 (the 96^3 stencil's) keeps its own launch and its own register
// budget.  kNumMultiMinWaves is each merged kernel's residency floor.
constexpr int kNumMultiSet0[] = {10, 11, 12, 13, 14, 16};
constexpr int kNumMultiSet1[] = {4, 5, 6, 7, 8};
constexpr int kNumMultiMinWaves[2] = {6, 3};
constexpr int kNumMultiMaxUnits1 = 4096;  // set 1 merges while its bins hold at most this many units
constexpr int num_multi_set(int v) {
    for (int m : kNumMultiSet0)
        if (m == v) return 0;
    for (int m : kNumMultiSet1)
        if (m == v) return 1;
    return -1;
}

// Bins of one merged launch: bin b's rows are items [off[b], off[b] +
// count[b]) in units of its variant's rows.  Every XCD takes one contiguous
// slice of every bin: the workgroups with local index (blockIdx.x / 8) in
// [slice0[b], slice0[b + 1]) work on bin b, slice0[b + 1] - slice0[b] =
// ceil(units of bin b / 8) of them per XCD.
struct NumMultiBins {
    int32_t n;
    int32_t variant[kMaxBins];
    int32_t off[kMaxBins];
    int32_t count[kMaxBins];
    int32_t slice0[kMaxBins + 1];
};

constexpr int kGlobalSortChunk = 4096;  // LDS chunk of the global-bin column sort

// Ints of one row slot in LDS: `bytes` rounded up to 8-byte words and
// padded so that the number of words is odd.  Slots of one wave then start
// in different LDS banks; with a power-of-two stride the same offset of
// every row of the wave falls in one bank, and the small-row variants (up
// to 64 rows per wave) serialise on bank conflicts.
constexpr int padded_row_ints(int bytes) {
    const int words = (bytes + 7) / 8;
    return 2 * (words % 2 ? words : words + 1);
}
// Fused variant (FP64 sizing): table 2p ints (reused for p values), sorted
// list p ints, chunk metadata when the row has more than one lane.
constexpr int num_row_ints(const NumVariant& v) {
    return padded_row_ints(12 * v.p + (v.wg / v.r > 1 ? 16 * v.ec : 0));
}
// Symbolic variant: table t ints, chunk metadata without values.
constexpr int sym_row_ints(const SymVariant& v) {
    return padded_row_ints(4 * v.t + (v.wg / v.r > 1 ? 8 * v.ec + 8 : 0));
}
constexpr int max_sym_multi_lds() {
    int m = 0;
    for (int i = 0; i < kSymMultiVariants; ++i) {
        const int b = 4 * kSymVariants[i].r * sym_row_ints(kSymVariants[i]);
        m = m > b ? m : b;
    }
    return m;
}
constexpr int max_num_lds() {
    int m = 0;
    for (int i = 0; i < kNumLdsVariants; ++i) {
        const int b = 4 * kNumVariants[i].r * num_row_ints(kNumVariants[i]);
        m = m > b ? m : b;
    }
    return m;
}
constexpr int max_num_multi_lds(int set) {
    int m = 0;
    for (int v = 0; v < kNumLdsVariants; ++v) {
        const int b = 4 * kNumVariants[v].r * num_row_ints(kNumVariants[v]);
        if (num_multi_set(v) == set) m = m > b ? m : b;
    }
    return m;
}
constexpr int kNumBigLds = max_num_lds();
static_assert(max_sym_multi_lds() <= kSymMultiLds, "symbolic multi-bin LDS budget");

// Slots of a global-memory hash table for a row with `keys` distinct keys.
__host__ __device__ __forceinline__ long long global_table_slots(long long keys) {
    long long t = kGlobalSortChunk;
    while (t < 2 * keys) t <<= 1;
    return t;
}

// ---- launchers (defined next to their kernels) ----
void launch_clear(DevMeta* meta, int from_word, hipStream_t s);
void launch_signal(int* host_flag, hipStream_t s);
void launch_readback(const int* src, int* host_dst, int words, const long long* src2, long long* host_dst2, int n2,
                     DevMeta* clear, int* host_flag, hipStream_t s);
void launch_total_nip(const int32_t* a_rpt, const int32_t* a_col, const int32_t* b_rpt, int m,
                      unsigned long long* out, hipStream_t s);

// NIP per row and the scatter into the symbolic bin lists (bin b's list
// starts at items + b * bin_cap)
template <typename V>
void launch_nip(const ProbSet<V>& ps, const BinTable& sym, DevMeta* meta, int2* items, int bin_cap, hipStream_t s);
template <typename V>
void launch_scan_count(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, long long* tile_sums,
                       int after_fail, hipStream_t s);
// with tile_prefix_min or more tiles, k_tile_prefix turns tile_sums into
// their exclusive prefix first (in place)
template <typename V>
void launch_scan_apply(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, long long* tile_sums,
                       RowItem* items, long long* gpool_off, long long* inst_nnz, int after_fail,
                       int tile_prefix_min, hipStream_t s);

// symbolic count: multi-bin launch over bins [0, n_multi) of `sym`
// part p: variants [kSymPartLo[p], kSymPartHi[p])
template <typename V>
void launch_sym_multi(int part, const ProbSet<V>& ps, const int2* items, int bin_cap, const BinTable& sym,
                      int n_multi, const DevMeta* meta, int grid, hipStream_t s);
template <typename V>
void launch_sym_max(const ProbSet<V>& ps, const int2* items, int bin_cap, const BinTable& sym, int bin,
                    DevMeta* meta, int2* fail_items, long long* fail_off, int grid, hipStream_t s);
template <typename V>
void launch_sym_global(const ProbSet<V>& ps, const int2* fail_items, const long long* fail_off, int* pool,
                       int count, hipStream_t s);

// fused pass: one launch per variant (0..9) over one bin's items
template <typename V>
hipError_t launch_symnum_bin(int variant, const ProbSet<V>& ps, const RowItem* items, int item_off, int count,
                             DevMeta* meta, int grid, hipStream_t s);
// fused pass: one launch over the units of several bins whose variants are
// in merged set `set` (a workgroup per unit)
template <typename V>
hipError_t launch_symnum_multi(int set, const ProbSet<V>& ps, const RowItem* items, const NumMultiBins& mb,
                               DevMeta* meta, hipStream_t s);
template <typename V>
void launch_symnum_global(const ProbSet<V>& ps, const RowItem* items, const long long* gpool_off, int item_off,
                          int count, int* pool, DevMeta* meta, hipStream_t s);

// occupancy (workgroups per CU) of the persistent launches
template <typename V>
int occupancy_sym_multi(int part);
template <typename V>
int occupancy_sym_max();
template <typename V>
int occupancy_symnum_bin(int variant);

}  // namespace spg
