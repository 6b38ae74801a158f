// Compiled kernel variants and host-side launchers.
//
// A variant fixes how a workgroup works on rows: R rows per workgroup,
// G lanes per entry of A's row (those G lanes stride one row of B), and the
// LDS table size.  The runtime bin tables (BinTable) map rows to variants;
// they replace the reference's hand-set per-arch tables
// (src/libproposal/include/proposal/parameters.cuh:22-413) and can be
// rewritten at run time by the tuner.
#pragma once

#include "spg_common.hpp"

namespace spg {

struct SymVariant {
    int wg, r, g, t;  // workgroup, rows per workgroup, lanes per A entry, table slots
};
struct NumVariant {
    int wg, r, g, p;  // ... p = sorted-list capacity (table has 2p slots)
};

// Symbolic (count) variants.  0..7 run inside one multi-bin launch of 256
// threads; 8 is the largest LDS table with overflow detection (sym1.cuh:169-245).
constexpr SymVariant kSymVariants[] = {
    {256, 256, 1, 16},  {256, 128, 1, 32}, {256, 64, 1, 64},  {256, 32, 2, 128},
    {256, 16, 4, 256},  {256, 4, 8, 512},  {256, 4, 16, 2048}, {256, 1, 32, 8192},
    {1024, 1, 64, 32768},
};
constexpr int kSymMultiVariants = 8;
constexpr int kSymMaxVariant = 8;
constexpr int kSymGlobalVariant = 9;     // rows that overflowed the largest table
constexpr int kSymMaxThreshold = 26214;  // distinct keys admitted in the 32768-slot table (80 %)
constexpr int kSymMultiLds = 32768;      // bytes: max over variants 0..7 of r * t * 4

// Fused sorted-columns + numeric variants.  0..8 share one multi-bin launch
// of 256 threads; 9 is a 1024-thread workgroup per row; rows above its
// capacity use global-memory tables.
constexpr NumVariant kNumVariants[] = {
    {256, 256, 1, 8},   {256, 128, 1, 16}, {256, 64, 1, 32},  {256, 32, 2, 64},
    {256, 16, 4, 128},  {256, 4, 16, 128}, {256, 4, 16, 512}, {256, 1, 32, 1024},
    {256, 1, 64, 2048}, {1024, 1, 64, 8192},
};
constexpr int kNumMultiVariants = 9;
constexpr int kNumBigVariant = 9;
constexpr int kNumGlobalVariant = 10;
constexpr int kNumMultiLds = 24576;  // bytes: max over variants 0..8 of r * 12 * p
constexpr int kNumBigLds = 12 * 8192;

constexpr int kGlobalSortChunk = 4096;  // LDS chunk of the global-bin column sort

// Slots of a global-memory hash table for a row with `keys` distinct keys.
__host__ __device__ __forceinline__ long long global_table_slots(long long keys) {
    long long t = kGlobalSortChunk;
    while (t < 2 * keys) t <<= 1;
    return t;
}

__host__ __device__ constexpr int sym_rows_per_unit(int v) { return kSymVariants[v].r; }
__host__ __device__ constexpr int num_rows_per_unit(int v) { return kNumVariants[v].r; }

// ---- launchers (defined next to their kernels) ----
void launch_clear(DevMeta* meta, int from_word, hipStream_t s);
void launch_total_nip(const int32_t* a_rpt, const int32_t* a_col, const int32_t* b_rpt, int m,
                      unsigned long long* out, hipStream_t s);

template <typename V>
void launch_nip(const ProbSet<V>& ps, const BinTable& sym, DevMeta* meta, hipStream_t s);
template <typename V>
void launch_scatter(const ProbSet<V>& ps, const BinTable& sym, DevMeta* meta, int2* items, hipStream_t s);
template <typename V>
void launch_scan_count(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, long long* tile_sums,
                       int after_fail, hipStream_t s);
template <typename V>
void launch_scan_apply(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, const long long* tile_sums,
                       int2* items, long long* gpool_off, long long* inst_nnz, int after_fail, hipStream_t s);

// symbolic count: multi-bin launch over bins [0, n_multi) of `sym`
template <typename V>
void launch_sym_multi(const ProbSet<V>& ps, const int2* items, const BinTable& sym, int n_multi,
                      const DevMeta* meta, int grid, hipStream_t s);
template <typename V>
void launch_sym_max(const ProbSet<V>& ps, const int2* items, const BinTable& sym, int bin, DevMeta* meta,
                    int2* fail_items, long long* fail_off, int grid, hipStream_t s);
template <typename V>
void launch_sym_global(const ProbSet<V>& ps, const int2* fail_items, const long long* fail_off, int* pool,
                       int count, hipStream_t s);

template <typename V>
void launch_symnum_multi(const ProbSet<V>& ps, const int2* items, const MultiArgs& ma, DevMeta* meta, int grid,
                         hipStream_t s);
template <typename V>
void launch_symnum_big(const ProbSet<V>& ps, const int2* items, int item_off, int count, DevMeta* meta,
                       int grid, hipStream_t s);
template <typename V>
void launch_symnum_global(const ProbSet<V>& ps, const int2* items, const long long* gpool_off, int item_off,
                          int count, int* pool, DevMeta* meta, hipStream_t s);

// occupancy (workgroups per CU) of the persistent launches
template <typename V>
int occupancy_sym_multi();
template <typename V>
int occupancy_sym_max();
template <typename V>
int occupancy_symnum_multi();
template <typename V>
int occupancy_symnum_big();

}  // namespace spg
