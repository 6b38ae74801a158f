// Symbolic count: exact nnz of every row of C from an open-addressing hash
// table of the row's B columns (sym1.cuh:28-306), written into c_rpt.
//
// Work decomposition: a workgroup unit holds R rows; each row owns L = WG/R
// consecutive lanes; those lanes split into L/G sub-groups that take the
// row's A entries in turn, and the G lanes of a sub-group stride the B row
// of that entry.  All tables live in LDS except for rows that overflow the
// largest table, which are recounted in global memory.
#include "spg_kernels.hpp"

namespace spg {

template <typename V, int WG, int R, int G, int T>
__device__ __forceinline__ void sym_rows(const ProbSet<V>& ps, const int2* __restrict__ items, int first,
                                         int end, int* smem, int* s_red) {
    constexpr int L = WG / R;
    constexpr int E = L / G;
    static_assert(WG % R == 0 && L % G == 0, "bad variant shape");
    static_assert(L <= kWave || R == 1, "a row spans several waves only when it owns the workgroup");
    const int slot = threadIdx.x / L;
    const int lr = threadIdx.x % L;
    const int sub = lr / G;
    const int g = lr % G;
    int* tab = smem + slot * T;
    for (int j = lr; j < T; j += L) tab[j] = kEmpty;
    if constexpr (L > kWave) {
        if (threadIdx.x == 0) *s_red = 0;
    }
    __syncthreads();

    const int idx = first + slot;
    const bool active = idx < end;
    const int2 it = active ? items[idx] : make_int2(0, 0);
    const Prob<V> P = ps.get(it.x);
    int mine = 0;
    if (active) {
        const int a1 = P.a_rpt[it.y + 1];
        for (int i = P.a_rpt[it.y] + sub; i < a1; i += E) {
            const int k = P.a_col[i];
            const int q1 = P.b_rpt[k + 1];
            for (int q = P.b_rpt[k] + g; q < q1; q += G) mine += table_insert<T>(tab, P.b_col[q]);
        }
    }
    int nnz;
    if constexpr (L <= kWave) {
        nnz = group_sum<L>(mine);
    } else {
        if (mine) atomicAdd(s_red, mine);
        __syncthreads();
        nnz = *s_red;
    }
    if (active && lr == 0) P.c_rpt[it.y] = nnz;
    __syncthreads();  // tables are reused by the next unit
}

#define SPG_SYM_CASE(v)                                                                                     \
    case v:                                                                                                 \
        sym_rows<V, kMultiThreads, kSymVariants[v].r, kSymVariants[v].g, kSymVariants[v].t>(ps, items, first, \
                                                                                            end, smem, &s_red); \
        break;

// One launch serves bins [0, n_multi) of the symbolic table.  Bin sizes are
// read on the device, so the launch needs no host round trip; the grid is
// persistent and walks the units of all bins.
template <typename V>
__global__ __launch_bounds__(kMultiThreads) void k_sym_multi(ProbSet<V> ps, const int2* __restrict__ items,
                                                             BinTable sym, int n_multi,
                                                             const DevMeta* __restrict__ meta) {
    __shared__ __align__(16) int smem[kSymMultiLds / 4];
    __shared__ int s_red;
    __shared__ int s_unit[kMaxBins + 1];
    __shared__ int s_off[kMaxBins];
    __shared__ int s_cnt[kMaxBins];
    if (threadIdx.x == 0) {
        int off = 0, units = 0;
        for (int b = 0; b < n_multi; ++b) {
            const int c = meta->sym_count[b];
            s_off[b] = off;
            s_cnt[b] = c;
            s_unit[b] = units;
            off += c;
            units += (c + sym.rows[b] - 1) / sym.rows[b];
        }
        s_unit[n_multi] = units;
    }
    __syncthreads();
    const UnitRange ur = xcd_units(s_unit[n_multi]);
    for (int u = ur.first; u < ur.end; u += ur.stride) {
        int b = 0;
        while (u >= s_unit[b + 1]) ++b;
        const int first = s_off[b] + (u - s_unit[b]) * sym.rows[b];
        const int end = s_off[b] + s_cnt[b];
        switch (sym.variant[b]) {
            SPG_SYM_CASE(0)
            SPG_SYM_CASE(1)
            SPG_SYM_CASE(2)
            SPG_SYM_CASE(3)
            SPG_SYM_CASE(4)
            SPG_SYM_CASE(5)
            SPG_SYM_CASE(6)
            SPG_SYM_CASE(7)
            default:
                break;
        }
    }
}

// Largest LDS table (32768 slots, 128 KiB) with overflow detection
// (sym1.cuh:169-245): a row whose distinct count passes 80 % of the table
// stops early and is queued, with its global-table slot range, for
// k_sym_global.
template <typename V>
__global__ __launch_bounds__(1024) void k_sym_max(ProbSet<V> ps, const int2* __restrict__ items, int bin,
                                                  DevMeta* __restrict__ meta, int2* __restrict__ fail_items,
                                                  long long* __restrict__ fail_off) {
    constexpr int T = kSymVariants[kSymMaxVariant].t;
    constexpr int G = kSymVariants[kSymMaxVariant].g;
    constexpr int WG = 1024;
    constexpr int E = WG / G;
    static_assert(T - kSymMaxThreshold > WG + 1, "table must keep room for in-flight inserts");
    __shared__ int tab[T];
    __shared__ int s_n;
    __shared__ int s_abort;
    int off = 0;
    for (int b = 0; b < bin; ++b) off += meta->sym_count[b];
    const int count = meta->sym_count[bin];
    const int sub = threadIdx.x / G;
    const int g = threadIdx.x % G;
    for (int u = blockIdx.x; u < count; u += gridDim.x) {
        for (int j = threadIdx.x; j < T; j += WG) tab[j] = kEmpty;
        if (threadIdx.x == 0) {
            s_n = 0;
            s_abort = 0;
        }
        __syncthreads();
        const int2 it = items[off + u];
        const Prob<V> P = ps.get(it.x);
        const int a1 = P.a_rpt[it.y + 1];
        for (int i = P.a_rpt[it.y] + sub; i < a1; i += E) {
            const int k = P.a_col[i];
            const int q1 = P.b_rpt[k + 1];
            bool stop = false;
            for (int q = P.b_rpt[k] + g; q < q1; q += G) {
                if (lds_load(&s_abort)) {
                    stop = true;
                    break;
                }
                if (table_insert<T>(tab, P.b_col[q]) && atomicAdd(&s_n, 1) >= kSymMaxThreshold)
                    __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (stop) break;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = s_n;
            if (n <= kSymMaxThreshold) {
                P.c_rpt[it.y] = n;
            } else {
                const int f = atomicAdd(&meta->fail_count, 1);
                fail_items[f] = it;
                fail_off[f] = static_cast<long long>(atomicAdd(
                    &meta->fail_pool, static_cast<unsigned long long>(global_table_slots(P.nip[it.y]))));
            }
        }
        __syncthreads();
    }
}

// Global-memory table per overflowed row (sym1.cuh:247-306), sized from the
// row's NIP (an upper bound of its distinct count) rather than from the
// largest NIP of the matrix.
template <typename V>
__global__ __launch_bounds__(1024) void k_sym_global(ProbSet<V> ps, const int2* __restrict__ fail_items,
                                                     const long long* __restrict__ fail_off, int* pool) {
    constexpr int WG = 1024;
    constexpr int G = 64;
    constexpr int E = WG / G;
    __shared__ int s_n;
    const int2 it = fail_items[blockIdx.x];
    const Prob<V> P = ps.get(it.x);
    const long long slots = global_table_slots(P.nip[it.y]);
    const int bits = __builtin_ctzll(static_cast<unsigned long long>(slots));
    int* tab = pool + fail_off[blockIdx.x];
    for (long long j = threadIdx.x; j < slots; j += WG) tab[j] = kEmpty;
    if (threadIdx.x == 0) s_n = 0;
    __threadfence();
    __syncthreads();
    const int sub = threadIdx.x / G;
    const int g = threadIdx.x % G;
    const uint32_t mask = static_cast<uint32_t>(slots - 1);
    int mine = 0;
    const int a1 = P.a_rpt[it.y + 1];
    for (int i = P.a_rpt[it.y] + sub; i < a1; i += E) {
        const int k = P.a_col[i];
        const int q1 = P.b_rpt[k + 1];
        for (int q = P.b_rpt[k] + g; q < q1; q += G) {
            const int key = P.b_col[q];
            uint32_t h = hash_slot_rt(key, bits);
            while (true) {
                const int cur = __hip_atomic_load(tab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == key) break;
                if (cur == kEmpty) {
                    const int old = atomicCAS(tab + h, kEmpty, key);
                    if (old == kEmpty) {
                        ++mine;
                        break;
                    }
                    if (old == key) break;
                }
                h = (h + 1) & mask;
            }
        }
    }
    if (mine) atomicAdd(&s_n, mine);
    __syncthreads();
    if (threadIdx.x == 0) P.c_rpt[it.y] = s_n;
}

// ----------------------------------------------------------------- launchers

template <typename V>
void launch_sym_multi(const ProbSet<V>& ps, const int2* items, const BinTable& sym, int n_multi,
                      const DevMeta* meta, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_sym_multi<V>, dim3(grid), dim3(kMultiThreads), 0, s, ps, items, sym, n_multi, meta);
}

template <typename V>
void launch_sym_max(const ProbSet<V>& ps, const int2* items, const BinTable& sym, int bin, DevMeta* meta,
                    int2* fail_items, long long* fail_off, int grid, hipStream_t s) {
    (void)sym;
    hipLaunchKernelGGL(k_sym_max<V>, dim3(grid), dim3(1024), 0, s, ps, items, bin, meta, fail_items, fail_off);
}

template <typename V>
void launch_sym_global(const ProbSet<V>& ps, const int2* fail_items, const long long* fail_off, int* pool,
                       int count, hipStream_t s) {
    hipLaunchKernelGGL(k_sym_global<V>, dim3(count), dim3(1024), 0, s, ps, fail_items, fail_off, pool);
}

template <typename V>
int occupancy_sym_multi() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sym_multi<V>, kMultiThreads, 0) != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}

template <typename V>
int occupancy_sym_max() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sym_max<V>, 1024, 0) != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}

#define SPG_INSTANTIATE_SYM(V)                                                                           \
    template void launch_sym_multi<V>(const ProbSet<V>&, const int2*, const BinTable&, int, const DevMeta*, \
                                      int, hipStream_t);                                                  \
    template void launch_sym_max<V>(const ProbSet<V>&, const int2*, const BinTable&, int, DevMeta*, int2*,  \
                                    long long*, int, hipStream_t);                                        \
    template void launch_sym_global<V>(const ProbSet<V>&, const int2*, const long long*, int*, int,        \
                                       hipStream_t);                                                      \
    template int occupancy_sym_multi<V>();                                                                \
    template int occupancy_sym_max<V>();

SPG_INSTANTIATE_SYM(float)
SPG_INSTANTIATE_SYM(double)

}  // namespace spg
