// Symbolic count: exact nnz of every row of C from an open-addressing hash
// table of the row's B columns (sym1.cuh:28-306), written into c_rpt.
//
// A workgroup unit holds R rows; each row owns L = WG/R consecutive lanes
// that enumerate its products flat (spg_rows.hpp) and insert their columns
// into the row's LDS table.  Rows that overflow the largest table are
// recounted in global memory.
#include "spg_kernels.hpp"
#include "spg_rows.hpp"

namespace spg {

template <typename V, int WG, int R, int T, int EC, int PMAX>
__device__ __forceinline__ void sym_rows(const ProbSet<V>& ps, const int2* __restrict__ items, int first,
                                         int end, int* smem, int* s_tmp) {
    constexpr int L = WG / R;
    static_assert(WG % R == 0, "bad variant shape");
    static_assert(L <= kWave || R == 1, "a row spans several waves only when it owns the workgroup");
    using Meta = ChunkMeta<V, EC, false>;
    constexpr int ROW_INTS = sym_row_ints(SymVariant{WG, R, T, EC, PMAX});
    const int slot = threadIdx.x / L;
    const int lr = threadIdx.x % L;
    int* tab = smem + slot * ROW_INTS;
    Meta* cm = reinterpret_cast<Meta*>(tab + T);
    lds_fill<T, L>(tab, kEmpty, lr);
    gsync<L>();

    const int idx = first + slot;
    const bool active = idx < end;
    const int2 it = active ? items[idx] : make_int2(0, 0);
    const GProb<V> P = ps.get(it.x);
    int a0 = 0, a1 = 0;
    if (active) {
        a0 = P.a_rpt[it.y];
        a1 = P.a_rpt[it.y + 1];
    }
    int mine = 0;
    for_each_round<V, L, EC, PMAX, false>(
        P, a0, a1, lr, cm, s_tmp,
        [&](const int(&key)[PMAX], const V(&)[PMAX], const bool(&ok)[PMAX]) { mine += table_insert_all<T>(tab, key, ok); });
    int nnz;
    if constexpr (L <= kWave) {
        nnz = group_sum<L>(mine);
    } else {
        int total;
        (void)group_excl_scan<L>(mine, lr, s_tmp, total);
        nnz = total;
    }
    if (active && lr == 0) P.c_rpt[it.y] = nnz;
    gsync<L>();  // the table is reused by the next unit
}

#define SPG_SYM_CASE(v)                                                                                   \
    case v:                                                                                               \
        if constexpr (v >= LO && v < HI)                                                                  \
            sym_rows<V, kMultiThreads, kSymVariants[v].r, kSymVariants[v].t, kSymVariants[v].ec,          \
                     kSymVariants[v].pmax>(ps, bin_items, first, end, smem, s_tmp);                       \
        break;

// LDS of a multi-bin launch over variants [LO, HI).
template <int LO, int HI>
constexpr int sym_multi_lds() {
    int m = 0;
    for (int i = LO; i < HI; ++i) {
        const int b = 4 * kSymVariants[i].r * sym_row_ints(kSymVariants[i]);
        m = m > b ? m : b;
    }
    return m;
}

// One launch serves the bins among [0, n_multi) of the symbolic table whose
// variant lies in [LO, HI).  Bin sizes are read on the device, so the
// launch needs no host round trip; the grid is persistent and walks the
// units of its bins.  The small-table variants get a launch of their own:
// a kernel takes the register budget of its largest variant, and theirs is
// far smaller, so twice as many of their waves fit on a CU.
template <typename V, int LO, int HI>
__global__ __launch_bounds__(kMultiThreads) SPG_WAVES_PER_EU(kSymMultiMinWaves[sym_part(LO)]) void k_sym_multi(ProbSet<V> ps, const int2* __restrict__ items,
                                                             int bin_cap, BinTable sym, int n_multi,
                                                             const DevMeta* __restrict__ meta) {
    static_assert(sym_multi_lds<LO, HI>() <= kSymMultiLds, "symbolic multi-bin LDS budget");
    __shared__ __align__(16) int smem[sym_multi_lds<LO, HI>() / 4];
    __shared__ int s_tmp[kMultiThreads / kWave];
    __shared__ int s_unit[kMaxBins + 1];
    __shared__ int s_cnt[kMaxBins];
    if (threadIdx.x == 0) {
        int units = 0;
        for (int b = 0; b < n_multi; ++b) {
            const int c = meta->sym_count[b].v;
            const bool mine = sym.variant[b] >= LO && sym.variant[b] < HI;
            s_cnt[b] = c;
            s_unit[b] = units;
            if (mine) units += (c + sym.rows[b] - 1) / sym.rows[b];
        }
        s_unit[n_multi] = units;
    }
    __syncthreads();
    const UnitRange ur = xcd_units(s_unit[n_multi]);
    for (int u = ur.first; u < ur.end; u += ur.stride) {
        int b = 0;
        while (u >= s_unit[b + 1]) ++b;
        const int2* bin_items = items + static_cast<long long>(b) * bin_cap;  // bin b's list
        const int first = (u - s_unit[b]) * sym.rows[b];
        const int end = s_cnt[b];
        switch (sym.variant[b]) {
            SPG_SYM_CASE(0)
            SPG_SYM_CASE(1)
            SPG_SYM_CASE(2)
            SPG_SYM_CASE(3)
            SPG_SYM_CASE(4)
            SPG_SYM_CASE(5)
            SPG_SYM_CASE(6)
            SPG_SYM_CASE(7)
            SPG_SYM_CASE(8)
            default:
                break;
        }
        // Consecutive units may belong to bins with different LDS layouts,
        // and rows of L <= 64 lanes only synchronise inside their wave: no
        // wave may start a unit of another layout while another wave still
        // uses this one.  Units of the same variant give every wave the same
        // LDS slots, so between them the waves run on without a barrier.
        const int un = u + ur.stride;
        if (un < ur.end) {
            int bn = b;
            while (un >= s_unit[bn + 1]) ++bn;
            if (sym.variant[bn] != sym.variant[b]) __syncthreads();
        }
    }
}

// Largest LDS table (32768 slots, 128 KiB) with overflow detection
// (sym1.cuh:169-245): a row whose distinct count passes 80 % of the table
// stops early and is queued, with its global-table slot range, for
// k_sym_global.
template <typename V>
__global__ __launch_bounds__(1024) void k_sym_max(ProbSet<V> ps, const int2* __restrict__ items, int bin_cap,
                                                  int bin, DevMeta* __restrict__ meta,
                                                  int2* __restrict__ fail_items, long long* __restrict__ fail_off) {
    constexpr SymVariant kV = kSymVariants[kSymMaxVariant];
    constexpr int T = kV.t;
    constexpr int WG = 1024;
    constexpr int EC = kV.ec;
    constexpr int PMAX = kV.pmax;
    static_assert(T - kSymMaxThreshold > WG * PMAX + 1, "table must keep room for in-flight inserts");
    using Meta = ChunkMeta<V, EC, false>;
    __shared__ __align__(16) int tab[T];
    __shared__ Meta cm;
    __shared__ int s_tmp[WG / kWave];
    __shared__ int s_n;
    const int2* bin_items = items + static_cast<long long>(bin) * bin_cap;
    const int count = meta->sym_count[bin].v;
    for (int u = blockIdx.x; u < count; u += gridDim.x) {
        lds_fill<T, WG>(tab, kEmpty, threadIdx.x);
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const int2 it = bin_items[u];
        const GProb<V> P = ps.get(it.x);
        const int a0 = P.a_rpt[it.y];
        const int a1 = P.a_rpt[it.y + 1];
        bool over = false;
        for (int ebase = a0; ebase < a1 && !over; ebase += EC) {
            int start[EC / WG];
            const int total = load_chunk<V, WG, EC, false>(P, ebase, a1, threadIdx.x, &cm, s_tmp, start);
            int prior = 0;
            for (int base = 0; base < total && !over; base += WG * PMAX) {
                int key[PMAX];
                V prod[PMAX];
                bool ok[PMAX];
                fetch_products<V, WG, EC, PMAX, false>(P, &cm, total, base, threadIdx.x, start, prior, key, prod,
                                                       ok);
                const int mine = table_insert_all<T>(tab, key, ok);
                if (mine) atomicAdd(&s_n, mine);
                __syncthreads();
                over = s_n > kSymMaxThreshold;  // uniform: every lane reads after the barrier
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            const int n = s_n;
            if (n <= kSymMaxThreshold) {
                P.c_rpt[it.y] = n;
            } else {
                const int f = atomicAdd(&meta->fail_count.v, 1);
                fail_items[f] = it;
                fail_off[f] = static_cast<long long>(atomicAdd(
                    &meta->fail_pool.v, static_cast<unsigned long long>(global_table_slots(P.nip[it.y]))));
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
    constexpr int EC = 1024;
    constexpr int PMAX = 4;
    using Meta = ChunkMeta<V, EC, false>;
    __shared__ Meta cm;
    __shared__ int s_tmp[WG / kWave];
    const int2 it = fail_items[blockIdx.x];
    const GProb<V> P = ps.get(it.x);
    const long long slots = global_table_slots(P.nip[it.y]);
    const int bits = __builtin_ctzll(static_cast<unsigned long long>(slots));
    int* tab = pool + fail_off[blockIdx.x];
    for (long long j = threadIdx.x; j < slots; j += WG) tab[j] = kEmpty;
    __threadfence();
    __syncthreads();
    const uint32_t mask = static_cast<uint32_t>(slots - 1);
    int mine = 0;
    for_each_product<V, WG, EC, PMAX, false>(P, P.a_rpt[it.y], P.a_rpt[it.y + 1], threadIdx.x, &cm, s_tmp,
                                             [&](int key, V) {
                                                 uint32_t h = hash_slot_rt(key, bits);
                                                 while (true) {
                                                     const int cur = __hip_atomic_load(
                                                         tab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                                             });
    int total;
    (void)group_excl_scan<WG>(mine, threadIdx.x, s_tmp, total);
    if (threadIdx.x == 0) P.c_rpt[it.y] = total;
}

// ----------------------------------------------------------------- launchers

template <typename V>
void launch_sym_multi(int part, const ProbSet<V>& ps, const int2* items, int bin_cap, const BinTable& sym,
                      int n_multi, const DevMeta* meta, int grid, hipStream_t s) {
    if (part == 0)
        hipLaunchKernelGGL((k_sym_multi<V, kSymPartLo[0], kSymPartHi[0]>), dim3(grid), dim3(kMultiThreads), 0, s, ps,
                           items, bin_cap, sym, n_multi, meta);
    else if (part == 1)
        hipLaunchKernelGGL((k_sym_multi<V, kSymPartLo[1], kSymPartHi[1]>), dim3(grid), dim3(kMultiThreads), 0, s, ps,
                           items, bin_cap, sym, n_multi, meta);
    else
        hipLaunchKernelGGL((k_sym_multi<V, kSymPartLo[2], kSymPartHi[2]>), dim3(grid), dim3(kMultiThreads), 0, s, ps,
                           items, bin_cap, sym, n_multi, meta);
}

template <typename V>
void launch_sym_max(const ProbSet<V>& ps, const int2* items, int bin_cap, const BinTable& sym, int bin,
                    DevMeta* meta, int2* fail_items, long long* fail_off, int grid, hipStream_t s) {
    (void)sym;
    hipLaunchKernelGGL(k_sym_max<V>, dim3(grid), dim3(1024), 0, s, ps, items, bin_cap, bin, meta, fail_items,
                       fail_off);
}

template <typename V>
void launch_sym_global(const ProbSet<V>& ps, const int2* fail_items, const long long* fail_off, int* pool,
                       int count, hipStream_t s) {
    hipLaunchKernelGGL(k_sym_global<V>, dim3(count), dim3(1024), 0, s, ps, fail_items, fail_off, pool);
}

template <typename V>
int occupancy_sym_multi(int part) {
    int n = 0;
    const hipError_t e =
        part == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &n, k_sym_multi<V, kSymPartLo[0], kSymPartHi[0]>, kMultiThreads, 0)
        : part == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &n, k_sym_multi<V, kSymPartLo[1], kSymPartHi[1]>, kMultiThreads, 0)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &n, k_sym_multi<V, kSymPartLo[2], kSymPartHi[2]>, kMultiThreads, 0);
    if (e != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}

template <typename V>
int occupancy_sym_max() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sym_max<V>, 1024, 0) != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}

#define SPG_INSTANTIATE_SYM(V)                                                                           \
    template void launch_sym_multi<V>(int, const ProbSet<V>&, const int2*, int, const BinTable&, int,      \
                                      const DevMeta*, int, hipStream_t);                                  \
    template void launch_sym_max<V>(const ProbSet<V>&, const int2*, int, const BinTable&, int, DevMeta*,   \
                                    int2*, long long*, int, hipStream_t);                                 \
    template void launch_sym_global<V>(const ProbSet<V>&, const int2*, const long long*, int*, int,        \
                                       hipStream_t);                                                      \
    template int occupancy_sym_multi<V>(int);                                                             \
    template int occupancy_sym_max<V>();

SPG_INSTANTIATE_SYM(float)
SPG_INSTANTIATE_SYM(double)

}  // namespace spg
