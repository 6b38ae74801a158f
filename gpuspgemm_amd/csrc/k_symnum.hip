// Fused second symbolic pass and numeric pass.  For every row:
//   1. insert the row's B columns into an LDS hash table of 2P slots
//      (sym2.cuh:27-121 without the count);
//   2. compact the occupied slots and sort them ascending with a bitonic
//      network in LDS (the reference sorts the whole table with CUB merge
//      sort, sym2.cuh:104-113); write them to C.col (sym2.cuh:116-120);
//   3. reuse the table's LDS as the row's value accumulator, zeroed
//      (num.cuh:97-100), and for every product a_ik * b_kj find the slot of
//      column j by binary search in the sorted LDS column list (find_key,
//      num.cuh:28-41) and add the product there (num.cuh:104-116);
//   4. write the values to C.val.
// Fusing the two passes keeps the sorted column list in LDS between them,
// so C.col is written once and never read back.  A row whose products fit
// in one chunk and in the lanes' PMAX register slots keeps (column, a*b)
// in registers from step 1 to step 3 and reads A and B only once.
#include "spg_kernels.hpp"
#include "spg_rows.hpp"

namespace spg {

// Compact the occupied slots of a row's table into cols[0..min(n, P))
// (unsorted); returns the number n of occupied slots.  Every lane reads its
// T/L slots first, then one group scan of the per-lane counts places them.
template <int L, int T, int P>
__device__ __forceinline__ int compact_table(const int* tab, int* cols, int lr, int* s_tmp) {
    static_assert(T % L == 0, "table must be a multiple of the group");
    constexpr int S = T / L;
    constexpr int B = L > kWave && S > 8 ? 8 : S;  // slots read per batch (workgroup-wide rows: bounded registers)
    int n = 0;
    for (int b0 = 0; b0 < S; b0 += B) {
        int key[B];
        int mine = 0;
#pragma unroll
        for (int s = 0; s < B; ++s) {
            key[s] = tab[lr + (b0 + s) * L];
            mine += key[s] != kEmpty;
        }
        int tot;
        int pos = n + group_excl_scan<L>(mine, lr, s_tmp, tot);
#pragma unroll
        for (int s = 0; s < B; ++s) {
            if (key[s] == kEmpty) continue;
            if (pos < P) cols[pos] = key[s];
            ++pos;
        }
        n += tot;
    }
    return n;
}

// Sorted column list of a row: the occupied slots of its table in
// ascending order in cols[0..n), padded with INT_MAX up to P; returns n.
// Groups inside one wave sort in registers (bitonic_sort_regs) and also
// leave lane lr's share cols[lr*E .. lr*E + E) in x; workgroup-wide rows
// sort wave blocks in registers and merge them in LDS (bitonic_sort_block).
template <int L, int T, int P>
__device__ __forceinline__ int sort_row(const int* tab, int* cols, int lr, int* s_tmp, int (&x)[L <= kWave ? P / L : 1]) {
    const int found = compact_table<L, T, P>(tab, cols, lr, s_tmp);
    const int n = min(found, P);
    if constexpr (L <= kWave) {
        constexpr int E = P / L;
        gsync<L>();
#pragma unroll
        for (int s = 0; s < E; ++s) x[s] = lr * E + s < n ? cols[lr * E + s] : INT_MAX;
        bitonic_sort_regs<P, L>(x, lr);
        gsync<L>();
#pragma unroll
        for (int s = 0; s < E; ++s) cols[lr * E + s] = x[s];
    } else {
        for (int j = n + lr; j < P; j += L) cols[j] = INT_MAX;
        gsync<L>();
        bitonic_sort_block<P, L>(cols, lr);
    }
    return found;
}

template <typename V, int WG, int R, int P, int EC, int PMAX>
__device__ __forceinline__ void symnum_rows(const ProbSet<V>& ps, const RowItem* __restrict__ items, int first,
                                            int end, int* smem, int* s_tmp, DevMeta* meta) {
    constexpr int L = WG / R;
    constexpr int T = 2 * P;
    constexpr int ROW_INTS = num_row_ints(NumVariant{WG, R, P, EC, PMAX});
    static_assert(WG % R == 0, "bad variant shape");
    static_assert(L <= kWave || R == 1, "a row spans several waves only when it owns the workgroup");
    static_assert(T * sizeof(int) >= P * sizeof(V), "values must fit in the table's LDS");
    static_assert(ROW_INTS % 2 == 0, "row slots must stay 8-byte aligned");
    using Meta = ChunkMeta<V, EC, true>;
    using KeyMeta = ChunkMeta<V, EC, false>;
    const int slot = threadIdx.x / L;
    const int lr = threadIdx.x % L;
    int* tab = smem + slot * ROW_INTS;
    int* cols = tab + T;
    Meta* cm = reinterpret_cast<Meta*>(cols + P);
    V* vals = reinterpret_cast<V*>(tab);

    lds_fill<T, L>(tab, kEmpty, lr);
    gsync<L>();

    const int idx = first + slot;
    const bool active = idx < end;
    const RowItem it = active ? load_item(items, idx) : RowItem{0, 0, 0, 0, 0, 0, 0, 0};
    const GProb<V> Pr = ps.get(it.inst);
    const int a0 = it.a0, a1 = it.a1, nip = it.nip, c0 = it.c0, cn = it.cn;
    // products staged in registers when the whole row fits in one round
    const bool fits = a1 - a0 <= EC && nip <= L * PMAX;
    int key[PMAX];
    V prod[PMAX];
    bool ok[PMAX];
    if (fits) {
        int total;
        if constexpr (L == 1) {
            Meta rm;
            total = load_chunk_lane<V, EC, true>(Pr, a0, a1, rm);
            fetch_products_lane<V, EC, PMAX, true>(Pr, rm, total, 0, key, prod, ok);
        } else {
            int start[EC / L];
            int prior = 0;
            total = load_chunk<V, L, EC, true>(Pr, a0, a1, lr, cm, s_tmp, start);
            fetch_products<V, L, EC, PMAX, true>(Pr, cm, total, 0, lr, start, prior, key, prod, ok);
        }
        table_insert_all<T>(tab, key, ok);
    } else {
        for_each_round<V, L, EC, PMAX, false>(
            Pr, a0, a1, lr, reinterpret_cast<KeyMeta*>(cm), s_tmp,
            [&](const int(&k)[PMAX], const V(&)[PMAX], const bool(&o)[PMAX]) { table_insert_all<T>(tab, k, o); });
    }
    gsync<L>();

    constexpr int E = L <= kWave ? P / L : 1;
    int x[E];
    const int found = sort_row<L, T, P>(tab, cols, lr, s_tmp, x);
    const int n = min(found, P);

    int w = 0;
    if (active) {
        if (lr == 0 && cn != found) atomicOr(&meta->flags.v, kFlagMismatch);
        w = min(n, cn);
    }
    if constexpr (L <= kWave) {
#pragma unroll
        for (int s = 0; s < E; ++s)
            if (lr * E + s < w) Pr.c_col[c0 + lr * E + s] = x[s];
    } else {
        for (int j = lr; j < w; j += L) Pr.c_col[c0 + j] = cols[j];
    }
    for (int j = lr; j < P; j += L) vals[j] = V(0);
    gsync<L>();

    // binary search of every product's column in the sorted list
    // (find_key, num.cuh:28-41), all of a lane's products in lockstep, then
    // LDS atomic adds that nothing waits for
    // (in groups of up to 8 keys, which bounds the registers in flight)
    auto accumulate = [&](const int(&k)[PMAX], const V(&v)[PMAX], const bool(&o)[PMAX]) {
        constexpr int G = PMAX % 8 == 0 ? 8 : (PMAX % 6 == 0 ? 6 : (PMAX % 4 == 0 ? 4 : PMAX));
        static_for<0, PMAX / G>([&](auto gi) {
            constexpr int J0 = decltype(gi)::value * G;
            int pos[G];
            lower_bound_all<P, G, J0>(cols, k, pos);
#pragma unroll
            for (int j = 0; j < G; ++j)
                if (o[J0 + j]) lds_add(vals + pos[j], v[J0 + j]);
        });
    };
    if (fits)
        accumulate(key, prod, ok);
    else
        for_each_round<V, L, EC, PMAX, true>(Pr, a0, a1, lr, cm, s_tmp, accumulate);
    gsync<L>();
    for (int j = lr; j < w; j += L) Pr.c_val[c0 + j] = vals[j];
    gsync<L>();  // LDS is reused by the next unit
}

// One launch per fused variant over the units (R rows each) of one bin; the
// host launches a workgroup per unit (residency is the hardware's), and the
// unit of a workgroup is chosen so that consecutive units stay on one XCD.
// A smaller grid walks the units persistently.
template <typename V, int VV>
__global__ __launch_bounds__(kNumVariants[VV].wg) SPG_WAVES_PER_EU(kNumMinWaves[VV]) void k_symnum_bin(ProbSet<V> ps, const RowItem* __restrict__ items,
                                                                    int item_off, int count, DevMeta* meta) {
    constexpr NumVariant kV = kNumVariants[VV];
    __shared__ __align__(16) int smem[kV.r * num_row_ints(kV)];
    __shared__ int s_tmp[kV.wg / kWave];
    const int units = (count + kV.r - 1) / kV.r;
    const UnitRange ur = xcd_units(units);
    for (int u = ur.first; u < ur.end; u += ur.stride)
        symnum_rows<V, kV.wg, kV.r, kV.p, kV.ec, kV.pmax>(ps, items, item_off + u * kV.r, item_off + count, smem,
                                                          s_tmp, meta);
}

#define SPG_NUM_VARIANTS(X) \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)

// Several bins of the fused pass in one launch, a workgroup per unit.
// Each XCD gets a contiguous slice of every bin (balanced work across the
// XCDs, and neighbouring rows, which read the same rows of B, meet in one
// L2).  Registers are the largest member's (kept within the set's
// residency floor), LDS the largest member's; a workgroup runs one unit, so
// no LDS layout is reused by another variant.
template <typename V, int SET>
__global__ __launch_bounds__(256) SPG_WAVES_PER_EU(kNumMultiMinWaves[SET]) void k_symnum_multi(
    ProbSet<V> ps, const RowItem* __restrict__ items, NumMultiBins mb, DevMeta* meta) {
    __shared__ __align__(16) int smem[max_num_multi_lds(SET) / 4];
    __shared__ int s_tmp[256 / kWave];
    const int xcd = blockIdx.x % kXcds;
    const int local = blockIdx.x / kXcds;
    int b = 0;
    while (b + 1 < mb.n && local >= mb.slice0[b + 1]) ++b;
    const int per = mb.slice0[b + 1] - mb.slice0[b];
    const int k = xcd * per + (local - mb.slice0[b]);  // unit of bin b
    const int end = mb.off[b] + mb.count[b];
#define SPG_MULTI_CASE(v)                                                                                   \
    case v:                                                                                                 \
        if constexpr (num_multi_set(v) == SET) {                                                            \
            static_assert(kNumVariants[v].wg == 256, "merged launches run 256-thread variants");           \
            if (mb.off[b] + k * kNumVariants[v].r < end)                                                    \
                symnum_rows<V, 256, kNumVariants[v].r, kNumVariants[v].p, kNumVariants[v].ec,               \
                            kNumVariants[v].pmax>(ps, items, mb.off[b] + k * kNumVariants[v].r, end, smem, s_tmp, \
                                                  meta);                                                    \
        }                                                                                                   \
        break;
    switch (mb.variant[b]) {
        SPG_NUM_VARIANTS(SPG_MULTI_CASE)
        default:
            break;
    }
#undef SPG_MULTI_CASE
}

// Lower bound in a sorted global list (values read coherently).
template <typename Ptr>
__device__ __forceinline__ int lower_bound_global(Ptr cols, int n, int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__hip_atomic_load(cols + mid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Rows above every LDS capacity: hash table in global memory, compaction
// into C.col, a sort of C.col in LDS-sized chunks followed by a merge by
// rank (each key's position = its rank in its chunk + the number of smaller
// keys in every other chunk), then binary search + global atomic adds into
// C.val (num.cuh:251-298).  The reference's O(nnz^2) rank sort
// (sym2.cuh:279-287) becomes O(nnz * chunks * log chunk).
template <typename V>
__global__ __launch_bounds__(1024) void k_symnum_global(ProbSet<V> ps, const RowItem* __restrict__ items,
                                                        const long long* __restrict__ gpool_off, int item_off,
                                                        int* pool, DevMeta* meta) {
    constexpr int WG = 1024;
    constexpr int EC = 1024;
    constexpr int PMAX = 4;
    constexpr int CH = kGlobalSortChunk;
    using Meta = ChunkMeta<V, EC, true>;
    __shared__ __align__(16) int chunk[CH];
    __shared__ Meta cm;
    __shared__ int s_tmp[WG / kWave];
    __shared__ int s_n;
    const RowItem it = load_item(items, item_off + blockIdx.x);
    const GProb<V> P = ps.get(it.inst);
    const int c0 = it.c0;
    const int nnz = it.cn;
    const long long slots = global_table_slots(nnz);
    const int bits = __builtin_ctzll(static_cast<unsigned long long>(slots));
    const uint32_t mask = static_cast<uint32_t>(slots - 1);
    int* tab = pool + gpool_off[item_off + blockIdx.x];
    auto ccol = P.c_col + c0;
    auto cval = P.c_val + c0;
    for (long long j = threadIdx.x; j < slots; j += WG) tab[j] = kEmpty;
    for (int j = threadIdx.x; j < nnz; j += WG) cval[j] = V(0);
    if (threadIdx.x == 0) s_n = 0;
    __threadfence();
    __syncthreads();

    const int a0 = it.a0;
    const int a1 = it.a1;
    for_each_product<V, WG, EC, PMAX, false>(
        P, a0, a1, threadIdx.x, reinterpret_cast<ChunkMeta<V, EC, false>*>(&cm), s_tmp, [&](int key, V) {
            uint32_t h = hash_slot_rt(key, bits);
            while (true) {
                const int cur = __hip_atomic_load(tab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (cur == key) break;
                if (cur == kEmpty) {
                    const int old = atomicCAS(tab + h, kEmpty, key);
                    if (old == kEmpty || old == key) break;
                }
                h = (h + 1) & mask;
            }
        });
    __threadfence();
    __syncthreads();

    // compact the occupied slots into C.col (unsorted)
    for (long long j = threadIdx.x; j < slots; j += WG) {
        const int key = __hip_atomic_load(tab + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (key != kEmpty) {
            const int pos = atomicAdd(&s_n, 1);
            if (pos < nnz) ccol[pos] = key;
        }
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0 && s_n != nnz) atomicOr(&meta->flags.v, kFlagMismatch);
    const int n = min(s_n, nnz);

    // sort chunks in LDS, staging them in the (now free) table
    const int nch = (n + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
        for (int j = threadIdx.x; j < CH; j += WG) {
            const int e = c * CH + j;
            chunk[j] = e < n ? __hip_atomic_load(ccol + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : INT_MAX;
        }
        __syncthreads();
        bitonic_sort_block<CH, WG>(chunk, threadIdx.x);
        for (int j = threadIdx.x; j < CH; j += WG) tab[static_cast<long long>(c) * CH + j] = chunk[j];
        __syncthreads();
    }
    __threadfence();
    __syncthreads();
    // merge by rank
    for (int e = threadIdx.x; e < n; e += WG) {
        const int c = e / CH;
        const int x = __hip_atomic_load(tab + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int rank = e - c * CH;
        for (int o = 0; o < nch; ++o) {
            if (o == c) continue;
            const int len = min(CH, n - o * CH);
            rank += lower_bound_global(tab + static_cast<long long>(o) * CH, len, x);
        }
        ccol[rank] = x;
    }
    __threadfence();
    __syncthreads();

    for_each_product<V, WG, EC, PMAX, true>(P, a0, a1, threadIdx.x, &cm, s_tmp, [&](int key, V v) {
        const int pos = lower_bound_global(ccol, n, key);
        if (pos < n) __hip_atomic_fetch_add(cval + pos, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    });
}

// ----------------------------------------------------------------- launchers

template <typename V, int VV>
static hipError_t launch_variant(const ProbSet<V>& ps, const RowItem* items, int item_off, int count, DevMeta* meta,
                                 int grid, hipStream_t s) {
    constexpr auto kernel = k_symnum_bin<V, VV>;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kNumVariants[VV].wg), 0, s, ps, items, item_off, count, meta);
    return hipGetLastError();
}

template <typename V, int VV>
static int occupancy_variant() {
    constexpr auto kernel = k_symnum_bin<V, VV>;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kNumVariants[VV].wg, 0) != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}


template <typename V>
hipError_t launch_symnum_bin(int variant, const ProbSet<V>& ps, const RowItem* items, int item_off, int count,
                             DevMeta* meta, int grid, hipStream_t s) {
#define SPG_LAUNCH_CASE(v) \
    case v:                \
        return launch_variant<V, v>(ps, items, item_off, count, meta, grid, s);
    switch (variant) {
        SPG_NUM_VARIANTS(SPG_LAUNCH_CASE)
        default:
            return hipErrorInvalidValue;
    }
#undef SPG_LAUNCH_CASE
}

template <typename V>
hipError_t launch_symnum_multi(int set, const ProbSet<V>& ps, const RowItem* items, const NumMultiBins& mb,
                               DevMeta* meta, hipStream_t s) {
    const int grid = kXcds * mb.slice0[mb.n];  // one unit per workgroup
    if (set == 0)
        hipLaunchKernelGGL((k_symnum_multi<V, 0>), dim3(grid), dim3(256), 0, s, ps, items, mb, meta);
    else if (set == 1)
        hipLaunchKernelGGL((k_symnum_multi<V, 1>), dim3(grid), dim3(256), 0, s, ps, items, mb, meta);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename V>
void launch_symnum_global(const ProbSet<V>& ps, const RowItem* items, const long long* gpool_off, int item_off,
                          int count, int* pool, DevMeta* meta, hipStream_t s) {
    hipLaunchKernelGGL(k_symnum_global<V>, dim3(count), dim3(1024), 0, s, ps, items, gpool_off, item_off, pool,
                       meta);
}

template <typename V>
int occupancy_symnum_bin(int variant) {
#define SPG_OCC_CASE(v) \
    case v:             \
        return occupancy_variant<V, v>();
    switch (variant) {
        SPG_NUM_VARIANTS(SPG_OCC_CASE)
        default:
            return 1;
    }
#undef SPG_OCC_CASE
}

#define SPG_INSTANTIATE_NUM(V)                                                                               \
    template hipError_t launch_symnum_bin<V>(int, const ProbSet<V>&, const RowItem*, int, int, DevMeta*, int,    \
                                             hipStream_t);                                                    \
    template hipError_t launch_symnum_multi<V>(int, const ProbSet<V>&, const RowItem*, const NumMultiBins&,       \
                                               DevMeta*, hipStream_t);                                        \
    template void launch_symnum_global<V>(const ProbSet<V>&, const RowItem*, const long long*, int, int, int*, \
                                          DevMeta*, hipStream_t);                                             \
    template int occupancy_symnum_bin<V>(int);

SPG_INSTANTIATE_NUM(float)
SPG_INSTANTIATE_NUM(double)

}  // namespace spg
