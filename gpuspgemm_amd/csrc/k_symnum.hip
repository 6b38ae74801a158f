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
// so C.col is written once and never read back.
#include "spg_kernels.hpp"

namespace spg {

// Ascending bitonic sort of P (a power of two) ints in LDS by L lanes.
template <int P, int L>
__device__ __forceinline__ void bitonic_sort(int* cols, int lr) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lr; t < P / 2; t += L) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int a = cols[i];
                const int b = cols[i + j];
                const bool asc = (i & k) == 0;
                if ((a > b) == asc) {
                    cols[i] = b;
                    cols[i + j] = a;
                }
            }
            __syncthreads();
        }
    }
}

template <typename V, int WG, int R, int G, int P>
__device__ __forceinline__ void symnum_rows(const ProbSet<V>& ps, const int2* __restrict__ items, int first,
                                            int end, int* smem, int* s_cnt, DevMeta* meta) {
    constexpr int L = WG / R;
    constexpr int E = L / G;
    constexpr int T = 2 * P;
    static_assert(WG % R == 0 && L % G == 0, "bad variant shape");
    static_assert(T * sizeof(int) >= P * sizeof(V), "values must fit in the table's LDS");
    const int slot = threadIdx.x / L;
    const int lr = threadIdx.x % L;
    const int sub = lr / G;
    const int g = lr % G;
    int* tab = smem + slot * T;
    int* cols = smem + R * T + slot * P;
    V* vals = reinterpret_cast<V*>(tab);

    for (int j = lr; j < T; j += L) tab[j] = kEmpty;
    if (lr == 0) s_cnt[slot] = 0;
    __syncthreads();

    const int idx = first + slot;
    const bool active = idx < end;
    const int2 it = active ? items[idx] : make_int2(0, 0);
    const Prob<V> Pr = ps.get(it.x);
    int a0 = 0, a1 = 0;
    if (active) {
        a0 = Pr.a_rpt[it.y];
        a1 = Pr.a_rpt[it.y + 1];
    }
    for (int i = a0 + sub; i < a1; i += E) {
        const int k = Pr.a_col[i];
        const int q1 = Pr.b_rpt[k + 1];
        for (int q = Pr.b_rpt[k] + g; q < q1; q += G) table_insert<T>(tab, Pr.b_col[q]);
    }
    __syncthreads();

    for (int j = lr; j < T; j += L) {
        const int key = tab[j];
        if (key != kEmpty) {
            const int pos = atomicAdd(&s_cnt[slot], 1);
            if (pos < P) cols[pos] = key;
        }
    }
    __syncthreads();
    const int n = min(s_cnt[slot], P);
    for (int j = n + lr; j < P; j += L) cols[j] = INT_MAX;
    __syncthreads();
    bitonic_sort<P, L>(cols, lr);

    int c0 = 0, w = 0;
    if (active) {
        c0 = Pr.c_rpt[it.y];
        const int cn = Pr.c_rpt[it.y + 1] - c0;
        if (lr == 0 && cn != s_cnt[slot]) atomicOr(&meta->flags, kFlagMismatch);
        w = min(n, cn);
    }
    for (int j = lr; j < w; j += L) Pr.c_col[c0 + j] = cols[j];
    for (int j = lr; j < P; j += L) vals[j] = V(0);
    __syncthreads();

    for (int i = a0 + sub; i < a1; i += E) {
        const int k = Pr.a_col[i];
        const V av = Pr.a_val[i];
        const int q1 = Pr.b_rpt[k + 1];
        for (int q = Pr.b_rpt[k] + g; q < q1; q += G) {
            const int pos = lower_bound_pow2<P>(cols, Pr.b_col[q]);
            const V prod = av * Pr.b_val[q];
            if constexpr (L == 1)
                vals[pos] += prod;
            else
                lds_add(vals + pos, prod);
        }
    }
    __syncthreads();
    for (int j = lr; j < w; j += L) Pr.c_val[c0 + j] = vals[j];
    __syncthreads();  // LDS is reused by the next unit
}

#define SPG_NUM_CASE(v)                                                                                   \
    case v:                                                                                               \
        symnum_rows<V, kMultiThreads, kNumVariants[v].r, kNumVariants[v].g, kNumVariants[v].p>(            \
            ps, items, first, end, smem, s_cnt, meta);                                                    \
        break;

template <typename V>
__global__ __launch_bounds__(kMultiThreads) void k_symnum_multi(ProbSet<V> ps, const int2* __restrict__ items,
                                                                MultiArgs ma, DevMeta* meta) {
    __shared__ __align__(16) int smem[kNumMultiLds / 4];
    __shared__ int s_cnt[kMultiThreads];
    const UnitRange ur = xcd_units(ma.unit_start[ma.n]);
    for (int u = ur.first; u < ur.end; u += ur.stride) {
        int b = 0;
        while (u >= ma.unit_start[b + 1]) ++b;
        const int first = ma.item_off[b] + (u - ma.unit_start[b]) * ma.rows[b];
        const int end = ma.item_off[b] + ma.count[b];
        switch (ma.variant[b]) {
            SPG_NUM_CASE(0)
            SPG_NUM_CASE(1)
            SPG_NUM_CASE(2)
            SPG_NUM_CASE(3)
            SPG_NUM_CASE(4)
            SPG_NUM_CASE(5)
            SPG_NUM_CASE(6)
            SPG_NUM_CASE(7)
            SPG_NUM_CASE(8)
            default:
                break;
        }
    }
}

// One 1024-lane workgroup per row, P = 8192 (96 KiB of LDS in FP64).
template <typename V>
__global__ __launch_bounds__(1024) void k_symnum_big(ProbSet<V> ps, const int2* __restrict__ items, int item_off,
                                                     int count, DevMeta* meta) {
    constexpr NumVariant kV = kNumVariants[kNumBigVariant];
    __shared__ __align__(16) int smem[kNumBigLds / 4];
    __shared__ int s_cnt[1];
    for (int u = blockIdx.x; u < count; u += gridDim.x)
        symnum_rows<V, 1024, kV.r, kV.g, kV.p>(ps, items, item_off + u, item_off + count, smem, s_cnt, meta);
}

// Lower bound in a sorted global list (values read coherently).
__device__ __forceinline__ int lower_bound_global(const int* cols, int n, int key) {
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
__global__ __launch_bounds__(1024) void k_symnum_global(ProbSet<V> ps, const int2* __restrict__ items,
                                                        const long long* __restrict__ gpool_off, int item_off,
                                                        int* pool, DevMeta* meta) {
    constexpr int WG = 1024;
    constexpr int G = 64;
    constexpr int E = WG / G;
    constexpr int CH = kGlobalSortChunk;
    __shared__ int chunk[CH];
    __shared__ int s_n;
    const int2 it = items[item_off + blockIdx.x];
    const Prob<V> P = ps.get(it.x);
    const int c0 = P.c_rpt[it.y];
    const int nnz = P.c_rpt[it.y + 1] - c0;
    const long long slots = global_table_slots(nnz);
    const int bits = __builtin_ctzll(static_cast<unsigned long long>(slots));
    const uint32_t mask = static_cast<uint32_t>(slots - 1);
    int* tab = pool + gpool_off[item_off + blockIdx.x];
    int* ccol = P.c_col + c0;
    V* cval = P.c_val + c0;
    for (long long j = threadIdx.x; j < slots; j += WG) tab[j] = kEmpty;
    for (int j = threadIdx.x; j < nnz; j += WG) cval[j] = V(0);
    if (threadIdx.x == 0) s_n = 0;
    __threadfence();
    __syncthreads();

    const int sub = threadIdx.x / G;
    const int g = threadIdx.x % G;
    const int a0 = P.a_rpt[it.y];
    const int a1 = P.a_rpt[it.y + 1];
    for (int i = a0 + sub; i < a1; i += E) {
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
                    if (old == kEmpty || old == key) break;
                }
                h = (h + 1) & mask;
            }
        }
    }
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
    if (threadIdx.x == 0 && s_n != nnz) atomicOr(&meta->flags, kFlagMismatch);
    const int n = min(s_n, nnz);

    // sort chunks in LDS, staging them in the (now free) table
    const int nch = (n + CH - 1) / CH;
    for (int c = 0; c < nch; ++c) {
        for (int j = threadIdx.x; j < CH; j += WG) {
            const int e = c * CH + j;
            chunk[j] = e < n ? __hip_atomic_load(ccol + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : INT_MAX;
        }
        __syncthreads();
        bitonic_sort<CH, WG>(chunk, threadIdx.x);
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

    for (int i = a0 + sub; i < a1; i += E) {
        const int k = P.a_col[i];
        const V av = P.a_val[i];
        const int q1 = P.b_rpt[k + 1];
        for (int q = P.b_rpt[k] + g; q < q1; q += G) {
            const int pos = lower_bound_global(ccol, n, P.b_col[q]);
            if (pos < n)
                __hip_atomic_fetch_add(cval + pos, av * P.b_val[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ----------------------------------------------------------------- launchers

template <typename V>
void launch_symnum_multi(const ProbSet<V>& ps, const int2* items, const MultiArgs& ma, DevMeta* meta, int grid,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_symnum_multi<V>, dim3(grid), dim3(kMultiThreads), 0, s, ps, items, ma, meta);
}

template <typename V>
void launch_symnum_big(const ProbSet<V>& ps, const int2* items, int item_off, int count, DevMeta* meta,
                       int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_symnum_big<V>, dim3(grid), dim3(1024), 0, s, ps, items, item_off, count, meta);
}

template <typename V>
void launch_symnum_global(const ProbSet<V>& ps, const int2* items, const long long* gpool_off, int item_off,
                          int count, int* pool, DevMeta* meta, hipStream_t s) {
    hipLaunchKernelGGL(k_symnum_global<V>, dim3(count), dim3(1024), 0, s, ps, items, gpool_off, item_off, pool,
                       meta);
}

template <typename V>
int occupancy_symnum_multi() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_symnum_multi<V>, kMultiThreads, 0) != hipSuccess)
        n = 1;
    return n > 0 ? n : 1;
}

template <typename V>
int occupancy_symnum_big() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_symnum_big<V>, 1024, 0) != hipSuccess) n = 1;
    return n > 0 ? n : 1;
}

#define SPG_INSTANTIATE_NUM(V)                                                                              \
    template void launch_symnum_multi<V>(const ProbSet<V>&, const int2*, const MultiArgs&, DevMeta*, int,    \
                                         hipStream_t);                                                       \
    template void launch_symnum_big<V>(const ProbSet<V>&, const int2*, int, int, DevMeta*, int, hipStream_t); \
    template void launch_symnum_global<V>(const ProbSet<V>&, const int2*, const long long*, int, int, int*,   \
                                          DevMeta*, hipStream_t);                                            \
    template int occupancy_symnum_multi<V>();                                                                \
    template int occupancy_symnum_big<V>();

SPG_INSTANTIATE_NUM(float)
SPG_INSTANTIATE_NUM(double)

}  // namespace spg
