// Row-parallel passes: counter reset, NIP + symbolic binning with the
// scatter into the bin lists, and the two-kernel exclusive scan that turns
// nnz per row into row pointers while binning rows for the fused pass.
//
// One workgroup handles one tile of kTileRows rows of one product; with a
// single product tiles are implicit, with a batch they come from a table.
#include "spg_kernels.hpp"

namespace spg {

__global__ __launch_bounds__(256) void k_clear(DevMeta* meta, int from_word) {
    int* w = reinterpret_cast<int*>(meta);
    constexpr int kWords = sizeof(DevMeta) / sizeof(int);
    for (int i = from_word + threadIdx.x; i < kWords; i += 256) w[i] = 0;
}

// NIP_i = sum over k in A(i,:) of nnz(B(k,:)) (setup.cuh:48-91): a lane
// per row, each wave walking the entries of its 64 rows together (below).
// Rows with NIP 0 get nnz 0 right away and join no bin
// (binning.cuh:60-66).  The other rows are scattered into the symbolic bin
// lists in the same pass (binning.cuh:76-120): bin b's list is the region
// items[b * bin_cap, (b + 1) * bin_cap) with bin_cap >= the total row
// count, so a tile's place in it is one atomic add on the bin's count and
// no offsets have to be known beforehand.
template <typename V>
__global__ __launch_bounds__(kRowThreads) void k_nip(ProbSet<V> ps, BinTable sym, DevMeta* meta,
                                                     int2* __restrict__ items, int bin_cap) {
    __shared__ int s_cnt[kMaxBins];
    __shared__ int s_base[kMaxBins];
    __shared__ int s_max;
    __shared__ unsigned long long s_total;
    __shared__ int s_flags;
    constexpr int kU = 8;  // entries per lane and round
    constexpr int kRound = kU * kWave;
    __shared__ long long s_pref[kRowThreads / kWave][kRound];  // each wave's prefix sums of a round
    if (threadIdx.x < kMaxBins) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_max = 0;
        s_total = 0;
        s_flags = 0;
    }
    __syncthreads();

    const Tile t = ps.tile(blockIdx.x);
    const GProb<V> P = ps.get(t.inst);
    // NIP of each lane's rows.  A wave takes 64 * kPer consecutive rows, kPer
    // consecutive ones per lane, so the wave's entries are one consecutive
    // range of A's columns.  The wave walks that range in rounds of kRound
    // entries side by side: coalesced loads of A's columns, a gather of B's
    // row pointers each (every load of the round in flight at once),
    // coalesced stores of the (start, length) of every entry's row of B (the
    // later passes read them).  The lengths then go through the wave's LDS
    // slice, kU consecutive ones to a lane, into a prefix sum of the round,
    // from which each lane takes its rows' shares.  Three dependent loads
    // cover up to kRound entries however they fall over the rows.  Rows past
    // the tile's end are empty rows at its last entry.
    constexpr int kPer = kTileRows / kRowThreads;
    long long* pref = s_pref[threadIdx.x / kWave];
    int* lens = reinterpret_cast<int*>(pref);  // the round's lengths, in its first half
    const int wrow = (threadIdx.x / kWave) * (kWave * kPer) + lane_id() * kPer;  // the lane's first row in the tile
    int bound[kPer + 1];  // entry bounds of the lane's rows
#pragma unroll
    for (int j = 0; j <= kPer; ++j) bound[j] = P.a_rpt[t.row0 + min(wrow + j, t.rows)];
    long long s[kPer] = {};
    const int g0 = __shfl(bound[0], 0), g1 = __shfl(bound[kPer], kWave - 1);
    for (int rb = g0; rb < g1; rb += kRound) {
        int kk[kU], b0[kU], len[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int e = rb + u * kWave + lane_id();
            kk[u] = e < g1 ? P.a_col[e] : -1;
        }
        // every gather is issued before the first store
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            b0[u] = 0;
            len[u] = 0;
            if (kk[u] >= 0) {
                b0[u] = P.b_rpt[kk[u]];
                len[u] = P.b_rpt[kk[u] + 1];
            }
        }
        wave_sync();  // the previous round's reads are done
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            len[u] -= b0[u];
            if (kk[u] >= 0) P.bref[rb + u * kWave + lane_id()] = pack_bref(b0[u], len[u]);
            lens[u * kWave + lane_id()] = len[u];
        }
        wave_sync();
        long long loc[kU];
        long long run = 0;
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            run += lens[lane_id() * kU + u];
            loc[u] = run;
        }
        // exclusive prefix of the lanes' totals
        long long x = run;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const long long y = __shfl_up(x, o);
            if (lane_id() >= o) x += y;
        }
        x -= run;
        wave_sync();  // every lane has read its lengths
#pragma unroll
        for (int u = 0; u < kU; ++u) pref[lane_id() * kU + u] = x + loc[u];
        wave_sync();
        // pref[i] = sum of the round's lengths 0..i; a row's share is the
        // difference at its clamped bounds
        long long at[kPer + 1];
#pragma unroll
        for (int j = 0; j <= kPer; ++j) {
            const int c = min(max(bound[j] - rb, 0), kRound);
            at[j] = c > 0 ? pref[c - 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) s[j] += at[j + 1] - at[j];
    }
    int my_max = 0;
    long long my_total = 0;
    int my_flags = 0;
    int bin[kPer], rank[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int r = wrow + j;
        const bool valid = r < t.rows;
        const int row = t.row0 + r;
        bin[j] = -1;
        if (valid) {
            if (s[j] > INT_MAX) {
                my_flags |= kFlagNipOverflow;
                s[j] = INT_MAX;
            }
            const int nip = static_cast<int>(s[j]);
            P.nip[row] = nip;
            if (nip == 0)
                P.c_rpt[row] = 0;
            else
                bin[j] = find_bin(sym, nip, 0);
            my_max = max(my_max, nip);
            my_total += s[j];
        }
        rank[j] = wave_rank_add(s_cnt, bin[j]);
    }
    my_max = wave_max(my_max);
    my_total = wave_sum64(my_total);
    if (lane_id() == 0) {
        atomicMax(&s_max, my_max);
        atomicAdd(&s_total, static_cast<unsigned long long>(my_total));
    }
    if (my_flags) atomicOr(&s_flags, my_flags);
    __syncthreads();
    if (threadIdx.x < sym.n && s_cnt[threadIdx.x] > 0)
        s_base[threadIdx.x] = atomicAdd(&meta->sym_count[threadIdx.x].v, s_cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
        atomicMax(&meta->max_nip.v, s_max);
        atomicAdd(&meta->total_nip.v, s_total);
        if (s_flags) atomicOr(&meta->flags.v, s_flags);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j)
        if (bin[j] >= 0)
            items[static_cast<long long>(bin[j]) * bin_cap + s_base[bin[j]] + rank[j]] =
                make_int2(t.inst, t.row0 + wrow + j);
}

// Scan part 1: per-tile sum of nnz, NUM-bin histogram, max nnz and the
// global-table slots the heaviest rows will need.
template <typename V>
__global__ __launch_bounds__(kRowThreads) void k_scan_count(ProbSet<V> ps, BinTable num, DevMeta* meta,
                                                            long long* __restrict__ tile_sums, int after_fail) {
    // Rows that overflowed the largest LDS table have no count yet: leave
    // c_rpt untouched; the host recounts them and runs the scan again.
    if (!after_fail && meta->fail_count.v > 0) return;
    __shared__ int s_cnt[kMaxBins];
    __shared__ long long s_sum;
    __shared__ unsigned long long s_pool;
    __shared__ int s_max;
    if (threadIdx.x < kMaxBins) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        s_sum = 0;
        s_pool = 0;
        s_max = 0;
    }
    __syncthreads();
    const Tile t = ps.tile(blockIdx.x);
    const GProb<V> P = ps.get(t.inst);
    long long sum = 0;
    unsigned long long pool = 0;
    int mx = 0;
    constexpr int kPer = kTileRows / kRowThreads;
    int nnz[kPer], nip[kPer];  // both loads of every row issued before any is used
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int r = threadIdx.x + j * kRowThreads;
        nnz[j] = r < t.rows ? P.c_rpt[t.row0 + r] : 0;
        nip[j] = r < t.rows ? P.nip[t.row0 + r] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        int b = -1;
        sum += nnz[j];
        mx = max(mx, nnz[j]);
        if (nnz[j] > 0) {
            b = find_bin(num, nip[j], nnz[j]);
            if (num.variant[b] == kNumGlobalVariant) pool += global_table_slots(nnz[j]);
        }
        wave_count_add(s_cnt, b);
    }
    sum = wave_sum64(sum);
    mx = wave_max(mx);
    long long p64 = wave_sum64(static_cast<long long>(pool));
    if (lane_id() == 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&s_sum), static_cast<unsigned long long>(sum));
        atomicAdd(&s_pool, static_cast<unsigned long long>(p64));
        atomicMax(&s_max, mx);
    }
    __syncthreads();
    if (threadIdx.x < num.n && s_cnt[threadIdx.x] > 0)
        atomicAdd(&meta->num_count[threadIdx.x].v, s_cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
        tile_sums[blockIdx.x] = s_sum;
        atomicMax(&meta->max_nnz.v, s_max);
        if (s_pool) atomicAdd(&meta->gnum_pool.v, s_pool);
    }
}

// Between the two scan parts when a multiply has many tiles: the tile sums,
// in place, become their exclusive prefix over all tiles (one workgroup,
// 1024 tiles a round), so that k_scan_apply reads a tile's prefix instead
// of summing every earlier tile of its product.
__global__ __launch_bounds__(kRowThreads) void k_tile_prefix(long long* __restrict__ tile_sums, int n) {
    constexpr int kWaves = kRowThreads / kWave;
    __shared__ long long s_wave[kWaves];
    long long carry = 0;
    for (int base = 0; base < n; base += kRowThreads) {
        const int i = base + threadIdx.x;
        const long long v = i < n ? tile_sums[i] : 0;
        long long incl = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const long long y = __shfl_up(incl, o);
            if (lane_id() >= o) incl += y;
        }
        if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
        __syncthreads();
        long long before = 0, round = 0;
        for (int w = 0; w < kWaves; ++w) {
            const long long x = s_wave[w];
            if (w < static_cast<int>(threadIdx.x / kWave)) before += x;
            round += x;
        }
        if (i < n) tile_sums[i] = carry + before + incl - v;
        carry += round;
        __syncthreads();  // s_wave is rewritten by the next round
    }
}

// Scan part 2: exclusive prefix of nnz into c_rpt (binning.cuh:219-224),
// c_rpt[m] = nnz(C), and the scatter of rows into the NUM bin lists.  The
// prefix of a tile is the sum of the earlier tiles of its product, which
// every workgroup reduces itself (no inter-workgroup waiting).
template <typename V>
__global__ __launch_bounds__(kRowThreads) void k_scan_apply(ProbSet<V> ps, BinTable num, DevMeta* meta,
                                                            const long long* __restrict__ tile_sums,
                                                            RowItem* __restrict__ items,
                                                            long long* __restrict__ gpool_off,
                                                            long long* __restrict__ inst_nnz, int after_fail,
                                                            int tiles_scanned) {
    if (!after_fail && meta->fail_count.v > 0) return;
    constexpr int kPer = kTileRows / kRowThreads;  // consecutive rows per thread
    constexpr int kWaves = kRowThreads / kWave;
    __shared__ long long s_wave[kWaves];
    __shared__ long long s_prefix;
    __shared__ int s_cnt[kMaxBins];
    __shared__ int s_base[kMaxBins];
    __shared__ int s_flags;

    const Tile t = ps.tile(blockIdx.x);
    const GProb<V> P = ps.get(t.inst);
    if (threadIdx.x < kMaxBins) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_flags = 0;

    // Every load that does not depend on another is issued first, so that
    // their latencies overlap: the rows' values, the bin sizes (for the
    // offset of each bin's list) and, below, the earlier tiles' sums.
    int v[kPer];
    int nip[kPer];
    int ar[kPer + 1];  // the rows' A row pointers, for the fused-pass items
    const int r0 = threadIdx.x * kPer;  // consecutive rows per thread
#pragma unroll
    for (int j = 0; j <= kPer; ++j) ar[j] = r0 + j <= t.rows ? P.a_rpt[t.row0 + r0 + j] : 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int r = r0 + j;
        v[j] = r < t.rows ? P.c_rpt[t.row0 + r] : 0;
        nip[j] = r < t.rows ? P.nip[t.row0 + r] : 0;
    }
    int bin_off = 0;  // thread b < num.n: rows of the bins before b
    if (threadIdx.x < num.n)
        for (int b = 0; b < static_cast<int>(threadIdx.x); ++b) bin_off += meta->num_count[b].v;

    // prefix over earlier tiles of this product: read off the scanned tile
    // sums (k_tile_prefix, many tiles), or summed here (few tiles; the
    // threads of a workgroup read them side by side)
    long long pre = 0;
    if (tiles_scanned) {
        if (threadIdx.x == 0) pre = tile_sums[blockIdx.x] - tile_sums[t.first_tile];
    } else {
        for (int i = t.first_tile + threadIdx.x; i < static_cast<int>(blockIdx.x); i += kRowThreads)
            pre += tile_sums[i];
    }
    pre = wave_sum64(pre);
    if (lane_id() == 0) s_wave[threadIdx.x / kWave] = pre;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long s = 0;
        for (int w = 0; w < kWaves; ++w) s += s_wave[w];
        s_prefix = s;
    }
    __syncthreads();
    const long long tile_base = s_prefix;

    // bins and ranks of the rows
    int bin[kPer];
    int rank[kPer];
    long long local = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (v[j] == 0) nip[j] = 0;
        local += v[j];
        bin[j] = v[j] > 0 ? find_bin(num, nip[j], v[j]) : -1;
        rank[j] = wave_rank_add(s_cnt, bin[j]);
    }
    // block exclusive scan of `local`
    long long incl = local;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const long long y = __shfl_up(incl, o);
        if (lane_id() >= o) incl += y;
    }
    __syncthreads();  // s_wave reuse
    if (lane_id() == kWave - 1) s_wave[threadIdx.x / kWave] = incl;
    __syncthreads();
    long long wave_base = 0;
    for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); ++w) wave_base += s_wave[w];
    long long run = tile_base + wave_base + incl - local;

    int flags = 0;
    int c0[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int r = r0 + j;
        c0[j] = static_cast<int>(run);
        if (r < t.rows) {
            if (run > INT_MAX) flags |= kFlagNnzOverflow;
            P.c_rpt[t.row0 + r] = static_cast<int>(run);
        }
        run += v[j];
    }
    // last row of the product: c_rpt[m] = nnz(C)
    if (t.rows > 0 && r0 < t.rows && r0 + kPer >= t.rows && t.row0 + t.rows == P.m) {
        if (run > INT_MAX) flags |= kFlagNnzOverflow;
        P.c_rpt[P.m] = static_cast<int>(run > INT_MAX ? INT_MAX : run);
        inst_nnz[t.inst] = run;
        if (ps.n_inst == 1) meta->nnz_total.v = run;
    }
    if (flags) atomicOr(&s_flags, flags);
    __syncthreads();

    if (threadIdx.x < num.n && s_cnt[threadIdx.x] > 0)
        s_base[threadIdx.x] = bin_off + atomicAdd(&meta->num_cursor[threadIdx.x].v, s_cnt[threadIdx.x]);
    if (threadIdx.x == 0 && s_flags) atomicOr(&meta->flags.v, s_flags);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (bin[j] < 0) continue;
        const int pos = s_base[bin[j]] + rank[j];
        items[pos] = RowItem{t.inst, t.row0 + r0 + j, ar[j], ar[j + 1], c0[j], v[j], nip[j], 0};
        if (num.variant[bin[j]] == kNumGlobalVariant)
            gpool_off[pos] = static_cast<long long>(
                atomicAdd(&meta->gnum_cursor.v, static_cast<unsigned long long>(global_table_slots(v[j]))));
    }
}

// Total NIP of A . B for the FLOP metric (nip.cuh:19-49): a tile of rows
// per workgroup, whose entries (one consecutive range of A's columns) the
// threads walk side by side, eight loads each in flight; one atomic per
// workgroup.
__global__ __launch_bounds__(kRowThreads) void k_total_nip(const int32_t* __restrict__ a_rpt,
                                                           const int32_t* __restrict__ a_col,
                                                           const int32_t* __restrict__ b_rpt, int m,
                                                           unsigned long long* out) {
    constexpr int kU = 8;
    __shared__ unsigned long long s_total;
    if (threadIdx.x == 0) s_total = 0;
    __syncthreads();
    const int r0 = blockIdx.x * kTileRows;
    const int e0 = a_rpt[r0], e1 = a_rpt[min(m, r0 + kTileRows)];
    long long s = 0;
    for (int base = e0 + threadIdx.x; base < e1; base += kU * kRowThreads) {
        int kk[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) kk[u] = base + u * kRowThreads < e1 ? a_col[base + u * kRowThreads] : -1;
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (kk[u] >= 0) s += b_rpt[kk[u] + 1] - b_rpt[kk[u]];
    }
    s = wave_sum64(s);
    if (lane_id() == 0 && s) atomicAdd(&s_total, static_cast<unsigned long long>(s));
    __syncthreads();
    if (threadIdx.x == 0 && s_total) atomicAdd(out, s_total);
}

// ----------------------------------------------------------------- launchers

void launch_total_nip(const int32_t* a_rpt, const int32_t* a_col, const int32_t* b_rpt, int m,
                      unsigned long long* out, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_total_nip, dim3((m + kTileRows - 1) / kTileRows), dim3(kRowThreads), 0, s, a_rpt,
                       a_col, b_rpt, m, out);
}

// Completion signal for the host: one lane stores 1 into a word of pinned
// host memory (a vector store at system scope) that the host cleared before
// queueing the work.  Queued behind a read-back on the same stream, it tells
// a polling host that the read-back has landed.  The value is a constant so
// that the launch can be replayed from a graph.
__global__ __launch_bounds__(64) void k_signal(int* host_flag) {
    if (threadIdx.x == 0) __hip_atomic_store(host_flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_signal(int* host_flag, hipStream_t s) {
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, s, host_flag);
}

// A read-back and its completion signal in one launch: copy `words` ints
// and `n2` long longs to pinned host memory (vector stores), optionally
// zero the device counters for the next multiply once they are copied,
// make the copies visible at system scope and signal.
__global__ __launch_bounds__(256) void k_readback(const int* __restrict__ src, int* host_dst, int words,
                                                  const long long* __restrict__ src2, long long* host_dst2, int n2,
                                                  DevMeta* clear, int* host_flag) {
    for (int i = threadIdx.x; i < words; i += 256) host_dst[i] = src[i];
    for (int i = threadIdx.x; i < n2; i += 256) host_dst2[i] = src2[i];
    if (clear) {
        __syncthreads();  // every copy has read its word
        int* w = reinterpret_cast<int*>(clear);
        constexpr int kWords = sizeof(DevMeta) / sizeof(int);
        for (int i = threadIdx.x; i < kWords; i += 256) w[i] = 0;
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(host_flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_readback(const int* src, int* host_dst, int words, const long long* src2, long long* host_dst2, int n2,
                     DevMeta* clear, int* host_flag, hipStream_t s) {
    hipLaunchKernelGGL(k_readback, dim3(1), dim3(256), 0, s, src, host_dst, words, src2, host_dst2, n2, clear,
                       host_flag);
}

void launch_clear(DevMeta* meta, int from_word, hipStream_t s) {
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(256), 0, s, meta, from_word);
}

template <typename V>
void launch_nip(const ProbSet<V>& ps, const BinTable& sym, DevMeta* meta, int2* items, int bin_cap, hipStream_t s) {
    hipLaunchKernelGGL(k_nip<V>, dim3(ps.n_tiles), dim3(kRowThreads), 0, s, ps, sym, meta, items, bin_cap);
}


template <typename V>
void launch_scan_count(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, long long* tile_sums,
                       int after_fail, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_count<V>, dim3(ps.n_tiles), dim3(kRowThreads), 0, s, ps, num, meta, tile_sums,
                       after_fail);
}

template <typename V>
void launch_scan_apply(const ProbSet<V>& ps, const BinTable& num, DevMeta* meta, long long* tile_sums,
                       RowItem* items, long long* gpool_off, long long* inst_nnz, int after_fail,
                       int tile_prefix_min, hipStream_t s) {
    const int scanned = ps.n_tiles >= tile_prefix_min ? 1 : 0;
    if (scanned) hipLaunchKernelGGL(k_tile_prefix, dim3(1), dim3(kRowThreads), 0, s, tile_sums, ps.n_tiles);
    hipLaunchKernelGGL(k_scan_apply<V>, dim3(ps.n_tiles), dim3(kRowThreads), 0, s, ps, num, meta, tile_sums,
                       items, gpool_off, inst_nnz, after_fail, scanned);
}

#define SPG_INSTANTIATE_ROWS(V)                                                                        \
    template void launch_nip<V>(const ProbSet<V>&, const BinTable&, DevMeta*, int2*, int, hipStream_t); \
    template void launch_scan_count<V>(const ProbSet<V>&, const BinTable&, DevMeta*, long long*, int,  \
                                       hipStream_t);                                                   \
    template void launch_scan_apply<V>(const ProbSet<V>&, const BinTable&, DevMeta*, long long*,       \
                                       RowItem*, long long*, long long*, int, int, hipStream_t);

SPG_INSTANTIATE_ROWS(float)
SPG_INSTANTIATE_ROWS(double)

}  // namespace spg
