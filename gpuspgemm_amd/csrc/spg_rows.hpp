// Row-level building blocks shared by the symbolic and fused passes.
//
// A row of C is worked on by a group of L consecutive lanes (L <= 64: the
// group lives inside one wave and synchronises at wave level; L > 64: the
// group is the whole workgroup).  Its products a_ik * b_kj are enumerated
// flat: the group loads the metadata of EC entries of A's row at once (B
// row start and length, a_ik), takes an exclusive prefix of the B-row
// lengths, and lane lr handles products p = lr, lr + L, ...  Consecutive
// lanes read consecutive entries of a B row (coalesced), every lane has up
// to PMAX independent loads in flight, and short B rows leave no lane idle.
// Product p belongs to the last entry that starts at or before p.  A group
// of one whole wave (64 lanes) finds it by counting: per round of 64 * PMAX
// products it sets one bit per entry start in an LDS bitmap, one 64-bit
// word per slot, and the entry of p is the number of set bits up to p
// (entries with empty B rows are dropped from the chunk first).  Narrower
// and wider groups binary-search the entries' product offsets (measured:
// for narrow groups, whose chunks are short, the counting costs more
// registers than the search saves).
#pragma once

#include "spg_common.hpp"

namespace spg {

template <int L>
__device__ __forceinline__ void gsync() {
    if constexpr (L > kWave)
        __syncthreads();
    else
        wave_sync();
}

// Exclusive prefix of v over the L lanes of a row group; `total` gets the
// group sum.  For L > 64 the group is the workgroup and s_tmp holds one int
// per wave.
template <int L>
__device__ __forceinline__ int group_excl_scan(int v, int lr, int* s_tmp, int& total) {
    if constexpr (L == 1) {
        total = v;
        return 0;
    } else if constexpr (L <= kWave) {
        int incl = v;
#pragma unroll
        for (int o = 1; o < L; o <<= 1) {
            const int y = __shfl_up(incl, o, L);
            if (lr >= o) incl += y;
        }
        total = __shfl(incl, L - 1, L);
        return incl - v;
    } else {
        constexpr int W = L / kWave;
        const int lane = lr % kWave;
        const int w = lr / kWave;
        int incl = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == kWave - 1) s_tmp[w] = incl;
        __syncthreads();
        int before = 0, all = 0;
#pragma unroll
        for (int i = 0; i < W; ++i) {
            const int s = s_tmp[i];
            before += i < w ? s : 0;
            all += s;
        }
        __syncthreads();
        total = all;
        return before + incl - v;
    }
}

// Metadata of one chunk of EC entries of A's row, per row group (in LDS for
// L > 1, in registers for a lane that owns its row).  Entry e's row of B
// comes from P.bref[e] (start, length), recorded by the NIP pass.
template <typename V, int EC, bool VALS>
struct alignas(8) ChunkMeta {
    // L != 64: exclusive prefix of B-row lengths, entries past the chunk
    // holding the chunk total.  L == 64: the bitmap of entry starts of the
    // current round (one 64-bit word per slot).
    int off[EC];
    // position in B of the entry's first product; L == 64: minus the
    // entry's product offset
    int d[EC];
    V av[VALS ? EC : 1];     // a_ik
};

// Load entries [ebase, min(ebase + EC, a1)) of A's row into `cm`; returns
// the chunk's product count.  Each lane loads EC/L consecutive entries.  A
// group of one whole wave keeps only entries with a non-empty B row, in
// order, and each lane gets the product offsets where its own entries
// start (start[t], -1 for an empty one).
template <typename V, int L, int EC, bool VALS>
__device__ __forceinline__ int load_chunk(const GProb<V>& P, int ebase, int a1, int lr,
                                          ChunkMeta<V, EC, VALS>* cm, int* s_tmp, int (&start)[EC / L]) {
    static_assert(EC % L == 0, "chunk must be a multiple of the group");
    constexpr int PER = EC / L;
    int len[PER], bs[PER];
    V av[PER];
    // all loads first, into registers with no default to merge with (a
    // merge at the end of the branch would wait for each load in turn)
    long long ref[PER];
    V a[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = ebase + lr * PER + t;
        if (e < a1) {
            ref[t] = P.bref[e];
            if constexpr (VALS) a[t] = P.a_val[e];
        }
    }
    int local = 0, filled = 0;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const bool in = ebase + lr * PER + t < a1;
        bs[t] = in ? bref_start(ref[t]) : 0;
        len[t] = in ? bref_len(ref[t]) : 0;
        av[t] = V(0);
        if constexpr (VALS) av[t] = in ? a[t] : V(0);
        local += len[t];
        filled += len[t] > 0 ? 1 : 0;
    }
    int total;
    int run = group_excl_scan<L>(local, lr, s_tmp, total);
    if constexpr (L == kWave) {
        int n_filled;
        int k = group_excl_scan<L>(filled, lr, s_tmp, n_filled);
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            start[t] = -1;
            if (len[t] > 0) {
                cm->d[k] = bs[t] - run;
                if constexpr (VALS) cm->av[k] = av[t];
                start[t] = run;
                ++k;
            }
            run += len[t];
        }
    } else {
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = lr * PER + t;
            cm->off[e] = run;
            cm->d[e] = bs[t];
            if constexpr (VALS) cm->av[e] = av[t];
            start[t] = run;
            run += len[t];
        }
    }
    gsync<L>();
    return total;
}

// Entry of product p[j] (0 <= p[j] < total) for every slot j: the largest
// e with off[e] <= p[j].  The slots' searches advance in lockstep, so each
// step issues PMAX independent LDS reads and waits once.
template <int EC, int PMAX>
__device__ __forceinline__ void chunk_find_all(const int* off, const int (&p)[PMAX], int (&e)[PMAX]) {
#pragma unroll
    for (int j = 0; j < PMAX; ++j) e[j] = 0;
#pragma unroll
    for (int step = EC / 2; step > 0; step >>= 1) {
        int o[PMAX];
#pragma unroll
        for (int j = 0; j < PMAX; ++j) o[j] = off[e[j] + step];
#pragma unroll
        for (int j = 0; j < PMAX; ++j) e[j] += o[j] <= p[j] ? step : 0;
    }
}

// Fetch up to PMAX products of a chunk: slot j holds product
// base + j*L + lr.  ok[j] tells which slots are real.  `prior` counts the
// entries that start before `base` (0 for a chunk's first round) and is
// advanced past this round.  Three phases, so that a lane's gathers are in
// flight together: the entry of every slot, then the loads of B's columns
// and values for all slots, then the products a_ik * b_kj (a multiply next
// to its load would wait for that load before the next one is issued).
template <typename V, int L, int EC, int PMAX, bool VALS>
__device__ __forceinline__ void fetch_products(const GProb<V>& P, ChunkMeta<V, EC, VALS>* cm, int total, int base,
                                               int lr, const int (&start)[EC / L], int& prior, int (&key)[PMAX],
                                               V (&prod)[PMAX], bool (&ok)[PMAX]) {
    int e[PMAX];
    if constexpr (L == kWave) {
        // slot j covers bitmap word j; lane lr counts the set bits at or
        // below its own position.  (Moving the word to scalar registers and
        // counting with mbcnt measured slower.)
        static_assert(kWave == 64, "one bitmap word per slot");
        static_assert(2 * PMAX <= EC, "the bitmap lives in the chunk's offset array");
        unsigned long long* bits = reinterpret_cast<unsigned long long*>(cm->off);
        if (lr < PMAX) bits[lr] = 0;
        wave_sync();
#pragma unroll
        for (int t = 0; t < EC / L; ++t) {
            const int q = start[t] - base;
            if (start[t] >= 0 && q >= 0 && q < L * PMAX) atomicOr(&bits[q >> 6], 1ull << (q & 63));
        }
        wave_sync();
        unsigned long long w[PMAX];
#pragma unroll
        for (int j = 0; j < PMAX; ++j) w[j] = bits[j];
        const unsigned long long upto = ~0ull >> (63 - lr);  // positions 0..lr
        int cnt = prior;  // entries starting before slot j's window
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            // clamped for slots past the chunk (int operands: a mixed
            // int / unsigned max resolves to the double overload)
            const int before = cnt + static_cast<int>(__popcll(w[j] & upto)) - 1;
            e[j] = before > 0 ? before : 0;
            cnt += static_cast<int>(__popcll(w[j]));
        }
        prior = cnt;
        wave_sync();  // the bitmap is rewritten by the next round
    } else {
        int p[PMAX];
#pragma unroll
        for (int j = 0; j < PMAX; ++j) p[j] = base + j * L + lr;
        chunk_find_all<EC, PMAX>(cm->off, p, e);
    }
    int at[PMAX];
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        const int p = base + j * L + lr;
        ok[j] = p < total;
        // L == 64: d holds the entry's B position minus its product offset
        at[j] = L == kWave ? p + cm->d[e[j]] : cm->d[e[j]] + p - cm->off[e[j]];
    }
    // The loads write registers that hold nothing else (no default to
    // merge with at the end of the branch, which would cost a copy and a
    // wait for the load inside the loop); defaults are selected afterwards.
    int kk[PMAX];
    V bv[PMAX];
#pragma unroll
    for (int j = 0; j < PMAX; ++j)
        if (ok[j]) {
            kk[j] = P.b_col[at[j]];
            if constexpr (VALS) bv[j] = P.b_val[at[j]];
        }
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        key[j] = ok[j] ? kk[j] : 0;
        if constexpr (VALS)
            prod[j] = ok[j] ? bv[j] * cm->av[e[j]] : V(0);
        else
            prod[j] = V(0);
    }
}

// A lane that owns its row keeps the chunk metadata in registers.
template <typename V, int EC, bool VALS>
__device__ __forceinline__ int load_chunk_lane(const GProb<V>& P, int ebase, int a1, ChunkMeta<V, EC, VALS>& cm) {
    int run = 0;
    int bs[EC], len[EC];
    long long ref[EC];
    V a[EC];
#pragma unroll
    for (int t = 0; t < EC; ++t)
        if (ebase + t < a1) {
            ref[t] = P.bref[ebase + t];
            if constexpr (VALS) a[t] = P.a_val[ebase + t];
        }
#pragma unroll
    for (int t = 0; t < EC; ++t) {
        const bool in = ebase + t < a1;
        bs[t] = in ? bref_start(ref[t]) : 0;
        len[t] = in ? bref_len(ref[t]) : 0;
        if constexpr (VALS) cm.av[t] = in ? a[t] : V(0);
    }
#pragma unroll
    for (int t = 0; t < EC; ++t) {
        cm.off[t] = run;
        cm.d[t] = bs[t];
        run += len[t];
    }
    return run;
}

// The same for a lane that owns its row (metadata in registers): every
// slot's B position first, then all loads, then the products.
template <typename V, int EC, int PMAX, bool VALS>
__device__ __forceinline__ void fetch_products_lane(const GProb<V>& P, const ChunkMeta<V, EC, VALS>& cm, int total,
                                                    int base, int (&key)[PMAX], V (&prod)[PMAX], bool (&ok)[PMAX]) {
    int q[PMAX];
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        const int p = base + j;
        ok[j] = p < total;
        int off = cm.off[0], bs = cm.d[0];
#pragma unroll
        for (int t = 1; t < EC; ++t)
            if (cm.off[t] <= p) {
                off = cm.off[t];
                bs = cm.d[t];
            }
        q[j] = bs + p - off;
    }
    int kk[PMAX];
    V bv[PMAX];
#pragma unroll
    for (int j = 0; j < PMAX; ++j)
        if (ok[j]) {
            kk[j] = P.b_col[q[j]];
            if constexpr (VALS) bv[j] = P.b_val[q[j]];
        }
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        key[j] = ok[j] ? kk[j] : 0;
        prod[j] = V(0);
        if constexpr (VALS) {
            const int p = base + j;
            V av = cm.av[0];
#pragma unroll
            for (int t = 1; t < EC; ++t)
                if (cm.off[t] <= p) av = cm.av[t];
            prod[j] = ok[j] ? bv[j] * av : V(0);
        }
    }
}

// Call g(key[], prod[], ok[]) for every round of up to PMAX products per
// lane of A's entries [a0, a1).
template <typename V, int L, int EC, int PMAX, bool VALS, typename G>
__device__ __forceinline__ void for_each_round(const GProb<V>& P, int a0, int a1, int lr,
                                               ChunkMeta<V, EC, VALS>* cm, int* s_tmp, G&& g) {
    for (int ebase = a0; ebase < a1; ebase += EC) {
        int key[PMAX];
        V prod[PMAX];
        bool ok[PMAX];
        if constexpr (L == 1) {
            ChunkMeta<V, EC, VALS> rm;
            const int total = load_chunk_lane<V, EC, VALS>(P, ebase, a1, rm);
            for (int base = 0; base < total; base += PMAX) {
                fetch_products_lane<V, EC, PMAX, VALS>(P, rm, total, base, key, prod, ok);
                g(key, prod, ok);
            }
        } else {
            int start[EC / L];
            const int total = load_chunk<V, L, EC, VALS>(P, ebase, a1, lr, cm, s_tmp, start);
            int prior = 0;
            for (int base = 0; base < total; base += L * PMAX) {
                fetch_products<V, L, EC, PMAX, VALS>(P, cm, total, base, lr, start, prior, key, prod, ok);
                g(key, prod, ok);
            }
            gsync<L>();  // the chunk metadata is overwritten next
        }
    }
}

// Call f(key, prod) for every product of A's entries [a0, a1).
template <typename V, int L, int EC, int PMAX, bool VALS, typename F>
__device__ __forceinline__ void for_each_product(const GProb<V>& P, int a0, int a1, int lr,
                                                 ChunkMeta<V, EC, VALS>* cm, int* s_tmp, F&& f) {
    for_each_round<V, L, EC, PMAX, VALS>(P, a0, a1, lr, cm, s_tmp,
                                         [&](const int(&key)[PMAX], const V(&prod)[PMAX], const bool(&ok)[PMAX]) {
#pragma unroll
                                             for (int j = 0; j < PMAX; ++j)
                                                 if (ok[j]) f(key[j], prod[j]);
                                         });
}

// Fill n ints of LDS with v using 8-byte stores where possible (p 8-byte aligned, n even).
template <int N, int L>
__device__ __forceinline__ void lds_fill(int* p, int v, int lr) {
    if constexpr (N % 2 == 0 && N >= 2 * L) {
        int2* q = reinterpret_cast<int2*>(p);
        const int2 w = make_int2(v, v);
#pragma unroll 4
        for (int j = lr; j < N / 2; j += L) q[j] = w;
    } else {
        for (int j = lr; j < N; j += L) p[j] = v;
    }
}

// Bitonic sorting network on P ints held in registers by the L lanes of a
// group (L <= 64): lane lr holds elements base + lr*E .. base + lr*E + E-1
// in x[0..E), E = P / L.  `base` is the index of the group's first element
// in a larger network: the direction of every compare-exchange follows the
// global index, so groups can run the first levels of a network that
// several groups finish together.  Stages whose partner lies in the same
// lane are register compare-exchanges; the others exchange with lane
// lr ^ (j / E) through cross-lane shuffles, E of them in flight at once.

// The value of lane lr ^ LJ (LJ < L <= 64, groups of L aligned lanes):
// a DPP quad permutation for distances 1 and 2, a swizzle (xor within 32
// lanes, no address operand) up to 16, a general cross-lane read for 32.
template <int LJ, int L>
__device__ __forceinline__ int lane_xor(int v) {
    static_assert(LJ >= 1 && LJ < L && L <= kWave, "bad lane distance");
    if constexpr (LJ == 1)
        return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    else if constexpr (LJ == 2)
        return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (LJ < 32)
        return __builtin_amdgcn_ds_swizzle(v, (LJ << 10) | 0x1F);
    else
        return __shfl_xor(v, LJ, L);
}

// One stage (level k, distance J) of the network.
template <int E, int L, int J>
__device__ __forceinline__ void bitonic_step_regs(int (&x)[E], int lr, int base, int k) {
    if constexpr (J >= E) {
        constexpr int LJ = J / E;
        const bool lower = (lr & LJ) == 0;
        int other[E];
#pragma unroll
        for (int s = 0; s < E; ++s) other[s] = lane_xor<LJ, L>(x[s]);
#pragma unroll
        for (int s = 0; s < E; ++s) {
            const bool asc = ((base + lr * E + s) & k) == 0;
            x[s] = (lower == asc) ? min(x[s], other[s]) : max(x[s], other[s]);
        }
    } else {
#pragma unroll
        for (int s = 0; s < E; ++s) {
            if (s & J) continue;
            const bool asc = ((base + lr * E + s) & k) == 0;
            const int lo = min(x[s], x[s + J]);
            const int hi = max(x[s], x[s + J]);
            x[s] = asc ? lo : hi;
            x[s + J] = asc ? hi : lo;
        }
    }
}

// Stages J, J/2, ..., 1 of level k.
template <int E, int L, int J>
__device__ __forceinline__ void bitonic_merge_regs(int (&x)[E], int lr, int base, int k) {
    if constexpr (J >= 1) {
        bitonic_step_regs<E, L, J>(x, lr, base, k);
        bitonic_merge_regs<E, L, J / 2>(x, lr, base, k);
    }
}

// Levels 2, 4, ..., K: a full sort of the group's P elements when K = P
// (ascending when base is a multiple of 2P).
template <int E, int L, int K>
__device__ __forceinline__ void bitonic_levels_regs(int (&x)[E], int lr, int base) {
    if constexpr (K >= 2) {
        bitonic_levels_regs<E, L, K / 2>(x, lr, base);
        bitonic_merge_regs<E, L, K / 2>(x, lr, base, K);
    }
}

template <int P, int L>
__device__ __forceinline__ void bitonic_sort_regs(int (&x)[P / L], int lr, int base = 0) {
    constexpr int E = P / L;
    static_assert(E >= 1 && E * L == P && L <= kWave, "bad register sort shape");
    bitonic_levels_regs<E, L, P>(x, lr, base);
}

// Ascending bitonic sort of P ints in LDS by a workgroup of L > 64 lanes.
// Every wave sorts its block of B = P / waves elements in registers; each
// further level runs its cross-block stages in LDS (a barrier each) and
// its in-block stages in registers again.  For P = 8192 on 1024 lanes this
// is 10 LDS stages and 5 register passes instead of 91 LDS stages.
template <int P, int L>
__device__ __forceinline__ void bitonic_sort_block(int* cols, int lr) {
    constexpr int W = L / kWave;
    constexpr int B = P / W;
    constexpr int E = B / kWave;
    static_assert(L > kWave && L % kWave == 0 && E >= 1 && E * kWave * W == P, "bad block sort shape");
    const int lane = lr % kWave;
    const int base = (lr / kWave) * B;
    int x[E];
#pragma unroll
    for (int s = 0; s < E; ++s) x[s] = cols[base + lane * E + s];
    bitonic_sort_regs<B, kWave>(x, lane, base);
#pragma unroll
    for (int s = 0; s < E; ++s) cols[base + lane * E + s] = x[s];
    __syncthreads();
    for (int k = 2 * B; k <= P; k <<= 1) {
        for (int j = k >> 1; j >= B; j >>= 1) {
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
#pragma unroll
        for (int s = 0; s < E; ++s) x[s] = cols[base + lane * E + s];
        bitonic_merge_regs<E, kWave, B / 2>(x, lane, base, k);
#pragma unroll
        for (int s = 0; s < E; ++s) cols[base + lane * E + s] = x[s];
        __syncthreads();
    }
}

}  // namespace spg
