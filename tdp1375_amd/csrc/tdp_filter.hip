// Lehmer-rank permutation filter for gfx950.
//
// Replaces pi_filter_kernel + performGpuPiFilter (reference
// src/native/backtracking/pi_filter.cu:38-260).  Same enumeration (pi[0]=0,
// ranks of {1..n-1} in lexicographic order, pi_filter.cu:38-53) and the same
// three predicates (pi_filter.cu:55-145):
//   1. the oriented triple is in cyclic order in pi;
//   2. the cycle is oriented: NOT in cyclic order in the reversed pi;
//   3. unless skipped, the cycle has no 2-move (a triple of the cycle in
//      cyclic order in pi whose three arcs add two odd cycles).
// Instead of an unordered atomic append capped at 2^20, pass 1 writes one
// 64-bit ballot mask per wave (64 consecutive ranks) plus a per-block count,
// pass 2 scans the counts, pass 3 writes the passing arrangements at their
// rank-ordered positions (coalesced rows), so the output is the first
// out_cap passing ranks and the uncapped count is exact.
//
// One thread, or for small totals four, take a group of 24 consecutive ranks
// (n >= 5): the ranks share every Lehmer digit but the last four, so the
// prefix (positions 1..n-5) is unranked once per thread instead of once per
// rank and the orders of the four remaining symbols come from a table; the
// mask is one 24-bit word per group.  For n <= 16 the
// unranking keeps the remaining-symbol list and the symbol positions as
// 4-bit fields of two 64-bit registers (no memory traffic); for 17 <= n <= 21
// (and n <= 4, groups of (n-1)! ranks) each rank is unranked on its own with
// a per-thread byte array.
#include <hip/hip_runtime.h>

#include "tdp_filter.h"

namespace tdp {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int nib(uint64_t w, int i) { return (int)((w >> (4 * i)) & 15ULL); }

// Positions of every symbol for rank r (n <= 16), as nibbles: pos(s) = nib(P, s).
__device__ __forceinline__ uint64_t unrank_pos16(uint64_t r, int n, const FilterParams &F,
                                                 uint8_t *row /* may be null */) {
    uint64_t avail = 0;
    for (int q = 0; q < n - 1; ++q) avail |= (uint64_t)(q + 1) << (4 * q);
    uint64_t pos = 0;
    if (row) row[0] = 0;
    for (int p = 0; p < n - 1; ++p) {
        const uint64_t f = F.fact[n - 2 - p];
        int choice;
        if (F.rank32) {
            const uint32_t r32 = (uint32_t)r, f32 = (uint32_t)f;
            choice = (int)(r32 / f32);
            r = r32 - (uint32_t)choice * f32;
        } else {
            choice = (int)(r / f);
            r -= (uint64_t)choice * f;
        }
        const int sym = nib(avail, choice);
        const uint64_t lowmask = (choice == 0) ? 0ULL : ((1ULL << (4 * choice)) - 1ULL);
        avail = (avail & lowmask) | ((avail >> (4 * (choice + 1))) << (4 * choice));
        pos |= (uint64_t)(p + 1) << (4 * sym);
        if (row) row[p + 1] = (uint8_t)sym;
    }
    return pos;
}

// The 24 orders of four items in lexicographic order (the last four Lehmer
// digits 0..23): kLexFwd4[j] holds the item at each place m (bits 2m..2m+1),
// kLexInv4[j] the place of each item k (bits 2k..2k+1).
__constant__ uint8_t kLexFwd4[24] = {0xE4, 0xB4, 0xD8, 0x78, 0x9C, 0x6C, 0xE1, 0xB1, 0xC9, 0x39, 0x8D, 0x2D,
                                     0xD2, 0x72, 0xC6, 0x36, 0x4E, 0x1E, 0x93, 0x63, 0x87, 0x27, 0x4B, 0x1B};
__constant__ uint8_t kLexInv4[24] = {0xE4, 0xB4, 0xD8, 0x9C, 0x78, 0x6C, 0xE1, 0xB1, 0xD2, 0x93, 0x72, 0x63,
                                     0xC9, 0x8D, 0xC6, 0x87, 0x4E, 0x4B, 0x39, 0x2D, 0x36, 0x27, 0x1E, 0x1B};

// Prefix of group g (ranks 24 g .. 24 g + 23), 5 <= n <= 16: the positions of
// the symbols placed at positions 1..n-5 as nibbles, the four symbols left in
// `avail` (nibbles 0..3, ascending), and in `row` the symbols at positions
// 0..n-5 as nibbles (position p in nibble p).
__device__ __forceinline__ uint64_t unrank_prefix16(uint64_t g, int n, const FilterParams &F, uint64_t &avail,
                                                    uint64_t &row) {
    avail = 0;
    for (int q = 0; q < n - 1; ++q) avail |= (uint64_t)(q + 1) << (4 * q);
    uint64_t pos = 0;
    row = 0;
    for (int p = 0; p < n - 5; ++p) {
        const uint64_t f = F.gfact[n - 2 - p];   // (n-2-p)! / 24
        int choice;
        if (F.rank32) {
            const uint32_t g32 = (uint32_t)g, f32 = (uint32_t)f;
            choice = (int)(g32 / f32);
            g = g32 - (uint32_t)choice * f32;
        } else {
            choice = (int)(g / f);
            g -= (uint64_t)choice * f;
        }
        const int sym = nib(avail, choice);
        const uint64_t lowmask = (choice == 0) ? 0ULL : ((1ULL << (4 * choice)) - 1ULL);
        avail = (avail & lowmask) | ((avail >> (4 * (choice + 1))) << (4 * choice));
        pos |= (uint64_t)(p + 1) << (4 * sym);
        row |= (uint64_t)sym << (4 * (p + 1));
    }
    return pos;
}

// The cycle's 2-move triples for the predicate: the first kMovesLds staged in
// the block's LDS by the mask pass, the rest read from device memory.
constexpr int kMovesLds = 256;
struct Moves {
    const uint32_t *lds;   // null: all from device memory
    int count;
    __device__ __forceinline__ uint32_t at(const FilterParams &F, int t) const {
        return (lds && t < kMovesLds) ? lds[t] : F.moves[t];
    }
};

__device__ __forceinline__ Moves global_moves(const FilterParams &F) {
    Moves m;
    m.lds = nullptr;
    m.count = F.move_count_dev ? (int)*F.move_count_dev : F.move_count;
    return m;
}

template <typename PosFn>
__device__ __forceinline__ bool accepts(const FilterParams &F, const Moves &M, PosFn pos) {
    // 1. oriented triple in cyclic order (pi_filter.cu:66-77,133)
    int leaps = 0;
    for (int i = 0; i < F.triple_len; ++i) {
        const int cur = pos(F.triple[i]);
        const int next = pos(F.triple[(i + 1) == F.triple_len ? 0 : i + 1]);
        leaps += cur > next;
    }
    if (leaps > 1) return false;
    // 2. cycle oriented: not in cyclic order in the reversed arrangement
    //    (positions n-1-pos, pi_filter.cu:135-142)
    leaps = 0;
    for (int i = 0; i < F.cycle_len; ++i) {
        const int cur = pos(F.cycle[i]);
        const int next = pos(F.cycle[(i + 1) == F.cycle_len ? 0 : i + 1]);
        leaps += cur < next;
    }
    if (leaps <= 1) return false;
    // 3. no 2-move among the host-listed cycle triples (pi_filter.cu:86-115)
    if (!F.skip_2move) {
        for (int t = 0; t < M.count; ++t) {
            const uint32_t m = M.at(F, t);
            const int ia = pos(m & 0xFF), ib = pos((m >> 8) & 0xFF), ic = pos((m >> 16) & 0xFF);
            if ((ia > ib) + (ib > ic) + (ic > ia) <= 1) return false;
        }
    }
    return true;
}

__device__ bool rank_accepts(const FilterParams &F, const Moves &M, uint64_t r) {
    if (F.n <= 16) {
        const uint64_t P = unrank_pos16(r, F.n, F, nullptr);
        return accepts(F, M, [&](int s) { return nib(P, s); });
    }
    uint8_t avail[64], pos[64];
    for (int i = 0; i < F.n - 1; ++i) avail[i] = (uint8_t)(i + 1);
    pos[0] = 0;
    for (int p = 0; p < F.n - 1; ++p) {
        const uint64_t f = F.fact[F.n - 2 - p];
        const int choice = (int)(r / f);
        r %= f;
        pos[avail[choice]] = (uint8_t)(p + 1);
        for (int j = choice; j < F.n - 2 - p; ++j) avail[j] = avail[j + 1];
    }
    return accepts(F, M, [&](int s) { return (int)pos[s]; });
}

__device__ void rank_row(const FilterParams &F, uint64_t r, uint8_t *row) {
    if (F.n <= 16) {
        unrank_pos16(r, F.n, F, row);
        return;
    }
    uint8_t avail[64];
    for (int i = 0; i < F.n - 1; ++i) avail[i] = (uint8_t)(i + 1);
    row[0] = 0;
    for (int p = 0; p < F.n - 1; ++p) {
        const uint64_t f = F.fact[F.n - 2 - p];
        const int choice = (int)(r / f);
        r %= f;
        row[p + 1] = avail[choice];
        for (int j = choice; j < F.n - 2 - p; ++j) avail[j] = avail[j + 1];
    }
}

// Thread q of a group takes the orders j = q * orders .. (q + 1) * orders - 1,
// orders = 24 >> F.split_log2.
__device__ __forceinline__ int orders_per_thread(const FilterParams &F) { return 24 >> F.split_log2; }

// Bit j of group g's mask for the orders of thread q: rank g * 24 + j passes
// (5 <= n <= 16).
__device__ __forceinline__ uint32_t group_mask16(const FilterParams &F, const Moves &M, uint64_t g, int q) {
    const int n = F.n;
    uint64_t avail, row;
    const uint64_t P = unrank_prefix16(g, n, F, avail, row);
    int sh[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[k] = 4 * nib(avail, k);
    uint32_t bits = 0;
    const int orders = orders_per_thread(F);
    for (int j = q * orders; j < (q + 1) * orders; ++j) {
        const uint32_t inv = kLexInv4[j];
        uint64_t Pj = P;
#pragma unroll
        for (int k = 0; k < 4; ++k) Pj |= (uint64_t)(n - 4 + (int)((inv >> (2 * k)) & 3u)) << sh[k];
        if (accepts(F, M, [&](int s) { return nib(Pj, s); })) bits |= 1u << j;
    }
    return bits;
}

__device__ __forceinline__ uint64_t group_count(const FilterParams &F, uint64_t R) {
    return (F.rank_end - F.rank_offset + R - 1) / R;
}

}  // namespace

// The orders of group `gid`'s quarter q as mask bits (bit j: rank r0 + j).
template <bool kNib>
__device__ __forceinline__ uint32_t quarter_bits(const FilterParams &F, const Moves &M, uint64_t R, uint64_t gid,
                                                 int q) {
    const uint64_t r0 = F.rank_offset + gid * R;
    if constexpr (kNib) return group_mask16(F, M, r0 / 24, q);
    uint32_t bits = 0;
    const uint64_t orders = (uint64_t)orders_per_thread(F);
    for (uint64_t j = (uint64_t)q * orders; j < R && j < (uint64_t)(q + 1) * orders; ++j)
        if (r0 + j < F.rank_end && rank_accepts(F, M, r0 + j)) bits |= 1u << j;
    return bits;
}

// Pass 1: 1 << F.split_log2 threads per group of ranks; a mask word per
// group, a popcount per block.  kNib: 5 <= n <= 16 (the group path in registers);
// otherwise rank by rank with byte arrays, in a kernel of its own so that
// their registers do not lower the occupancy of the common one.
template <bool kNib>
__device__ __forceinline__ void mask_pass(const FilterParams &F) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    __shared__ uint32_t smoves[kMovesLds];
    Moves M = global_moves(F);
    if (!F.skip_2move) {
        for (int i = threadIdx.x; i < M.count && i < kMovesLds; i += kBlock) smoves[i] = F.moves[i];
        M.lds = smoves;
    }
    __syncthreads();
    const uint64_t R = (uint64_t)filter_group_size(F.n);
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t gid = tid >> F.split_log2;
    const int q = (int)(tid & ((1u << F.split_log2) - 1u));
    const bool valid = gid < group_count(F, R);
    uint32_t bits = valid ? quarter_bits<kNib>(F, M, R, gid, q) : 0u;
    // the group's word from its adjacent lanes
    uint32_t word = bits;
    for (int off = 1; off < (1 << F.split_log2); off <<= 1) word |= (uint32_t)__shfl_xor((int)word, off);
    if (valid && q == 0) F.mask[gid] = word;
    int c = __popc(bits);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    const int wib = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wave_counts[wib] = (uint32_t)c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += wave_counts[w];
        F.block_counts[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void filter_mask_kernel(FilterParams F) { mask_pass<true>(F); }
__global__ __launch_bounds__(kBlock) void filter_mask_any_kernel(FilterParams F) { mask_pass<false>(F); }


// Pass 2: exclusive scan of the block counts (single block, 1024 threads),
// continuing the running total of earlier rank chunks held in *total.
__global__ __launch_bounds__(1024) void filter_scan_kernel(const uint32_t *counts, uint64_t nblocks,
                                                           uint64_t *offsets, uint64_t *total) {
    __shared__ uint64_t wave_sums[1024 / 64];
    const uint64_t per = (nblocks + 1023) / 1024;
    const uint64_t lo = threadIdx.x * per;
    const uint64_t hi = (lo + per < nblocks) ? lo + per : nblocks;
    uint64_t s = 0;
#pragma unroll 8
    for (uint64_t b = lo; b < hi; ++b) s += counts[b];
    // exclusive scan of the 1024 partial sums: within each wave by shuffles,
    // then over the 16 wave totals
    const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    uint64_t incl = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wave_sums[wib] = incl;
    __syncthreads();
    const uint64_t start = *total;
    uint64_t run = start + incl - s;
    uint64_t all = 0;
    for (int w = 0; w < 1024 / 64; ++w) {
        if (w < wib) run += wave_sums[w];
        all += wave_sums[w];
    }
    __syncthreads();   // every thread has read *total
    if (threadIdx.x == 0) *total = start + all;
#pragma unroll 8
    for (uint64_t b = lo; b < hi; ++b) {
        offsets[b] = run;
        run += counts[b];
    }
}

// Pass 3: rank-ordered output of the passing arrangements below out_cap: a
// thread's rows start at its block's offset plus the passing ranks of the
// block's earlier groups (a scan of the groups' popcounts).
template <bool kNib>
__device__ __forceinline__ void write_pass(const FilterParams &F, const uint64_t *offsets) {
    __shared__ uint32_t wave_sums[kBlock / 64];
    const uint64_t block_off = offsets[blockIdx.x];
    if (block_off >= F.out_cap) return;   // the whole block
    const uint64_t R = (uint64_t)filter_group_size(F.n);
    const uint64_t tid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const uint64_t gid = tid >> F.split_log2;
    const int q = (int)(tid & ((1u << F.split_log2) - 1u));
    const int orders = orders_per_thread(F);
    // this thread's orders of the group (threads in rank order)
    const uint32_t qmask = ((1u << orders) - 1u) << (q * orders);
    const uint32_t bits = gid < group_count(F, R) ? (F.mask[gid] & qmask) : 0u;
    const int c = __popc(bits);
    const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
    int incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wave_sums[wib] = (uint32_t)incl;
    __syncthreads();
    uint64_t idx = block_off + (uint64_t)(incl - c);
    for (int w = 0; w < wib; ++w) idx += wave_sums[w];
    if (bits == 0 || idx >= F.out_cap) return;
    const uint64_t r0 = F.rank_offset + gid * R;
    const int n = F.n;
    if constexpr (kNib) {
        uint64_t avail, prefix;
        unrank_prefix16(r0 / 24, n, F, avail, prefix);
        // rows of a multiple of 4 bytes go out as words (the buffer is
        // 4-byte aligned: device allocations are)
        const bool words = (n & 3) == 0 && (reinterpret_cast<uintptr_t>(F.out) & 3u) == 0;
        for (int j = q * orders; j < (q + 1) * orders && idx < F.out_cap; ++j) {
            if (!((bits >> j) & 1u)) continue;
            const uint32_t fwd = kLexFwd4[j];
            uint64_t row = prefix;
#pragma unroll
            for (int m = 0; m < 4; ++m)
                row |= (uint64_t)nib(avail, (int)((fwd >> (2 * m)) & 3u)) << (4 * (n - 4 + m));
            uint8_t *dst = F.out + idx * (uint64_t)n;
            if (words) {
                for (int w = 0; w < n / 4; ++w) {
                    const uint32_t v = (uint32_t)nib(row, 4 * w) | ((uint32_t)nib(row, 4 * w + 1) << 8) |
                                       ((uint32_t)nib(row, 4 * w + 2) << 16) | ((uint32_t)nib(row, 4 * w + 3) << 24);
                    reinterpret_cast<uint32_t *>(dst)[w] = v;
                }
            } else {
                for (int p = 0; p < n; ++p) dst[p] = (uint8_t)nib(row, p);
            }
            ++idx;
        }
    } else {
        uint8_t row[64];
        for (uint64_t j = (uint64_t)(q * orders); j < R && j < (uint64_t)((q + 1) * orders) && idx < F.out_cap;
             ++j) {
            if (!((bits >> j) & 1u)) continue;
            rank_row(F, r0 + j, row);
            uint8_t *dst = F.out + idx * (uint64_t)n;
            for (int p = 0; p < n; ++p) dst[p] = row[p];
            ++idx;
        }
    }
}

__global__ __launch_bounds__(kBlock) void filter_write_kernel(FilterParams F, const uint64_t *offsets) {
    write_pass<true>(F, offsets);
}
__global__ __launch_bounds__(kBlock) void filter_write_any_kernel(FilterParams F, const uint64_t *offsets) {
    write_pass<false>(F, offsets);
}

// The host list of tdp_filter as a kernel (pi_filter.cu:86-115): triples
// i<j<k of cycle indices whose arcs, measured through the symbol -> index
// map, turn the cycle's odd count `before` into before + 2.
__global__ __launch_bounds__(64) void filter_moves_kernel(FilterParams F, uint32_t *moves, uint32_t *count,
                                                          uint64_t *total) {
    __shared__ uint8_t cidx[64];
    const int lane = threadIdx.x;
    const int L = F.cycle_len;
    cidx[lane] = 0;
    __syncthreads();
    // sequential, so a repeated symbol keeps its last index as on the host
    if (lane == 0)
        for (int i = 0; i < L; ++i) cidx[F.cycle[i]] = (uint8_t)i;
    __syncthreads();
    auto dist = [&](int a, int b) {
        const int ai = cidx[a], bi = cidx[b];
        return bi >= ai ? bi - ai : (L - ai) + bi;
    };
    const int before = (L % 2 == 1) ? 1 : 0;
    uint32_t cnt = 0;
    for (int i = 0; i < L - 2; ++i) {
        for (int j = i + 1; j < L - 1; ++j) {
            const int k = j + 1 + lane;
            bool ok = false;
            uint32_t mv = 0;
            if (k < L) {
                const int a = F.cycle[i], b = F.cycle[j], c = F.cycle[k];
                const int after = (dist(a, b) % 2 == 1) + (dist(b, c) % 2 == 1) + (dist(c, a) % 2 == 1);
                ok = after - before == 2;
                mv = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16);
            }
            const unsigned long long m = __ballot(ok);
            const unsigned long long below = (lane == 0) ? 0ULL : (m & (~0ULL >> (64 - lane)));
            if (ok) moves[cnt + (uint32_t)__popcll(below)] = mv;
            cnt += (uint32_t)__popcll(m);
        }
    }
    if (lane == 0) {
        *count = cnt;
        *total = 0;
    }
}

}  // namespace tdp
