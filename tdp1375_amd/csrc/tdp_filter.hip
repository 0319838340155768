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
// For n <= 16 the unranking keeps the remaining-symbol list and the symbol
// positions as 4-bit fields of two 64-bit registers (no memory traffic); for
// 17 <= n <= 21 a per-thread byte array is used.
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

template <typename PosFn>
__device__ __forceinline__ bool accepts(const FilterParams &F, PosFn pos) {
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
        const int move_count = F.move_count_dev ? (int)*F.move_count_dev : F.move_count;
        for (int t = 0; t < move_count; ++t) {
            const uint32_t m = F.moves[t];
            const int ia = pos(m & 0xFF), ib = pos((m >> 8) & 0xFF), ic = pos((m >> 16) & 0xFF);
            if ((ia > ib) + (ib > ic) + (ic > ia) <= 1) return false;
        }
    }
    return true;
}

__device__ bool rank_accepts(const FilterParams &F, uint64_t r) {
    if (F.n <= 16) {
        const uint64_t P = unrank_pos16(r, F.n, F, nullptr);
        return accepts(F, [&](int s) { return nib(P, s); });
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
    return accepts(F, [&](int s) { return (int)pos[s]; });
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

}  // namespace

// Pass 1: one thread per rank; ballot mask per wave, popcount per block.
__global__ __launch_bounds__(kBlock) void filter_mask_kernel(FilterParams F) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    const uint64_t r = F.rank_offset + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool pass = (r < F.rank_end) && rank_accepts(F, r);
    const unsigned long long mask = __ballot(pass);
    const int wib = threadIdx.x >> 6;
    const uint64_t word = ((uint64_t)blockIdx.x * kBlock) / 64 + wib;
    if ((threadIdx.x & 63) == 0) {
        F.mask[word] = mask;
        wave_counts[wib] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        for (int w = 0; w < kBlock / 64; ++w) c += wave_counts[w];
        F.block_counts[blockIdx.x] = c;
    }
}

// Pass 2: exclusive scan of the block counts (single block, 1024 threads),
// continuing the running total of earlier rank chunks held in *total.
__global__ __launch_bounds__(1024) void filter_scan_kernel(const uint32_t *counts, uint64_t nblocks,
                                                           uint64_t *offsets, uint64_t *total) {
    __shared__ uint64_t partial[1024];
    const uint64_t per = (nblocks + 1023) / 1024;
    const uint64_t lo = threadIdx.x * per;
    const uint64_t hi = (lo + per < nblocks) ? lo + per : nblocks;
    uint64_t s = 0;
    for (uint64_t b = lo; b < hi; ++b) s += counts[b];
    partial[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = *total;
        for (int t = 0; t < 1024; ++t) {
            const uint64_t v = partial[t];
            partial[t] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    uint64_t run = partial[threadIdx.x];
    for (uint64_t b = lo; b < hi; ++b) {
        offsets[b] = run;
        run += counts[b];
    }
}

// Pass 3: rank-ordered output of the passing arrangements below out_cap.
__global__ __launch_bounds__(kBlock) void filter_write_kernel(FilterParams F, const uint64_t *offsets) {
    const uint64_t block_off = offsets[blockIdx.x];
    if (block_off >= F.out_cap) return;
    const int wib = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const uint64_t word = ((uint64_t)blockIdx.x * kBlock) / 64;
    uint64_t before = 0;
    for (int w = 0; w < wib; ++w) before += __popcll(F.mask[word + w]);
    const unsigned long long m = F.mask[word + wib];
    if (!((m >> lane) & 1ULL)) return;
    const unsigned long long below = (lane == 0) ? 0ULL : (m & (~0ULL >> (64 - lane)));
    const uint64_t idx = block_off + before + (uint64_t)__popcll(below);
    if (idx >= F.out_cap) return;
    const uint64_t r = F.rank_offset + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint8_t row[64];
    rank_row(F, r, row);
    uint8_t *dst = F.out + idx * (uint64_t)F.n;
    for (int p = 0; p < F.n; ++p) dst[p] = row[p];
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
