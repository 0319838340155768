// Launch interface of the filter kernels (tdp_filter.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdp {

struct FilterParams {
    int n;
    int cycle_len;
    int triple_len;
    int move_count;          // 2-move candidate triples of the cycle
    int skip_2move;
    int rank32;              // (n-1)! < 2^32: 32-bit rank arithmetic
    uint8_t cycle[64];
    uint8_t triple[64];
    const uint32_t *moves;   // packed a | b << 8 | c << 16 symbols
    const uint32_t *move_count_dev;  // when set: the number of moves, in device memory
    uint64_t fact[24];
    uint64_t gfact[24];      // fact[k] / group size, k >= filter_tail(n)
    uint64_t rank_offset;    // first rank of this launch (a multiple of the group size)
    uint64_t rank_end;       // one past the last rank of this launch (likewise, or the total)
    uint32_t *mask;          // one word per group: bit j = rank group * size + j passes
    int split_log2;          // log2 of the threads per group (0..2): 24 >> split_log2 ranks each
    uint32_t *block_counts;
    uint8_t *out;
    uint64_t out_cap;
};

// Each thread of the mask and write passes takes one group of consecutive
// ranks: the t! orders of the last t = min(4, n-1) positions after a common
// prefix (ranks g t! .. g t! + t! - 1 share every Lehmer digit but the last
// t), so the prefix is unranked once per group.
__host__ __device__ inline int filter_tail(int n) { return n - 1 < 4 ? (n > 0 ? n - 1 : 0) : 4; }
__host__ __device__ inline int filter_group_size(int n) {
    const int t = filter_tail(n);
    return t == 4 ? 24 : t == 3 ? 6 : t == 2 ? 2 : 1;
}

// Threads per group in the mask and write passes: one (24 ranks each, the
// prefix unranked once per group) when a launch has groups enough to fill the
// GPU, four (6 ranks each) for small totals.
__host__ __device__ inline int filter_split_log2(uint64_t groups) { return groups >= (1ULL << 20) ? 0 : 2; }
constexpr int kFilterMaxThreadsPerGroup = 4;

// The mask and write passes: *_kernel for 5 <= n <= 16 (filter_nibbles),
// *_any_kernel for the other n.
__host__ __device__ inline bool filter_nibbles(int n) { return n >= 5 && n <= 16; }
__global__ void filter_mask_kernel(FilterParams F);
__global__ void filter_mask_any_kernel(FilterParams F);
__global__ void filter_scan_kernel(const uint32_t *counts, uint64_t nblocks, uint64_t *offsets,
                                   uint64_t *total);
__global__ void filter_write_kernel(FilterParams F, const uint64_t *offsets);
__global__ void filter_write_any_kernel(FilterParams F, const uint64_t *offsets);
// Set-up pass of the allocation-free filter: the cycle's 2-move candidate
// triples into moves[0..*count) (lexicographic order of cycle indices) and
// *total cleared, by one 64-lane wave.
__global__ void filter_moves_kernel(FilterParams F, uint32_t *moves, uint32_t *count, uint64_t *total);

}  // namespace tdp
