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
    uint64_t rank_offset;    // first rank of this launch
    uint64_t rank_end;       // one past the last rank of this launch
    unsigned long long *mask;
    uint32_t *block_counts;
    uint8_t *out;
    uint64_t out_cap;
};

__global__ void filter_mask_kernel(FilterParams F);
__global__ void filter_scan_kernel(const uint32_t *counts, uint64_t nblocks, uint64_t *offsets,
                                   uint64_t *total);
__global__ void filter_write_kernel(FilterParams F, const uint64_t *offsets);
// Set-up pass of the allocation-free filter: the cycle's 2-move candidate
// triples into moves[0..*count) (lexicographic order of cycle indices) and
// *total cleared, by one 64-lane wave.
__global__ void filter_moves_kernel(FilterParams F, uint32_t *moves, uint32_t *count, uint64_t *total);

}  // namespace tdp
