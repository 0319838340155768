// Launch interface of the search kernels (tdp_search.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdp_common.h"

namespace tdp {

struct SearchParams {
    const ConfigDesc *cfgs;
    ConfigResult *results;
    int num_configs;
    int num_waves;


    uint32_t *stacks;             // num_waves x stack_cap x node_words
    uint32_t stack_cap;           // power of two
    uint32_t node_words;
    uint32_t spi_words;

    uint32_t *pool_nodes;         // pool_cap x node_words
    unsigned long long *pool_seq; // pool_cap
    unsigned long long pool_cap;  // power of two

    ControlLine *ctrl;            // kNumControl lines

    unsigned long long *dd_tag;   // dd_cap
    uint32_t *dd_best;            // dd_cap: epoch << 16 | depth
    uint32_t *dd_cfg;             // dd_cap
    uint32_t *dd_key;             // dd_cap x 16 words (spi bytes)
    unsigned long long dd_cap;    // 0: no dedup table
    uint32_t epoch;               // 1..65535

    const int *cancel_flag;
    int donate;                   // 0: every configuration stays on one wave
    unsigned long long timeout_ticks;  // watchdog: 100 MHz ticks per wave, 0 = none
};

__global__ void search_kernel(SearchParams P);
__global__ void reset_kernel(SearchParams P);
__global__ void eval_children_kernel(const uint8_t *pi, const uint8_t *spi, int n,
                                     int T, int e, int32_t *out);

}  // namespace tdp
