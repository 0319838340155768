// Launch interface of the search kernels (tdp_search.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdp_common.h"
#include "tdp_learn.h"

namespace tdp {

// Most nodes a wave expands before one dedup pass over all their children
// (the learned ordering, which keeps each node's cycle info, at most
// kLearnedBatchMax); a batch also fits the LDS node pool of kNodePoolWords
// words at the launch's node size.
constexpr int kExpansionBatchMax = 16;
constexpr int kLearnedBatchMax = 8;
constexpr int kNodePoolWords = 448;
// Children staged per wave before a dedup pass.
constexpr int kStageEntries = 512;

struct SearchParams {
    const ConfigDesc *cfgs;
    ConfigResult *results;
    int num_configs;
    int num_waves;


    uint32_t *stacks;             // num_waves x stack_cap x node_words
    uint32_t stack_cap;           // power of two
    uint32_t node_words;
    uint32_t spi_words;

    // work sharing (parallel mode): per wave an inbox of share_max nodes and
    // an inbox line (count of published nodes, 128-byte stride); one idle bit
    // per wave in idle_words words (128-byte stride)
    uint32_t *inbox_nodes;        // num_waves x share_max x node_words
    unsigned long long *inbox;    // num_waves x 16
    unsigned long long *idle_bits;  // idle_words x 16
    uint32_t idle_words;
    // spill pool for full stacks when no wave is idle
    uint32_t *pool_nodes;         // pool_cap x node_words
    unsigned long long *pool_seq; // pool_cap
    unsigned long long pool_cap;  // power of two

    ControlLine *ctrl;            // kNumControl lines

    // dedup table: dd_buckets lines of dd_width 64-bit tags (one 64-byte
    // line per bucket) and one 64-byte key line (child spi words) per entry
    unsigned long long *dd_tag;   // dd_buckets x dd_width
    uint32_t *dd_key;             // dd_buckets x dd_width x 16 words
    unsigned long long dd_buckets;  // 0: no dedup table
    uint32_t dd_width;            // tags per bucket, 1..8
    uint32_t epoch;               // 1..65535

    const int *cancel_flag;
    int donate;                   // 0: every configuration stays on one wave
    uint32_t share_mask;          // donation check every share_mask + 1 expansions
    int batch;                    // nodes expanded before one dedup pass, 1..kExpansionBatchMax
    uint32_t share_max;           // nodes handed over per donation
    uint32_t idle_backoff;        // longest idle sleep, in units of s_sleep 127
    unsigned long long timeout_ticks;  // watchdog: 100 MHz ticks per wave, 0 = none
    MlpWeights mlp;               // learned ordering (search_learned_kernel only)

    // device-resident serving (serve kernels only): num_configs device slots
    // (cfgs, results), their state words, the request index each holds, the
    // pinned host request ring (ring entries), reply slots and control words
    int serve;
    unsigned long long *slot_state;     // num_configs
    unsigned long long *req_index;      // num_configs
    const ServeRequest *host_req;       // ring entries
    ServeReply *host_reply;             // num_configs
    unsigned long long *host_ctl;       // kServeHostWords
    uint32_t *ring_slot;                // ring entries: device slot of each published request
    uint32_t ring;
};

__global__ void search_kernel(SearchParams P);
__global__ void search_learned_kernel(SearchParams P);
// Device-resident serving: the search waves (num_waves, in blocks of 4) plus
// one more block whose first wave feeds requests from the host ring.  Runs
// until the host sets kServeHostStop, the queues overflow or the watchdog
// fires.
__global__ void serve_kernel(SearchParams P);
__global__ void serve_learned_kernel(SearchParams P);
__global__ void reset_kernel(SearchParams P);
__global__ void eval_pair_kernel(const uint8_t *pia, const uint8_t *spia, const uint8_t *pib, const uint8_t *spib,
                                 int n, int ea, int eb, uint32_t ga, uint32_t aa, uint32_t gb, uint32_t ab,
                                 int32_t *out);
__global__ void eval_children_kernel(const uint8_t *pi, const uint8_t *spi, int n,
                                     int T, int e, int32_t *out, int e0, int d, int D, uint32_t *features);

}  // namespace tdp
