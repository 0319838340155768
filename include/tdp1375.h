/*
 * tdp1375 MI355X proof-search path: C ABI of libtdp1375_jni.so.
 *
 * Plain pointers and sizes only.  Every function returns a status code
 * (TDP_OK on success); on failure tdp_last_error() returns the message of the
 * calling thread's last error, with the reference's wording
 * (sorting_search.cu:710-715,842-845; cuda_error.cuh:56-62 with "HIP" for
 * "CUDA"; cuda_memory.cuh:82-85; pi_filter.cu:165-166).
 *
 * The JNI entry points of br.unb.cic.tdp.base.GPUSortingSearch exported by
 * the same library (jni_shim.cpp) are thin marshalling over these calls.
 */
#ifndef TDP1375_H
#define TDP1375_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define TDP_API __attribute__((visibility("default")))
#else
#define TDP_API
#endif

#define TDP_MAX_N 63
#define TDP_MAX_DEPTH 15
#define TDP_FILTER_MAX_OUTPUT (1 << 20)

enum tdp_status {
    TDP_OK = 0,
    TDP_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference      */
    TDP_ERR_CANCELLED = 2,        /* GpuSortCancelledException                   */
    TDP_ERR_QUEUE_OVERFLOW = 3,   /* GpuMemoryExhaustedException (queue overflow) */
    TDP_ERR_OUT_OF_MEMORY = 4,    /* GpuMemoryExhaustedException (allocation)     */
    TDP_ERR_HIP = 5,              /* CudaException                               */
    TDP_ERR_INTERNAL = 6
};

/* Counters of one search (no counterpart in the reference, which keeps none). */
typedef struct tdp_counters {
    uint64_t expansions;  /* popped nodes with depth < D and can_still_reach   */
    uint64_t generated;   /* children evaluated: C(n,3) per expansion          */
    uint64_t admitted;    /* non-root children pushed                          */
    uint64_t dominated;   /* children rejected by the dedup table              */
    uint64_t improved;    /* admitted because they lowered a stored depth      */
    uint64_t inserted;    /* admitted as a new table entry                     */
    uint64_t uninserted;  /* admitted without a table entry (probe limit)      */
    uint64_t roots_live;  /* roots that are goals or pass can_still_reach      */
} tdp_counters;

/* Result of one configuration of a batch. */
typedef struct tdp_case_result {
    int32_t found;                       /* 1 found, 0 not found              */
    int32_t depth;                       /* moves in the sequence; -1: undecided */
                                         /* (TDP_FLAG_SOFT_DEADLINE)            */
    int32_t moves[TDP_MAX_DEPTH][3];     /* positions i<j<k, in move order    */
    tdp_counters counters;
} tdp_case_result;

/* Launch statistics of the last search or batch on a slot. */
typedef struct tdp_launch_stats {
    double kernel_ms;        /* search kernel time (HIP events)              */
    int32_t waves;           /* waves launched                               */
    int32_t overflowed;      /* queue overflow flag                          */
    uint64_t donations;      /* nodes handed to idle waves                   */
    uint64_t steals;         /* nodes received by idle waves                 */
    uint64_t node_bytes;     /* bytes per node                               */
    uint64_t pops;           /* nodes read from a wave's stack               */
    uint64_t pushes;         /* nodes written to a wave's stack              */
    uint64_t probes;         /* dedup tag words read                         */
    uint64_t key_reads;      /* dedup keys compared                          */
    uint64_t key_writes;     /* dedup entries written                        */
    uint64_t key_bytes;      /* bytes of one stored key                      */
    uint64_t evictions;      /* dedup entries replaced while live (full line) */
    uint64_t spills;         /* nodes moved from full stacks to the spill pool */
    uint64_t spill_takes;    /* nodes taken from the spill pool              */
    uint64_t overflow_info;  /* on overflow: reason << 56 (1 no sharing, 2 pool */
                             /* full) | stack fill << 32 | pool fill           */
    uint64_t phase_ticks[8]; /* profiling builds only: 10 ns ticks summed over */
                             /* waves for pop, checks, cycle analysis, scan,  */
                             /* dedup, push, sharing, root expansion          */
} tdp_launch_stats;

/* search flags */
#define TDP_FLAG_DETERMINISTIC 1u   /* one wave per configuration, no sharing: */
                                    /* counters are reproducible                */
#define TDP_FLAG_SINGLE_WAVE   2u   /* the whole launch runs on one wave        */
#define TDP_FLAG_SOFT_DEADLINE 16u  /* tdp_search_batch: each launch stops after the */
                                    /* tdp_set_launch_deadline budget without an    */
                                    /* error; configurations not found by then get  */
                                    /* depth -1 (undecided) instead of 0            */
#define TDP_FLAG_LEARNED       8u   /* order each node's admitted children by the */
                                    /* move-ordering model (tdp_mlp_*), best first */
#define TDP_FLAG_GRAPH         32u  /* run each launch round (reset, search,      */
                                    /* control-line copy) as the slot's HIP graph */
                                    /* (instantiated once, updated per round)     */
#define TDP_FLAG_RESIDENT      4u   /* tdp_search_batch: reuse the configurations */
                                    /* already in device memory from the previous */
                                    /* batch on this slot (same count); the host  */
                                    /* arrays are not read or uploaded            */

TDP_API const char *tdp_last_error(void);

/* Replaces initGpuSortSlots (slot_manager.cu:62-98) behind JNI initSlots
 * (jni_sorter.cu:25-34).  Slot i lives on device i % devices_count. */
TDP_API int tdp_init_slots(int total_slots, int devices_count, int64_t queue_bytes_budget,
                   int dedup_table_size);

/* Replaces destroyGpuSortSlots (slot_manager.cu:22-60) behind JNI
 * destroySlots (jni_sorter.cu:36-43). */
TDP_API int tdp_destroy_slots(void);

/* Replaces performGpuSortingSearch (sorting_search.cu:701-856) behind JNI
 * searchForSortingSeq (jni_sorter.cu:45-104).  pi/spi: n bytes each.  On
 * TDP_OK, *out_depth is the sequence length (0 = none found) and out_moves
 * holds *out_depth position triples (i,j,k).  Blocks until the search ends or
 * is cancelled. */
TDP_API int tdp_search(int slot, const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles,
               float min_rate, int max_moves, int full_sorting, int32_t *out_moves,
               int32_t *out_depth);

/* tdp_search with flags and counters (extension). */
TDP_API int tdp_search_ex(int slot, const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles,
                  float min_rate, int max_moves, int full_sorting, uint32_t flags,
                  int32_t *out_moves, int32_t *out_depth, tdp_counters *out_counters);

/* Many configurations in one persistent launch on one slot (extension; no
 * reference counterpart).  pis/spis: count x 64 bytes (first n[c] used);
 * count <= 65536. */
TDP_API int tdp_search_batch(int slot, int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns,
                     const int32_t *initial_even_cycles, const float *min_rates,
                     const int32_t *max_moves, const int32_t *full_sorting, uint32_t flags,
                     tdp_case_result *out_results);

/* ---- device-resident serving (extension; replaces the one-launch-per-call
 * path of performGpuSortingSearch, sorting_search.cu:701-856, for the
 * ForkJoin callers of GPUSortingSearch.search, GPUSortingSearch.java:47-79,
 * SortOrExtend.java:96-104).  tdp_serve_start launches one persistent search
 * on the current device that runs until tdp_serve_stop (or lifetime_s
 * seconds): its waves take requests as they arrive, share their work, the
 * queue budget and a dedup table of dedup_table_size entries, and reply
 * through pinned host memory.  While a session runs, tdp_search (and so
 * searchForSortingSeq) on a slot of that device is served by it, and
 * tdp_cancel cancels the slot's running request.  A request is admitted when
 * one of `slots` (<= 65536) device slots is free and the roots of its start
 * node (C(n,3) nodes) fit beside those of the requests in flight within half
 * of the spill pool (a quarter of queue_bytes); slots bounds the requests in
 * flight and so their demand on the shared dedup table; max_n (<= 63) bounds
 * n and sets the node size.  flags:
 * TDP_FLAG_LEARNED orders children by the model set at start.  Verdicts are
 * the plain search's; counters in replies are those of the batches ended
 * before the verdict.  initSlots starts a session itself when TDP_SERVE=1 is
 * in the environment, with the slots' total queue budget and dedup entries
 * and TDP_SERVE_SLOTS (default 256) requests in flight; the slots then hold
 * no queue memory of their own. */
TDP_API int tdp_serve_start(int64_t queue_bytes, int dedup_table_size, int slots, int max_n, uint32_t flags,
                            int lifetime_s);
TDP_API int tdp_serve_stop(void);
/* One request through the running session; blocks until its verdict.  Same
 * arguments and results as tdp_search; thread-safe. */
TDP_API int tdp_serve_search(const uint8_t *pi, const uint8_t *spi, int n, int initial_even_cycles,
                             float min_rate, int max_moves, int full_sorting, int32_t *out_moves,
                             int32_t *out_depth, tdp_counters *out_counters);
/* A batch through the running session from one caller (same layout and
 * results as tdp_search_batch): requests are admitted as slots and queue
 * memory free up while the earlier ones run, and the call returns when every
 * verdict is in. */
TDP_API int tdp_serve_batch(int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns,
                            const int32_t *initial_even_cycles, const float *min_rates, const int32_t *max_moves,
                            const int32_t *full_sorting, tdp_case_result *out_results);
/* Requests completed (slot handed back), slots in use and queue bytes
 * reserved by the requests in flight. */
TDP_API int tdp_serve_stats(uint64_t *completed, uint64_t *in_flight, int64_t *reserved_bytes);

/* Statistics of the last launch on a slot. */
TDP_API int tdp_last_launch_stats(int slot, tdp_launch_stats *out);

/* Replaces cancelGpuSort (slot_manager.cu:100-125) behind JNI cancelSort
 * (jni_sorter.cu:106-109).  Never fails. */
TDP_API void tdp_cancel(int slot);

/* Replaces performGpuPiFilter (pi_filter.cu:157-260) behind JNI
 * filterPiPermutations (jni_sorter.cu:111-150).  Enumerates the (n-1)!
 * arrangements with pi[0]=0 in Lehmer-rank order and keeps those passing the
 * three predicates.  Writes the first min(count, out_rows_cap) passing rows
 * (n bytes each, rank order) to out; *out_total is the uncapped count. */
TDP_API int tdp_filter(const uint8_t *cycle, int cycle_len, const uint8_t *triple, int triple_len, int n,
               int skip_2move_check, uint8_t *out, int64_t out_rows_cap, int64_t *out_total,
               int64_t *out_written);

/* tdp_filter without any allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as tdp_filter, but the outputs are in device memory and
 * the call only enqueues work on `stream` (a hipStream_t; NULL = the null
 * stream) and returns: out_rows_dev receives the first min(count,
 * out_rows_cap) passing rows (n bytes each, rank order), *out_total_dev the
 * uncapped count once the stream reaches that point.  Working memory is a
 * fixed 8.5 MiB device array inside the library (passes of 2^21 groups of ranks, the
 * cycle's 2-move table), one per GPU for the life of the process, so any
 * valid input fits it; calls that overlap on one GPU (another thread or
 * stream) race on that array and must be serialised by the caller. */
TDP_API int tdp_filter_noalloc(const uint8_t *cycle, int cycle_len, const uint8_t *triple, int triple_len, int n,
                               int skip_2move_check, uint8_t *out_rows_dev, int64_t out_rows_cap,
                               uint64_t *out_total_dev, void *stream);

/* Kernel time of the last tdp_filter call on this thread, milliseconds. */
TDP_API double tdp_filter_last_kernel_ms(void);

/* Validation entry (extension): odd-cycle count of each of the C(n,3)
 * children of (pi, spi), computed on the GPU with the search kernel's LDS
 * cycle analysis and O(1) parity delta. out: C(n,3) ints. */
TDP_API int tdp_eval_children(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out);

/* Validation entry (extension, no reference counterpart): the cycle-local
 * enumeration of two nodes a and b (same n, 3..32) that the search kernel
 * runs in one pass when both admit only +2 children, and the same two nodes
 * enumerated one after the other.  goal_x / admit_x: bit 2 set when a +2
 * child of node x is a goal / admissible.  out: 2 records of 4 + 512 ints
 * (status 1 = done, goal node (1 = a, 0 = b, -1 none), goal triple index,
 * staged count, stage entries i | j << 6 | k << 12 | odd cycles << 18 |
 * node << 24); record 0 the paired pass, record 1 the sequential passes. */
TDP_API int tdp_eval_pair(const uint8_t *pi_a, const uint8_t *spi_a, const uint8_t *pi_b, const uint8_t *spi_b,
                          int n, int e_a, int e_b, uint32_t goal_a, uint32_t admit_a, uint32_t goal_b,
                          uint32_t admit_b, int32_t *out);

/* Validation entry (extension): the learned ordering's 128 feature bits of
 * each of the C(n,3) moves of the node (pi, spi) at depth `depth` of a search
 * to max_depth with start odd-cycle count e0, as the search kernel computes
 * them.  out: C(n,3) x 4 words, lexicographic triple order. */
TDP_API int tdp_eval_features(const uint8_t *pi, const uint8_t *spi, int n, int e0, int depth, int max_depth,
                              uint32_t *out);

/* Host-only helpers (no GPU): the same parity delta on the CPU, and the
 * predicate tables the kernel uses. */
TDP_API int tdp_host_eval_children(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out);
TDP_API int tdp_host_predicate_tables(int n, int initial_even_cycles, float min_rate, int max_depth,
                              int full_sorting, uint64_t *accept16, uint64_t *reach16);

/* Work-sharing launch tuning for later launches on every slot (extension;
 * no reference counterpart).  0 keeps a value's default: waves per CU = the
 * kernel's occupancy limit, donation check every 16 expansions, 8 nodes per
 * donation, idle sleep up to 16 x 127 s_sleep units, at most 1024
 * configurations per persistent launch (a larger batch runs as consecutive
 * launches).  share_interval must be a power of two. */
TDP_API int tdp_set_tuning(int waves_per_cu, int share_interval, int share_max, int idle_backoff,
                           int max_active);

/* Nodes a wave of the work-sharing mode takes from its stack and expands
 * before one dedup pass over all their children (1..8; 0 = default 8; the
 * deterministic modes always take one).  Extension;
 * no reference counterpart. */
TDP_API int tdp_set_expansion_batch(int nodes);

/* ---- learned move ordering (extension; no reference counterpart) ----
 * A 128 -> 256 (ReLU) -> 1 MLP scores the candidate moves of a node from 128
 * feature bits (tdp_learn.h); searches with TDP_FLAG_LEARNED push the admitted
 * children of every node in ascending score order, so the best-scored child
 * is taken first.  Verdicts do not depend on the order.  Weights are bf16 on
 * the device (fp32 master copy for training), one model per process.
 * Layout: w1[k * 256 + j] (input k, hidden j), b1[256], w2[256], b2.
 * Feature rows are 4 uint32 words (bit b of the row = word b/32, bit b%32). */
TDP_API int tdp_mlp_set_weights(const float *w1, const float *b1, const float *w2, float b2);
TDP_API int tdp_mlp_get_weights(float *w1, float *b1, float *w2, float *b2);
TDP_API int tdp_mlp_init(uint64_t seed);
/* logits of N rows */
TDP_API int tdp_mlp_forward(const uint32_t *rows, int N, float *logits);
/* Loss and gradients (flat [w1 | b1 | w2 | b2], 33,281 floats) of the softmax
 * cross-entropy of each group's logits against its one labelled row, mean over
 * the G groups; group g holds rows group_start[g] .. group_start[g+1]-1. */
TDP_API int tdp_mlp_gradients(const uint32_t *rows, const uint8_t *labels, const int32_t *group_start, int G,
                              float *grads, float *loss);
/* Adam training on the GPU: `epochs` passes over the groups in random order
 * (seed), batches of groups_per_batch groups; epoch_loss[e] = mean loss of
 * epoch e (may be NULL). */
TDP_API int tdp_mlp_train(const uint32_t *rows, const uint8_t *labels, const int32_t *group_start, int G,
                          int epochs, float lr, int groups_per_batch, uint64_t seed, float *epoch_loss);

/* Time budget per persistent launch for TDP_FLAG_SOFT_DEADLINE (extension;
 * default 1000 ms). */
TDP_API int tdp_set_launch_deadline(int milliseconds);

/* Canonical form of `count` configurations (extension; replaces
 * Configuration.getSignature / getEquivalentSignatures / getCanonical /
 * hashCode, Configuration.java:53-198,243-297, for the Java caller's case
 * keys).  Configuration c has ns[c] <= 64 symbols, spi as a one-line array
 * spis[c*64 + x] = image of x (no fixed points) and pi as the symbols by
 * position pis[c*64 + i].  Writes the least of its 2n equivalent signatures
 * by CRC32 hash then content (out_signatures[c*64 + i], float), that hash
 * (= Configuration.hashCode), the index 2*rotation + mirror of the
 * equivalent signature it is, and the configuration it stands for
 * (ofSignature, pi the identity) as a one-line spi. */
TDP_API int tdp_canonical_batch(int count, const uint8_t *spis, const uint8_t *pis, const int32_t *ns,
                                float *out_signatures, int32_t *out_hashes, int32_t *out_index,
                                uint8_t *out_canonical_spi);

/* Allocation retry timeout (replaces gGpuMemoryRetryTimeoutMs,
 * cuda_memory_pressure.cuh:10, read by allocateDeviceBuffer,
 * cuda_memory.cuh:72): a device allocation that fails for lack of memory is
 * retried with backoff (0.1 s doubling to 5 s) until this many milliseconds
 * have passed, then fails with TDP_ERR_OUT_OF_MEMORY and "HIP out of memory
 * while allocating N bytes for <label>" (cuda_memory.cuh:82-85).  Default
 * 60000; a negative value acts as 0 (no retry). */
TDP_API int tdp_set_alloc_timeout(int64_t milliseconds);
/* Allocation retries so far in this process (gGpuMemoryRetries,
 * cuda_memory_pressure.cuh:9). */
TDP_API uint64_t tdp_alloc_retries(void);

/* Number of visible GPUs (0 when none). */
TDP_API int tdp_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
