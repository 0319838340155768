/*
 * CPU oracle for the tdp1375 proof-search path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline.  The product
 * (tdp1375_amd) never links or calls it.
 *
 * Every function restates the reference's algorithm from its source text
 * (citations are relative to the reference repository root):
 *   search      src/native/backtracking/sorting_search.cu:230-615 (Q=1 order)
 *   predicates  src/native/backtracking/sorting_search.cu:321-362
 *   moves       src/native/backtracking/transposition.cuh:5-19, common.cuh:25-44
 *   full sort   src/main/java/br/unb/cic/tdp/base/SortingSearch.java:17-104
 *   filter      src/native/backtracking/pi_filter.cu:38-154
 *   replay      src/main/java/br/unb/cic/tdp/base/GPUSortingSearch.java:65-87
 *
 * Parity pins (which check fixes which function) are listed in the header of
 * tdp_oracle.c.
 */
#ifndef TDP_ORACLE_H
#define TDP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_N 63
#define OR_MAX_DEPTH 15

typedef struct {
    int32_t found;        /* 1 when an accepted sequence was found            */
    int32_t depth;        /* its length                                       */
    int32_t moves[OR_MAX_DEPTH][3]; /* positions i<j<k per move (reference   */
                                    /* PackedMove order)                      */
    int32_t aborted;      /* 1 when max_expansions was hit (verdict unknown)  */
    int32_t status;       /* 0 ok, <0 invalid argument                        */
    uint64_t expansions;  /* popped nodes with depth<D && can_still_reach     */
    uint64_t generated;   /* children evaluated (T per expansion)             */
    uint64_t admitted;    /* non-root children pushed                         */
    uint64_t dominated;   /* non-root children rejected by dedup              */
    uint64_t improved;    /* admitted because they lowered a stored depth     */
    uint64_t inserted;    /* admitted as a new dedup entry                    */
    uint64_t roots_live;  /* roots that are goals or pass can_still_reach     */
    uint64_t pops;        /* all popped nodes                                 */
} or_result;

int or_count_odd_cycles(const uint8_t *spi, int n);
int or_satisfies_rate(int even_cycles, int depth, int n, int e0, float min_rate, int full);
int or_can_still_reach(int even_cycles, int depth, int n, int e0, float min_rate,
                       int max_depth, int full);
int or_triple_count(int n);
/* lexicographic (i<j<k) triple t of n -> positions */
void or_triple_unrank(int n, int t, int32_t *i, int32_t *j, int32_t *k);
/* apply the move at positions i<j<k: spi' = spi o (a b c)^-1, pi' = block swap */
void or_apply_move(uint8_t *pi, uint8_t *spi, int n, int i, int j, int k);

/* odd-cycle count of each child (lexicographic triples), O(n) recount each */
int or_children_even(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out);

/* Deterministic single-queue search (reference with Q=1).  use_dedup=0 runs
 * with no table (reference dedupTable == nullptr).  max_expansions=0: no
 * limit. */
int or_search(const uint8_t *pi, const uint8_t *spi, int n, int e0, float min_rate,
              int max_depth, int full, int use_dedup, uint64_t max_expansions,
              or_result *out);

/* or_search with flags.  OR_FLAG_ARC_CHILDREN evaluates each child's
 * odd-cycle count by re-linking the arcs of the three touched cycles
 * (child_even_arcs) instead of recounting the child's cycles: same results,
 * same counters (checked against the recount on whole workloads), several
 * times faster.  It exists for the verdict fixtures of large workloads. */
#define OR_FLAG_ARC_CHILDREN 1
int or_search_ex(const uint8_t *pi, const uint8_t *spi, int n, int e0, float min_rate,
                 int max_depth, int full, int use_dedup, uint64_t max_expansions, int flags,
                 or_result *out);

/* Many independent searches over `threads` host threads (CPU baseline).
 * pis/spis are count x 64 bytes. Returns wall seconds. */
double or_search_many(int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns,
                      const int32_t *e0s, const float *min_rates, const int32_t *max_depths,
                      const int32_t *fulls, int use_dedup, uint64_t max_expansions,
                      int threads, or_result *outs);

double or_search_many_ex(int count, const uint8_t *pis, const uint8_t *spis, const int32_t *ns,
                         const int32_t *e0s, const float *min_rates, const int32_t *max_depths,
                         const int32_t *fulls, int use_dedup, uint64_t max_expansions, int flags,
                         int threads, or_result *outs);

/* odd-cycle count of each child by the arc re-linking of OR_FLAG_ARC_CHILDREN */
int or_children_even_arcs(const uint8_t *pi, const uint8_t *spi, int n, int32_t *out);

/* SortingSearch.java (fullSorting CPU search, no dedup).  Writes symbol
 * triples (a,b,c); returns the depth found (0 = none).  float_prune=0 is the
 * Java arithmetic (double product in the prune bound); 1 rounds that product
 * to float32 as the device kernel does.  Not thread-safe. */
int or_java_full_sort_search(const uint8_t *pi, const uint8_t *spi, int n, int e0,
                             float min_rate, int max_moves, int float_prune,
                             int32_t *out_symbol_moves);

/* Brute-force verdict by breadth-first search over spi states (n <= 9). */
int or_bfs_verdict(const uint8_t *pi, const uint8_t *spi, int n, int e0, float min_rate,
                   int max_depth, int full);

/* Replays position moves; returns the final odd-cycle count or -1 if a move
 * is not i<j<k<n. pi/spi are updated in place. */
int or_replay_positions(uint8_t *pi, uint8_t *spi, int n, const int32_t *moves, int depth);

/* Replays symbol 3-cycles (a b c) (any rotation) on linear pi / spi; returns
 * the final odd-cycle count or -1 if a move is not a transposition of pi. */
int or_replay_symbols(uint8_t *pi, uint8_t *spi, int n, const int32_t *moves, int depth);

/* pi filter: all (n-1)! arrangements with pi[0]=0 in Lehmer-rank order.
 * Writes the first out_cap passing permutations (n bytes each); returns the
 * uncapped count, or -1 on invalid arguments. */
int64_t or_filter(const uint8_t *cycle, int cycle_len, const uint8_t *triple, int triple_len,
                  int n, int skip_2move_check, uint8_t *out, int64_t out_cap, int threads,
                  int64_t *written);

/* predicate of one permutation (used by tests on hand-made arrangements) */
int or_filter_accepts(const uint8_t *pi, int n, const uint8_t *cycle, int cycle_len,
                      const uint8_t *triple, int triple_len, int skip_2move_check);

#ifdef __cplusplus
}
#endif
#endif
