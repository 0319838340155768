// Shared host/device definitions for the MI355X (gfx950) proof-search path.
//
// Vocabulary follows the reference: "even cycles" = number of odd-length
// cycles of spi (sorting_search.cu:302-317, Cycle.java:149-151); a move is a
// triple of positions i<j<k of the linear pi (common.cuh:4-44).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define TDP_HD __host__ __device__ __forceinline__
#else
#define TDP_HD inline
#endif

namespace tdp {

constexpr int kMaxN = 63;       // common.cuh:14-16
constexpr int kMaxDepth = 15;   // common.cuh:10-12
constexpr int kWave = 64;

// C(n,3) (common.cuh:25-30)
TDP_HD int triple_count(int n) { return n < 3 ? 0 : n * (n - 1) * (n - 2) / 6; }

// Triples are stored packed as i | j << 8 | k << 16 in lexicographic order
// (the order of sorting_search.cu:752-758).
TDP_HD uint32_t pack_triple(int i, int j, int k) {
    return (uint32_t)i | ((uint32_t)j << 8) | ((uint32_t)k << 16);
}

// Reference PackedMove (common.cuh:32-37): i | j << 6 | k << 12.
TDP_HD uint32_t pack_move(int i, int j, int k) {
    return (uint32_t)(i & 63) | ((uint32_t)(j & 63) << 6) | ((uint32_t)(k & 63) << 12);
}

// Per-symbol cycle descriptor used by the O(1) parity delta:
//   bits 0..5   id of the symbol's cycle (its smallest symbol)
//   bits 8..13  position of the symbol along the cycle, counted from the id
//   bits 16..22 cycle length
TDP_HD uint32_t cycle_info(int root, int pos, int len) {
    return (uint32_t)root | ((uint32_t)pos << 8) | ((uint32_t)len << 16);
}

// Change of the odd-length-cycle count when the move with symbols
// a=pi[i], b=pi[j], c=pi[k] is applied (spi' = spi o (a b c)^-1,
// sorting_search.cu:547-557).  Derived from how the three arcs of the cycle
// structure are re-linked (a->spi(c), b->spi(a), c->spi(b)):
//   * a,b,c in three cycles: the cycles merge into one;
//   * a,b,c in one cycle of length L: if a->b->c follows the cycle the cycle
//     splits into arcs of lengths k_ab, k_bc, k_ca, otherwise it stays whole;
//   * two symbols x<y (in triple order) of one cycle (length L) and the third
//     in another cycle (length L'): two cycles of lengths {k, L+L'-k}, with
//     k = k_ab or k_bc, or {L-k_ac, L'+k_ac} for the pair (a,c).
// k_xy is the number of spi-steps from x to y.  Checked exhaustively against
// an O(n) recount on every triple of random permutations in the tests.
TDP_HD int parity_delta(uint32_t ia, uint32_t ib, uint32_t ic) {
    const int ra = ia & 63, rb = ib & 63, rc = ic & 63;
    const int pa = (ia >> 8) & 63, pb = (ib >> 8) & 63, pc = (ic >> 8) & 63;
    const int la = (ia >> 16) & 127, lb = (ib >> 16) & 127, lc = (ic >> 16) & 127;
    const bool ab = ra == rb, bc = rb == rc, ac = ra == rc;
    if (ab && bc) {
        const int L = la;
        int kab = pb - pa; kab += (kab < 0) ? L : 0;
        int kbc = pc - pb; kbc += (kbc < 0) ? L : 0;
        int kca = pa - pc; kca += (kca < 0) ? L : 0;
        if (kab + kbc + kca != L) return 0;
        return (kab & 1) + (kbc & 1) + (kca & 1) - (L & 1);
    }
    if (!ab && !bc && !ac) {
        return ((la + lb + lc) & 1) - (la & 1) - (lb & 1) - (lc & 1);
    }
    int L, Lo, x, y;
    if (ab) { L = la; Lo = lc; x = pa; y = pb; }
    else if (bc) { L = lb; Lo = la; x = pb; y = pc; }
    else { L = la; Lo = lb; x = pa; y = pc; }
    int k = y - x; k += (k < 0) ? L : 0;
    int after;
    if (ac) after = ((L - k) & 1) + ((Lo + k) & 1);
    else after = (k & 1) + ((L + Lo - k) & 1);
    return after - (L & 1) - (Lo & 1);
}

// One search configuration of a launch (uploaded once per launch).
struct alignas(16) ConfigDesc {
    uint32_t n;
    uint32_t max_depth;
    uint32_t e_start;        // odd-cycle count of spi0 (start node)
    uint32_t triple_count;
    uint32_t reserved0;
    uint32_t flags;          // bit 0: full sorting
    uint32_t pad[2];
    uint64_t accept[16];     // bit e of accept[d]: satisfies_rate(e, d)
    uint64_t reach[16];      // bit e of reach[d]: can_still_reach(e, d)
    uint8_t spi0[64];
    uint8_t pi0[64];
};

enum CounterIndex {
    kCntExpansions = 0,  // popped with depth < D and can_still_reach
    kCntGenerated,       // children evaluated (C(n,3) per expansion)
    kCntAdmitted,        // non-root children pushed
    kCntDominated,       // non-root children rejected by the dedup table
    kCntImproved,        // admitted because they lowered a stored depth
    kCntInserted,        // admitted as a new table entry
    kCntUninserted,      // admitted without a table entry (probe limit)
    kCntRootsLive,       // roots that are goals or pass can_still_reach
    kNumCounters
};

struct alignas(16) ConfigResult {
    int32_t found;
    int32_t depth;
    int32_t pad[2];
    uint32_t moves[16];  // triple indices of the solution, in move order
    unsigned long long counters[kNumCounters];
    // nodes of the configuration pushed minus nodes taken (mod 2^64): zero
    // once its search space is exhausted, whatever stopped the launch
    unsigned long long live;
    unsigned long long pad2;
};

// Node layout in HBM (32-bit words):
//   w0        configuration index
//   w1        depth | even << 8
//   w2..w9    moves as u16 triple indices (16 slots)
//   w10..     spi bytes (spi_words words), then pi bytes (spi_words words)
constexpr int kNodeHeaderWords = 10;
TDP_HD int spi_words_for(int max_n) { return (max_n + 3) / 4; }
TDP_HD int node_words_for(int max_n) {
    int w = kNodeHeaderWords + 2 * spi_words_for(max_n);
    return (w + 3) & ~3;  // 16-byte stride
}

// Dedup tag word: epoch << 48 | configuration << 32 | hash27 << 5 |
// ready << 4 | depth.  An entry whose epoch is not the launch's is free; the
// configuration field limits a launch to 65,536 configurations.
constexpr int kDedupBucket = 8;
constexpr unsigned long long kTagReady = 16ULL;
TDP_HD unsigned long long dedup_ident(uint32_t epoch, uint32_t cfg, uint64_t hash) {
    return ((unsigned long long)(epoch & 0xFFFFu) << 48) | ((unsigned long long)(cfg & 0xFFFFu) << 32) |
           ((unsigned long long)(hash & 0x7FFFFFFu) << 5);
}
constexpr int kMaxBatchConfigs = 65536;

// Launch-wide control block; hot words sit on their own 128-byte lines.
struct alignas(128) ControlLine {
    unsigned long long v;
    unsigned long long pad[15];
};
enum ControlIndex {
    kCtlNextConfig = 0,
    kCtlIdle,
    kCtlRegistered,
    kCtlPoolHead,      // spill pool: nodes taken
    kCtlPoolTail,      // spill pool: slots reserved
    kCtlSpills,        // nodes moved to the spill pool
    kCtlStop,          // 1 overflow
    kCtlResolved,      // configurations found
    kCtlDonations,
    kCtlSteals,
    kCtlPops,          // nodes read from a stack (including discarded ones)
    kCtlPushes,        // nodes written to a stack
    kCtlProbes,        // dedup tag words read
    kCtlKeyReads,      // dedup entries whose stored key was compared
    kCtlKeyWrites,     // dedup entries written
    kCtlEvictions,     // dedup entries replaced while live (bucket full)
    kCtlSpillTakes,    // nodes taken from the spill pool
    kCtlOverflowInfo,  // why the overflow flag was raised: reason << 56 | stack fill << 32 | pool fill
    kCtlPublished,     // serving: requests copied to device slots by the feeder
    kCtlPhase0,        // per-phase time of profiling builds (kNumPhases lines)
    kNumControl = kCtlPhase0 + 8
};

// Phases timed by profiling builds of the search kernel.
enum PhaseIndex {
    kPhasePop = 0,      // stack pop into LDS
    kPhaseCheck,        // stop / found checks and config switch
    kPhaseAnalyse,      // cycle analysis of the node
    kPhaseScan,         // child evaluation over all triples
    kPhaseDedup,        // dedup of staged children
    kPhasePush,         // pushes of admitted children (and their learned scores)
    kPhaseShare,        // donation, idle polling, inbox receipt and pool takes
    kPhaseRoot,         // root expansions
    kNumPhases
};

// ---- device-resident serving (serve_kernel)
// A request as the host publishes it in its pinned request ring: the search
// descriptor, the device slot it is to run in (chosen by the host among the
// drained slots), and seq = request index + 1 written last.
struct alignas(16) ServeRequest {
    ConfigDesc desc;
    uint32_t slot;
    uint32_t pad0;
    unsigned long long seq;
};

// The reply of a slot's current request in pinned host memory.  seq =
// request index + 1 once the verdict is in (found: as soon as a wave finds
// it; not found / cancelled: once the request's last node is gone), drained
// = request index + 1 once no node of the request is left anywhere, after
// which the host may give the slot to another request.
struct alignas(16) ServeReply {
    int32_t found;           // 1 found, 0 not found, 2 cancelled
    int32_t depth;
    uint32_t moves[16];      // triple indices of the sequence, in move order
    unsigned long long counters[kNumCounters];   // so far (found) / final (not found)
    unsigned long long seq;
    unsigned long long drained;
};

// Host control words of a serving session (pinned host memory).
enum ServeHostIndex {
    kServeHostStop = 0,      // set by the host: the session ends
    kServeHostConsumed = 1,  // written by the feeder: requests copied so far
    kServeHostCancels = 2,   // set by the host: cancellations published so far
    kServeHostCancelRing = 16,   // kServeCancelEntries words: request << 16 | slot
    kServeHostWords = 16 + 256
};
constexpr int kServeCancelEntries = 256;

// slot_state word of a serving slot: generation << 2 | done (bit 0: the
// request is found, cancelled or drained, so its dedup entries may go).
constexpr unsigned long long kSlotDone = 1ULL;

}  // namespace tdp
