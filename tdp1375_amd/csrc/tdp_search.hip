// Persistent wave64 proof-search kernel for gfx950.
//
// Replaces sorting_search_persistent_kernel + initialize_queues_kernel
// (reference src/native/backtracking/sorting_search.cu:377-615).  Semantics
// kept from the reference:
//   * roots are all C(n,3) children of the start node, admitted at depth 1
//     without pruning and goal-checked when popped (sorting_search.cu:403-441);
//   * a popped node is accepted when satisfies_rate holds, expanded when
//     depth < D and can_still_reach (:519-527);
//   * children are goal-checked on generation; the reported goal child is the
//     first one in descending triple order (:533-535, :580-592);
//   * non-goal children with depth < D passing can_still_reach go through the
//     dedup table (dominated -> dropped; improved or new -> pushed) (:559-577);
//   * pushes happen in descending triple order so the LIFO pop order ascends.
//
// MI355X design (not a translation):
//   * one 64-lane wave per worker, each with a private node stack in HBM
//     (no locks on the hot path) and a shared MPMC pool fed by donation when
//     other waves are idle; termination by idle/registered counters;
//   * the node's cycle structure is analysed once per expansion with
//     ds_bpermute pointer jumping, so each child's odd-cycle count is an O(1)
//     parity delta of LDS lookups instead of an O(n) recount;
//   * predicates are host-built bit tables (accept[d], reach[d] over e);
//   * children passing the prune are staged in LDS with ballot compaction and
//     deduplicated lane-parallel in an open-addressing table whose entries are
//     claimed by a 64-bit atomicCAS on (epoch | hash) and hold the full packed
//     spi, so the table is exact; epochs replace the per-search memset;
//   * many configurations run in one persistent launch (batch), keyed by a
//     configuration index in every node and table entry.
#include <hip/hip_runtime.h>

#include "tdp_common.h"
#include "tdp_search.h"
#include "tdp_learn.h"

namespace tdp {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kStage = kStageEntries;
constexpr int kBatchMax = kExpansionBatchMax;
// Rounds of one dedup lookup before the child is admitted without an entry
// (a round is lost only to a concurrent update of the same bucket).
constexpr int kMaxDedupRounds = 16;

enum DedupOutcome { kDedupPending = -1, kDedupDominated = 0, kDedupImproved = 1,
                    kDedupInserted = 2, kDedupUninserted = 3 };

// Per-wave LDS slice.  The 64-bit counters live here rather than in
// registers (lane 0 updates them with wave-uniform values).
struct WaveLds {
    uint32_t nodes[kNodePoolWords];          // the nodes being expanded, node_words apart (node 0 alone
                                             // outside batches)
    uint32_t posinfo[kLearnedBatchMax][64];  // cycle info per position of each node of a learned batch
                                             // (row 0 only for the plain ordering)
    union {
        struct {
            uint32_t memb[64];               // cycle members in spi order, cycles of length >= 2:
                                             // cycle info | position in pi << 24
            uint32_t ememb[64];              // the members of the even cycles again, contiguous
        };
        uint32_t twin[2 * kWave];            // dedup of a chunk of staged children (the members are
                                             // dead once the batch's children are staged)
    };
    uint32_t stage[kStage];
    uint8_t rowsrc[64];                      // learned ordering: lane of each compacted admitted child
    float carry_score[kWave];                // learned ordering: the best-scored admitted children of
    uint32_t carry_st[kWave];                // the batch so far (score, stage entry), pushed at its end
    uint32_t ncarry;
    uint32_t ckey[16][kWave];                // child spi words of each lane's staged child
    unsigned long long cnt[kNumCounters];    // counters of the current configuration
    unsigned long long pend[kNumCounters];   // counters of the current expansion
    unsigned long long live;                 // pushes - pops of the current configuration (serving:
                                             // pushes not yet added to the request's count)
    unsigned long long takes;                // serving: nodes taken, not yet subtracted
    unsigned long long traffic[6];           // pops, pushes, tag reads, key compares, key writes,
                                             // evictions
    unsigned long long phase[kNumPhases];    // 10 ns ticks per phase (TDP_PHASE_TIMING builds)
};

// Node b of the batch in the wave's LDS pool (nodes of sw spi words are
// node_words_for(4 sw) words apart).
__device__ __forceinline__ uint32_t *node_at(WaveLds &L, int b, int sw) {
    return L.nodes + b * ((kNodeHeaderWords + 2 * sw + 3) & ~3);
}
__device__ __forceinline__ const uint32_t *node_at(const WaveLds &L, int b, int sw) {
    return L.nodes + b * ((kNodeHeaderWords + 2 * sw + 3) & ~3);
}

// An LDS atomic add whose result is unused (ds_add_u64 without return), so
// the wave does not wait for the read of a read-modify-write.
__device__ __forceinline__ void lds_add(unsigned long long *p, unsigned long long v) {
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Optional per-phase wall time (s_memrealtime, 100 MHz) for profiling builds
// (-DTDP_PHASE_TIMING); it adds timing reads only and changes no work.
#ifdef TDP_PHASE_TIMING
#define PHASE_T0() const unsigned long long phase_t0_ = __builtin_amdgcn_s_memrealtime()
#define PHASE_ADD(L, k) lds_add(&(L).phase[k], __builtin_amdgcn_s_memrealtime() - phase_t0_)
#else
#define PHASE_T0() (void)0
#define PHASE_ADD(L, k) (void)0
#endif

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Lane 0's value as a wave-uniform (scalar) value: the broadcast of what lane
// 0 alone computed, which keeps values derived from it (stack bounds, loop
// counts, masks) in scalar registers.
__device__ __forceinline__ int lane0(int v) { return __builtin_amdgcn_readlane(v, 0); }
__device__ __forceinline__ unsigned long long lane0_64(unsigned long long v) {
    return ((unsigned long long)(uint32_t)lane0((int)(v >> 32)) << 32) | (uint32_t)lane0((int)(uint32_t)v);
}
// Fields of the wave's LDS descriptor copy as wave-uniform (scalar) values.
__device__ __forceinline__ int cfg_int(const uint32_t &f) { return uniform((int)f); }
__device__ __forceinline__ uint64_t cfg_bits(const uint64_t &f) {
    return ((uint64_t)(uint32_t)uniform((int)(f >> 32)) << 32) | (uint32_t)uniform((int)(uint32_t)f);
}

// Inclusive prefix sum / running maximum over the 64 lanes with DPP row
// shifts and row broadcasts (a few VALU cycles each, instead of LDS-latency
// cross-lane shuffles); lane 63 holds the total / maximum.
__device__ __forceinline__ int wave_scan_add(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x111, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x112, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x114, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x118, 0xF, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x142, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp(-1, v, 0x143, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Cross-wave data (dedup entries, pool nodes, flags) is only ever accessed
// with relaxed device-scope atomics (sc1 loads and stores), which are
// coherent per location across the XCDs' L2 caches.  Publication order comes
// from waiting, in this wave, for every outstanding vector memory access to
// complete before the flag access is issued (s_waitcnt vmcnt(0); gfx9 counts
// stores in vmcnt too), so no L2 writeback or invalidate is ever needed.  A
// workgroup-scope fence is not enough: on gfx950 it emits no vmcnt wait, and
// two sc1 stores to different channels then become visible in either order.
// The "memory" clobber keeps the compiler from moving accesses across it.
__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void vm_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// payload accesses complete before the following publishing store
__device__ __forceinline__ void complete_before() { vm_complete(); }
// the preceding load completes before any following access is issued
__device__ __forceinline__ void complete_after() { vm_complete(); }
__device__ __forceinline__ unsigned long long fetch_add(unsigned long long *p, unsigned long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas64(unsigned long long *p, unsigned long long expected,
                                      unsigned long long desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pinned host memory of a serving session (request ring, replies, control
// words) is accessed with system-scope atomics, which bypass the GPU caches.
__device__ __forceinline__ unsigned long long ld_sys(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_sys32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ unsigned long long ctl_load(const SearchParams &P, int idx) {
    return ld_relaxed(&P.ctrl[idx].v);
}
// A control word read by lane 0 and broadcast (wave-uniform).
__device__ __forceinline__ unsigned long long ctl_load_uniform(const SearchParams &P, int idx) {
    unsigned long long v = 0;
    if ((threadIdx.x & 63) == 0) v = ctl_load(P, idx);
    return lane0_64(v);
}

__device__ __forceinline__ uint64_t fmix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

__device__ __forceinline__ uint32_t put_byte(uint32_t w, int byte, uint32_t v) {
    const int sh = (byte & 3) * 8;
    return (w & ~(0xFFu << sh)) | (v << sh);
}

// Word q of the child spi: parent word with spi'[a]=spi[c], spi'[b]=spi[a],
// spi'[c]=spi[b] (sorting_search.cu:551-555).
struct ChildMove {
    int a, b, c;     // symbols at positions i, j, k
    uint32_t na, nb, nc;  // new images of a, b, c
};

__device__ __forceinline__ uint32_t child_spi_word(const uint32_t *pw, int q, const ChildMove &m) {
    uint32_t w = pw[q];
    w = ((m.a >> 2) == q) ? put_byte(w, m.a, m.na) : w;
    w = ((m.b >> 2) == q) ? put_byte(w, m.b, m.nb) : w;
    w = ((m.c >> 2) == q) ? put_byte(w, m.c, m.nc) : w;
    return w;
}

// Triple enumeration without memory traffic.  In descending lexicographic
// order the triples with first index i = n-1-m occupy offsets
// [C(m,3), C(m+1,3)), and inside that block the pairs (j, k) follow the colex
// order of (u, v) = (n-1-k, n-1-j), the same sequence for every i and n.  One
// table of the C(63,2) colex pairs (u | v << 8) per workgroup in LDS therefore
// yields any triple from its offset o; the lexicographic index is
// t = C(n,3) - 1 - o.
constexpr int kColexPairs = 1953;   // C(63, 2)
// Most candidates of the cycle-local enumeration (more: the full scan runs)
constexpr int kCycleLocalMax = 512;

__device__ __forceinline__ int choose3(int m) { return m * (m - 1) * (m - 2) / 6; }
__device__ __forceinline__ int choose2(int v) { return v * (v - 1) / 2; }

__device__ __forceinline__ int lex_index(int n, int T, int i, int j, int k) {
    return T - 1 - choose3(n - 1 - i) - choose2(n - 1 - j) - (n - 1 - k);
}

// Triple at descending offset o; m is the caller's block cursor, advanced to
// the block holding o (offsets only grow along a lane's scan).
__device__ __forceinline__ void triple_at(const uint16_t *colex, int n, int &m, int o, bool valid, int &i,
                                          int &j, int &k) {
    while (m + 1 < n && choose3(m + 1) <= o) ++m;
    const uint32_t pair = colex[valid ? o - choose3(m) : 0];
    i = n - 1 - m;
    j = n - 1 - (int)(pair >> 8);
    k = n - 1 - (int)(pair & 0xFF);
}

// Parity delta of one child from the posinfo of its three symbols: three
// different cycles merge into one (delta from the lengths only); the cases
// where a cycle is hit twice take the full rule.
__device__ __forceinline__ bool hits_a_cycle_twice(uint32_t ia, uint32_t ib, uint32_t ic) {
    return ((ia ^ ib) & 63u) == 0 || ((ib ^ ic) & 63u) == 0 || ((ia ^ ic) & 63u) == 0;
}

__device__ __forceinline__ int child_delta(uint32_t ia, uint32_t ib, uint32_t ic) {
    // three cycles merge into one: -2 when at least two of them are odd, else 0
    const uint32_t odd = ((ia >> 16) & 1u) + ((ib >> 16) & 1u) + ((ic >> 16) & 1u);
    int delta = odd >= 2u ? -2 : 0;
    if (hits_a_cycle_twice(ia, ib, ic)) delta = parity_delta(ia, ib, ic);
    return delta;
}

// Enumerates the children in descending lexicographic order, one chunk of
// 64 consecutive offsets per call of visit(valid, i, j, k, obase) (lane l
// holds offset obase + l).  Blocks m <= 11 (first index i = n-1-m) hold fewer
// than 64 pairs each, so the first C(12,3) = 220 offsets run with a per-lane
// block cursor; larger blocks run one wave-uniform block at a time with lanes
// over its colex pairs.  visit returns kScanGo to continue, kScanStop to end
// the scan, kScanPause to end it with (mb, qb) set to the next chunk so that
// a later call resumes there (mb < 0: inside the first 220 offsets, qb the
// next offset).  Pausing lets the caller flush the stage outside the scan, so
// the scan's and the flush's registers are never live together.
enum { kScanGo = 0, kScanStop = 1, kScanPause = 2 };

template <typename Visit>
__device__ __forceinline__ int scan_children(const uint16_t *colex, int n, int T, int lane, int &mb, int &qb,
                                             Visit visit) {
    if (mb < 0) {
        const int prefix = T < 220 ? T : 220;
        int m = 2;
        for (int obase = qb; obase < prefix; obase += kWave) {
            const int o = obase + lane;
            const bool valid = o < prefix;
            int i, j, k;
            triple_at(colex, n, m, o, valid, i, j, k);
            const int r = visit(valid, i, j, k, obase);
            if (r == kScanStop) return r;
            if (r == kScanPause) {
                qb = obase + kWave;
                return r;
            }
        }
        mb = 12;
        qb = 0;
    }
    for (; mb < n; ++mb, qb = 0) {
        const int pairs = choose2(mb), o0 = choose3(mb), i = n - 1 - mb;
        for (; qb < pairs; qb += kWave) {
            const int q = qb + lane;
            const bool valid = q < pairs;
            const uint32_t pair = colex[valid ? q : 0];
            const int r = visit(valid, i, n - 1 - (int)(pair >> 8), n - 1 - (int)(pair & 0xFF), o0 + qb);
            if (r == kScanStop) return r;
            if (r == kScanPause) {
                qb += kWave;
                if (qb >= pairs) { ++mb; qb = 0; }
                return r;
            }
        }
    }
    return kScanGo;
}

__device__ __forceinline__ void fill_colex(uint16_t *colex, uint32_t *colex3 = nullptr) {
    // q -> (u, v) with v the largest such that C(v,2) <= q
    for (int q = threadIdx.x; q < kColexPairs; q += blockDim.x) {
        int v = 1;
        while (choose2(v + 1) <= q) ++v;
        colex[q] = (uint16_t)((q - choose2(v)) | (v << 8));
    }
    // r -> (p0, p1, p2), p0 < p1 < p2, with C(p2,3) + C(p1,2) + p0 = r
    if (colex3) {
        for (int r = threadIdx.x; r < kCycleLocalMax; r += blockDim.x) {
            int p2 = 2;
            while (choose3(p2 + 1) <= r) ++p2;
            const int r1 = r - choose3(p2);
            int p1 = 1;
            while (choose2(p1 + 1) <= r1) ++p1;
            colex3[r] = (uint32_t)(r1 - choose2(p1)) | ((uint32_t)p1 << 8) | ((uint32_t)p2 << 16);
        }
    }
    __syncthreads();
}

// Stage entries: i | j << 6 | k << 12 | even << 18 | parent node << 24.
__device__ __forceinline__ uint32_t stage_entry(int i, int j, int k, int even, int b) {
    return (uint32_t)i | ((uint32_t)j << 6) | ((uint32_t)k << 12) | ((uint32_t)even << 18) | ((uint32_t)b << 24);
}

// transposition.cuh:5-11
__device__ __forceinline__ int transposed_src_index(int p, int i, int j, int k) {
    const int lenA = j - i, lenB = k - j;
    const int inLow = (p >= i) & (p < i + lenB);
    const int inHigh = (p >= i + lenB) & (p < k);
    return p + inLow * lenA - inHigh * lenB;
}

// Lane s handles symbol s: cycle id (min symbol) and distance to it by
// pointer jumping over ds_bpermute, then posinfo[p] = info(pi[p]).
__device__ __forceinline__ uint32_t build_posinfo(WaveLds &L, const uint32_t *nd, int b, int n, int sw, int lane) {
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(nd + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    const int s = lane;
    const int nxt = (s < n) ? (int)spi[s] : s;
    // w: the minimum over the window of 2^r symbols from s on (bits 0..7)
    // and the steps from s to its first occurrence (bits 8..15); p: the
    // symbol 2^r steps on.  Once every window has the same minimum as the
    // next one, each holds its cycle's minimum (otherwise some window
    // without it is followed by one with it), so its steps field is the
    // distance from s to the cycle's minimum: short cycles finish in fewer
    // than the 6 rounds that lengths up to 64 need.
    int w = s, p = nxt;
    for (int r = 0; r < 6; ++r) {
        const int wo = __shfl(w, p);
        if (__ballot((wo & 0xFF) != (w & 0xFF)) == 0) break;
        const int pp = __shfl(p, p);
        if ((wo & 0xFF) < (w & 0xFF)) w = wo + (1 << (8 + r));
        p = pp;
    }
    const int root = w & 0xFF;
    // f is the number of steps to the root, so the root's successor has
    // f = len - 1 (a fixed point is its own successor with f = 0)
    const int f = w >> 8;
    const int f_next = __shfl(f, nxt);
    const int len = __shfl(f_next + 1, root);
    const int pos = (f == 0) ? 0 : len - f;
    const uint32_t info = cycle_info(root, pos, len);
    const int x = (s < n) ? (int)pi[s] : s;
    L.posinfo[b][lane] = (uint32_t)__shfl((int)info, x);
    wave_sync();
    return info;   // of symbol `lane`
}

struct WaveState {
    int wave, lane;
    const uint16_t *colex;         // workgroup's colex pair table (LDS)
    const uint32_t *colex3;        // workgroup's colex triple table (LDS)
    uint32_t lo, hi;               // private stack window [lo, hi), wraps mod 2^32
    int cur_cfg;
    uint32_t cur_gen;              // serving: generation of the slot cur_cfg
    bool configs_left;
    bool stop;
    unsigned long long deadline;   // s_memrealtime ticks (100 MHz); 0 = none
};

__device__ __forceinline__ bool past_deadline(const WaveState &W) {
    return W.deadline != 0 && __builtin_amdgcn_s_memrealtime() > W.deadline;
}

__device__ __forceinline__ void set_watchdog(const SearchParams &P, WaveState &W) {
    if (W.lane == 0) __hip_atomic_store(&P.ctrl[kCtlStop].v, 2ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    W.stop = true;
}

template <bool kServe>
__device__ __forceinline__ void flush_counters(const SearchParams &P, WaveState &W, WaveLds &L) {
    wave_sync();
    if (W.cur_cfg >= 0 && W.lane < kNumCounters) {
        const unsigned long long v = L.cnt[W.lane];
        if (v) fetch_add(&P.results[W.cur_cfg].counters[W.lane], v);
    }
    if (!kServe) {
        if (W.cur_cfg >= 0 && W.lane == kNumCounters && L.live) fetch_add(&P.results[W.cur_cfg].live, L.live);
        if (W.lane == kNumCounters) L.live = 0;
    }
    if (W.lane < kNumCounters) L.cnt[W.lane] = 0;
    wave_sync();
}

// ---- serving: the outstanding-node count of a request.  results[c].live
// counts the request's nodes that exist (on a stack, in an inbox or the spill
// pool, or taken by a wave whose batch has not ended).  A push is added
// before the node can leave its wave (serve_flush_pushes ahead of every
// hand-off or spill) and a take is subtracted only at the end of the batch,
// together with the pushes of its children, so the count never falls below
// the true number: it reaches zero exactly once, when the last node is gone,
// and the wave whose subtraction gets it there reports the request drained.
__device__ __forceinline__ void serve_flush_pushes(const SearchParams &P, const WaveState &W, WaveLds &L) {
    wave_sync();
    const unsigned long long add = L.live;
    wave_sync();
    if (W.lane == 0 && add) {
        L.live = 0;
        fetch_add(&P.results[W.cur_cfg].live, add);
    }
    wave_sync();
}

__device__ __forceinline__ uint32_t *stack_slot(const SearchParams &P, const WaveState &W, uint32_t idx) {
    return P.stacks + ((size_t)W.wave * P.stack_cap + (size_t)(idx & (P.stack_cap - 1))) * P.node_words;
}

__device__ __forceinline__ void set_overflow(const SearchParams &P, WaveState &W) {
    if (W.lane == 0) __hip_atomic_store(&P.ctrl[kCtlStop].v, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    W.stop = true;
}

// ---- work sharing between waves (parallel mode)
// A wave without work counts itself idle, sets its bit in idle_bits and then
// polls only its own inbox line.  A busy wave that sees a set bit claims that
// wave (fetch_and on the bit; the claim also takes the wave out of the idle
// count), copies some of its oldest nodes (the largest subtrees) into the
// claimed wave's inbox and publishes their number on the inbox line.  So the
// idle waves, which are most of the grid for long stretches, never share a
// polled line, and the launch is over exactly when every registered wave is
// idle: work only moves from a busy wave, which is not counted idle.

__device__ __forceinline__ unsigned long long *inbox_line(const SearchParams &P, int wave) {
    return &P.inbox[(size_t)wave * 16];
}
__device__ __forceinline__ unsigned long long *idle_word(const SearchParams &P, int word) {
    return &P.idle_bits[(size_t)word * 16];
}

// Claims one idle wave; returns its index or -1 (wave-uniform).  Up to four
// idle words are read at once (one lane each), so a busy grid costs one
// memory round trip; the words with set bits are then tried in turn.
__device__ __forceinline__ int claim_idle_wave(const SearchParams &P, const WaveState &W, unsigned salt) {
    const int words = (int)P.idle_words;
    const int tries = words < 4 ? words : 4;
    const int w0 = (int)(((unsigned)(W.wave >> 6) + salt) % (unsigned)words);
    int word = w0 + W.lane;
    word = word >= words ? word - words : word;
    unsigned long long bits = 0;
    if (W.lane < tries) bits = ld_relaxed(idle_word(P, word));
    int r = -1;
    for (unsigned long long nz = __ballot(bits != 0); nz && r < 0; nz &= nz - 1) {
        const int c = __ffsll((long long)nz) - 1;
        int got = -1;
        if (W.lane == c) {
            while (bits && got < 0) {
                const int b = __ffsll((long long)bits) - 1;
                const unsigned long long bit = 1ULL << b;
                const unsigned long long old =
                    __hip_atomic_fetch_and(idle_word(P, word), ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (old & bit) got = word * 64 + b;
                else bits = old & ~bit;
            }
            if (got >= 0) fetch_add(&P.ctrl[kCtlIdle].v, ~0ULL);
        }
        r = __builtin_amdgcn_readlane(got, c);
    }
    return r;
}

// Hands up to `most` of the oldest nodes of this wave's stack to an idle wave.
// Returns the number handed over (0: no idle wave).
__device__ __forceinline__ uint32_t share_nodes(const SearchParams &P, WaveState &W, uint32_t most, unsigned salt) {
    if (most == 0) return 0;
    const int r = claim_idle_wave(P, W, salt);
    if (r < 0) return 0;
    const int nw = (int)P.node_words;
    uint32_t *dst = P.inbox_nodes + (size_t)r * P.share_max * nw;
    for (uint32_t g = 0; g < most; ++g) {
        const uint32_t *src = stack_slot(P, W, W.lo + g);
        if (W.lane < nw) st32(dst + g * nw + W.lane, src[W.lane]);
    }
    W.lo += most;
    complete_before();
    if (W.lane == 0) {
        st64(inbox_line(P, r), (unsigned long long)most);
        fetch_add(&P.ctrl[kCtlDonations].v, (unsigned long long)most);
    }
    return most;
}

// Moves the nodes published in this wave's inbox onto its stack (which is
// empty: the wave was idle).  Returns false when the inbox is empty.
__device__ __forceinline__ bool receive_nodes(const SearchParams &P, WaveState &W) {
    unsigned long long cnt = 0;
    if (W.lane == 0) cnt = ld_relaxed(inbox_line(P, W.wave));
    const int k = lane0((int)cnt);
    if (k == 0) return false;
    complete_after();
    const int nw = (int)P.node_words;
    const uint32_t *src = P.inbox_nodes + (size_t)W.wave * P.share_max * nw;
    for (int g = 0; g < k; ++g) {
        uint32_t *dst = stack_slot(P, W, W.hi);
        if (W.lane < nw) dst[W.lane] = ld32(src + g * nw + W.lane);
        W.hi++;
    }
    complete_before();
    if (W.lane == 0) {
        st64(inbox_line(P, W.wave), 0ULL);
        fetch_add(&P.ctrl[kCtlSteals].v, (unsigned long long)k);
    }
    return true;
}

// ---- spill pool: a bounded MPMC ring (per-slot sequence numbers) that takes
// the oldest nodes of a full stack when no wave is idle; idle waves drain it.

// Moves up to `most` of the oldest nodes of the private stack into the
// spill pool with one reservation of consecutive tickets on the tail (one
// CAS for the whole group, so waves spilling at once do not serialise node by
// node).  Each ticket's slot is waited for (its previous node taken), filled,
// and published with the ticket's sequence number.  Returns the number of
// nodes moved (0: the pool is full).
__device__ __forceinline__ uint32_t donate_nodes(const SearchParams &P, WaveState &W, uint32_t most) {
    if (!P.pool_seq || most == 0) return 0;
    unsigned long long ticket = 0;
    uint32_t got = 0;
    if (W.lane == 0) {
        // a lost CAS only means another wave moved the tail: retry until the
        // pool is really full (or the watchdog fires)
        for (unsigned attempt = 0;; ++attempt) {
            // head first, and complete before the tail is read: loads to
            // different lines complete in any order, and a head newer than
            // the tail would make an empty pool look full
            const unsigned long long h = ctl_load(P, kCtlPoolHead);
            complete_after();
            const unsigned long long t = ctl_load(P, kCtlPoolTail);
            const unsigned long long used = t >= h ? t - h : P.pool_cap;
            if (used >= P.pool_cap) break;
            // at most an eighth of the free room (one node at least): a
            // nearly full pool is shared out in small spills, not taken
            // whole by the first waves to overflow
            const unsigned long long room = P.pool_cap - used;
            const unsigned long long share = room >= 8 ? room >> 3 : 1;
            const uint32_t take = (uint32_t)(share < most ? share : most);
            if (cas64(&P.ctrl[kCtlPoolTail].v, t, t + take)) {
                ticket = t;
                got = take;
                break;
            }
            if ((attempt & 63) == 63 && past_deadline(W)) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    got = (uint32_t)lane0((int)got);
    if (got == 0) return 0;
    ticket = lane0_64(ticket);
    for (uint32_t g = 0; g < got; ++g) {
        const size_t slot = (size_t)((ticket + g) & (P.pool_cap - 1));
        int late = 0;
        if (W.lane == 0) {
            while (ld_relaxed(&P.pool_seq[slot]) != ticket + g) {
                if (past_deadline(W)) { late = 1; break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (lane0(late)) { set_watchdog(P, W); return 0; }
        complete_after();
        const uint32_t *src = stack_slot(P, W, W.lo + g);
        uint32_t *dst = P.pool_nodes + slot * P.node_words;
        if (W.lane < (int)P.node_words) st32(dst + W.lane, src[W.lane]);
        // published node by node: a wave that took this ticket waits for one
        // copy, not for the whole group
        complete_before();
        if (W.lane == 0) st64(&P.pool_seq[slot], ticket + g + 1);
    }
    if (W.lane == 0) fetch_add(&P.ctrl[kCtlSpills].v, (unsigned long long)got);
    W.lo += got;
    return got;
}

// Takes one node from the shared pool into L.node.
__device__ __forceinline__ bool take_one(const SearchParams &P, WaveState &W, WaveLds &L) {
    unsigned long long ticket = 0;
    int ok = 0;
    if (W.lane == 0) {
        for (int attempt = 0; attempt < 8; ++attempt) {
            const unsigned long long h = ctl_load(P, kCtlPoolHead);
            complete_after();
            const unsigned long long t = ctl_load(P, kCtlPoolTail);
            if (h >= t) break;
            if (cas64(&P.ctrl[kCtlPoolHead].v, h, h + 1)) { ticket = h; ok = 1; break; }
        }
    }
    ok = lane0(ok);
    if (!ok) return false;
    ticket = lane0_64(ticket);
    const size_t slot = (size_t)(ticket & (P.pool_cap - 1));
    int late = 0;
    if (W.lane == 0) {
        while (ld_relaxed(&P.pool_seq[slot]) != ticket + 1) {
            if (past_deadline(W)) { late = 1; break; }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    if (lane0(late)) { set_watchdog(P, W); return false; }
    complete_after();
    const uint32_t *src = P.pool_nodes + slot * P.node_words;
    if (W.lane < (int)P.node_words) node_at(L, 0, (int)P.spi_words)[W.lane] = ld32(src + W.lane);
    complete_before();
    if (W.lane == 0) {
        st64(&P.pool_seq[slot], ticket + P.pool_cap);
        fetch_add(&P.ctrl[kCtlSpillTakes].v, 1ULL);
    }
    wave_sync();
    return true;
}

// Ensures room for one push on the private stack: a full stack hands its
// oldest nodes to an idle wave, else spills them to the pool, else the launch
// stops with a queue overflow.
template <bool kServe>
__device__ __forceinline__ bool make_room(const SearchParams &P, WaveState &W, WaveLds &L) {
    if (W.hi - W.lo < P.stack_cap) return true;
    unsigned long long reason = 1;   // 1: no sharing, 2: pool full
    if (kServe) serve_flush_pushes(P, W, L);   // before any node leaves the wave
    if (P.donate) {
        uint32_t most = (W.hi - W.lo) / 2;
        if (most > P.share_max) most = P.share_max;
        if (share_nodes(P, W, most, W.hi)) return true;
        // the oldest nodes to the pool.  A launch spills an eighth of the
        // stack, 1..32 nodes, one at least staying: with one configuration
        // every wave is busy and spills, and one reservation per group keeps
        // them from serialising on the tail (a larger share of a small stack
        // would leave the stacks half empty while the pool fills with their
        // halves).  The serving loop spills one node per push: its waves free
        // up as requests finish, and the hand-off tried before every spill
        // then gives them the nodes directly, where the pool would hold them
        // behind newly admitted requests (measured at 10,000 requests/s: a
        // tenth of the spills, p50 3 ms against 1.2 s).
        uint32_t spill = kServe ? 1u : P.stack_cap >> 3;
        spill = spill < 1u ? 1u : spill > 32u ? 32u : spill;
        if (spill > W.hi - W.lo - 1) spill = W.hi - W.lo - 1;
        (void)donate_nodes(P, W, spill);
        if (W.stop) return false;
        if (W.hi - W.lo < P.stack_cap) return true;
        reason = 2;
    }
    if (W.lane == 0) {
        const unsigned long long h = ctl_load(P, kCtlPoolHead);
        complete_after();
        const unsigned long long t = ctl_load(P, kCtlPoolTail);
        st64(&P.ctrl[kCtlOverflowInfo].v, (reason << 56) | ((t - h) & 0xFFFFFFFFULL) |
                                              ((unsigned long long)(W.hi - W.lo) << 32));
    }
    set_overflow(P, W);
    return false;
}

// Word `lane` of the child (t, i, j, k, even) of node pw.  Straight-line:
// the lane reads its parent word and the four pi bytes its word would take
// (one LDS round trip), then selects its word.
__device__ __forceinline__ uint32_t child_word(const uint32_t *pw, int lane, int n, int sw, int t, int i, int j,
                                               int k, int even, const ChildMove &m) {
    const uint8_t *pi = reinterpret_cast<const uint8_t *>(pw + kNodeHeaderWords) + 4 * sw;
    const int d = (int)(pw[1] & 0xFF);
    const uint32_t own = pw[lane < kNodeHeaderWords + 2 * sw ? lane : 0];
    const int q = lane - kNodeHeaderWords - sw;   // pi word of this lane (when 0 <= q < sw)
    uint32_t piw = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int p = (4 * q + b) & 63;
        const uint32_t v = pi[transposed_src_index(p, i, j, k)];
        piw |= (4 * q + b < n ? v : 0u) << (8 * b);
    }
    const int sq = lane - kNodeHeaderWords;       // spi word of this lane (when 0 <= sq < sw)
    uint32_t spw = own;
    spw = ((m.a >> 2) == sq) ? put_byte(spw, m.a, m.na) : spw;
    spw = ((m.b >> 2) == sq) ? put_byte(spw, m.b, m.nb) : spw;
    spw = ((m.c >> 2) == sq) ? put_byte(spw, m.c, m.nc) : spw;
    const int m0 = 2 * (lane - 2);                // header word: moves m0, m0 + 1
    uint32_t hw = own;
    hw = (m0 == d) ? ((hw & 0xFFFF0000u) | (uint32_t)t) : hw;
    hw = (m0 + 1 == d) ? ((hw & 0xFFFFu) | ((uint32_t)t << 16)) : hw;
    return lane == 1 ? ((uint32_t)(d + 1) | ((uint32_t)even << 8))
         : lane < kNodeHeaderWords ? (lane == 0 ? own : hw)
         : sq < sw ? spw : piw;
}

// Pushes the child (t, i, j, k, even) of node pw.
template <bool kServe>
__device__ __forceinline__ void push_child(const SearchParams &P, WaveState &W, WaveLds &L, const uint32_t *pw,
                                           int n, int sw, int t, int i, int j, int k, int even, const ChildMove &m) {
    if (!make_room<kServe>(P, W, L)) return;
    const uint32_t w = child_word(pw, W.lane, n, sw, t, i, j, k, even, m);
    uint32_t *dst = stack_slot(P, W, W.hi);
    if (W.lane < (int)P.node_words) dst[W.lane] = w;
    W.hi++;
    lds_add(&L.traffic[1], 1);
    lds_add(&L.live, 1ULL);
}


// Pushes a copy of node nd (a node of a batch that was not expanded).
template <bool kServe>
__device__ __forceinline__ void push_node(const SearchParams &P, WaveState &W, WaveLds &L, const uint32_t *nd) {
    if (!make_room<kServe>(P, W, L)) return;
    uint32_t *dst = stack_slot(P, W, W.hi);
    if (W.lane < (int)P.node_words) dst[W.lane] = nd[W.lane];
    W.hi++;
    lds_add(&L.traffic[1], 1);
    lds_add(&L.live, 1ULL);
}

// Pushes the staged child with stage entry st (wave-uniform) of a node of the
// batch, its move recomputed from the parent in L.node.
template <bool kServe>
__device__ __forceinline__ void push_staged(const SearchParams &P, WaveState &W, WaveLds &L, int n, int sw, int T,
                                            uint32_t st) {
    const int i = (int)(st & 63u), j = (int)((st >> 6) & 63u), k = (int)((st >> 12) & 63u);
    const uint32_t *nd = node_at(L, st >> 24, sw);
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(nd + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    ChildMove m;
    m.a = pi[i]; m.b = pi[j]; m.c = pi[k];
    m.na = spi[m.c]; m.nb = spi[m.a]; m.nc = spi[m.b];
    push_child<kServe>(P, W, L, nd, n, sw, lex_index(n, T, i, j, k), i, j, k, (int)((st >> 18) & 63u), m);
}

// Ascending bitonic sort of 128 (key, payload) pairs held two per lane
// (element e = slot * 64 + lane); element e holds rank e afterwards.  A
// compare-exchange swaps only on a strict inequality, which both partners
// see alike, so equal keys keep a consistent order.
__device__ __forceinline__ void sort128(float (&key)[2], uint32_t (&val)[2], int lane) {
#pragma unroll
    for (int k = 2; k <= 2 * kWave; k <<= 1) {
#pragma unroll
        for (int d = k >> 1; d > 0; d >>= 1) {
            if (d == kWave) {
                // the lane's two slots, ascending
                const bool sw = key[1] < key[0];
                const float k0 = sw ? key[1] : key[0], k1 = sw ? key[0] : key[1];
                const uint32_t v0 = sw ? val[1] : val[0], v1 = sw ? val[0] : val[1];
                key[0] = k0; key[1] = k1;
                val[0] = v0; val[1] = v1;
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int e = s * kWave + lane;
                    const float pk = __shfl_xor(key[s], d);
                    const uint32_t pv = (uint32_t)__shfl_xor((int)val[s], d);
                    const bool keep_min = ((e & d) == 0) == ((e & k) == 0);
                    const bool sw = keep_min ? (pk < key[s]) : (pk > key[s]);
                    key[s] = sw ? pk : key[s];
                    val[s] = sw ? pv : val[s];
                }
            }
        }
    }
}

// Learned ordering over a whole batch.  The admitted children of one chunk of
// the stage (lane l: score, stage entry, when adm) join the children held back
// so far (L.carry_*, ascending); one sort of the 128 candidates ranks them,
// all but the `keep` best are pushed in ascending score (two per store when a
// node fits in half a wave), and the best `keep` are held back again.
// keep = kWave between chunks and flushes of a batch, 0 at its end, so the
// best kWave children of the batch end on top of the stack in score order
// wherever the stage was cut into chunks and flushes.
template <bool kServe>
__device__ __forceinline__ void learned_push(const SearchParams &P, WaveState &W, WaveLds &L, int n, int sw, int T,
                                             bool adm, float score, uint32_t st, int keep) {
    const int nc = (int)L.ncarry;
    const int total = __popcll(__ballot(adm)) + nc;
    if (total == 0) return;
    // valid keys within +-2e38 (a NaN score ranks lowest), empty entries 3e38,
    // so ranks 0 .. total-1 are the candidates
    float key[2];
    uint32_t val[2];
    key[0] = adm ? fminf(fmaxf(score, -2.0e38f), 2.0e38f) : 3.0e38f;
    val[0] = st;
    const bool vb = W.lane < nc;
    key[1] = vb ? L.carry_score[W.lane] : 3.0e38f;
    val[1] = vb ? L.carry_st[W.lane] : 0u;
    sort128(key, val, W.lane);
    const int losers = total > keep ? total - keep : 0;
    const bool pairs = (int)P.node_words <= kWave / 2;
    for (int r = 0; r < losers && !W.stop;) {
        const uint32_t st0 = (uint32_t)__builtin_amdgcn_readlane((int)(r >= kWave ? val[1] : val[0]), r & (kWave - 1));
        const bool two = pairs && r + 1 < losers && W.hi - W.lo + 2 <= P.stack_cap;
        if (two) {
            const int r1 = r + 1;
            const uint32_t st1 =
                (uint32_t)__builtin_amdgcn_readlane((int)(r1 >= kWave ? val[1] : val[0]), r1 & (kWave - 1));
            const bool upper = W.lane >= kWave / 2;
            const uint32_t cst = upper ? st1 : st0;
            const int i = (int)(cst & 63u), j = (int)((cst >> 6) & 63u), k = (int)((cst >> 12) & 63u);
            const uint32_t *nd = node_at(L, cst >> 24, sw);
            const uint8_t *cspi = reinterpret_cast<const uint8_t *>(nd + kNodeHeaderWords);
            const uint8_t *cpi = cspi + 4 * sw;
            ChildMove cm;
            cm.a = cpi[i]; cm.b = cpi[j]; cm.c = cpi[k];
            cm.na = cspi[cm.c]; cm.nb = cspi[cm.a]; cm.nc = cspi[cm.b];
            const int wl = W.lane & (kWave / 2 - 1);
            const uint32_t w = child_word(nd, wl, n, sw, lex_index(n, T, i, j, k), i, j, k,
                                          (int)((cst >> 18) & 63u), cm);
            uint32_t *dst = stack_slot(P, W, W.hi + (upper ? 1u : 0u));
            if (wl < (int)P.node_words) dst[wl] = w;
            W.hi += 2;
            lds_add(&L.traffic[1], 2);
            lds_add(&L.live, 2ULL);
            r += 2;
        } else {
            push_staged<kServe>(P, W, L, n, sw, T, st0);
            r += 1;
        }
    }
    wave_sync();   // every lane has read the held-back list before it is rewritten
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int r = s * kWave + W.lane;
        if (r >= losers && r < total) {
            L.carry_score[r - losers] = key[s];
            L.carry_st[r - losers] = val[s];
        }
    }
    if (W.lane == 0) L.ncarry = (uint32_t)(total - losers);
    wave_sync();
}

// Lane-parallel dedup of staged children (sorting_search.cu:261-298
// semantics: a child already stored at a depth <= its own is dominated, one
// stored deeper is admitted and lowers the stored depth, an unknown one is
// inserted).  Each lane reads its bucket's eight tags as independent loads
// (one 64-byte line, one memory round trip), compares the stored key of every
// ready tag with its identity (epoch, configuration, 27 hash bits), and
// settles the depth from that tag: dominated, or a CAS from it that lowers the
// depth (a failed CAS repeats the lookup).  A full bucket evicts an entry
// of a configuration already found, else the deepest entry not shallower than
// the child (its subtree is the cheapest to search again); a bucket of
// shallower entries admits the child without an entry, and no entry with the
// child's own identity is evicted for it.  Losing an entry only
// costs repeated work: a verdict never depends on what the table holds.
// While no bucket fills the table is exact, as the CPU oracle's.
// Counts of tag reads, key compares, key writes and evictions go to
// L.traffic[2..5] as popcounts of the lanes taking part (first active lane
// adds, no wait).
__device__ __forceinline__ void count_active(unsigned long long *p, unsigned long long lanes) {
    const unsigned long long ex = __builtin_amdgcn_read_exec();
    if ((int)(threadIdx.x & 63) == __ffsll((long long)ex) - 1)
        __hip_atomic_fetch_add(p, lanes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}


// t[q] for a wave-varying q without dynamic register indexing (which would
// put the array in scratch memory)
__device__ __forceinline__ unsigned long long pick_tag(const unsigned long long (&t)[kDedupBucket], int q) {
    unsigned long long v = 0;
#pragma unroll
    for (int i = 0; i < kDedupBucket; ++i) v = (q == i) ? t[i] : v;
    return v;
}

// The child's key words are in L.ckey[0..nw)[lane] (written by the caller),
// so they do not occupy registers across the probe loop.
// Key words 2w and 2w+1 of a lane's child as one 64-bit word (0 past nw).
__device__ __forceinline__ unsigned long long key_pair(const WaveLds &L, int lane, int w, int nw) {
    const unsigned long long lo = L.ckey[2 * w][lane];
    const unsigned long long hi = (2 * w + 1 < nw) ? L.ckey[2 * w + 1][lane] : 0u;
    return lo | (hi << 32);
}

// Whether the entry with tag tv may be replaced before a live one: its
// configuration is found (batch launches), or its slot has moved to another
// request or its request is done (serving).  f is the loaded found flag or
// slot state word of the entry's configuration.
template <bool kServe>
__device__ __forceinline__ bool entry_done(unsigned long long tv, unsigned long long f) {
    if (kServe) return (f >> 2) != (tv >> 48) || (f & kSlotDone);
    return f != 0;
}
template <bool kServe>
__device__ __forceinline__ unsigned long long entry_done_word(const SearchParams &P, unsigned long long tv) {
    const unsigned long long c = (tv >> 32) & 0xFFFFu;
    if (kServe) return ld_relaxed(&P.slot_state[c]);
    return (unsigned long long)__hip_atomic_load(&P.results[c].found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Serving (P.serve): the epoch field of an identity is the generation of the
// request's slot (gen), a free entry is a zero tag, and an entry is evictable
// first when its slot has moved to another request (other generation) or its
// request is done (found, cancelled or drained).
//
// One round per lookup in the common cases, with the memory accesses of the
// lanes that compare a stored key and of the lanes that insert issued
// together: (1) the bucket's tags; (2) the key of the first entry of this
// identity, or the found flags of a full bucket's configurations; (3) the
// key compare and the tag re-read, or the claim of a victim entry by CAS;
// (4) dominated / improved, or the key written and the entry published.  A
// lane whose stored key of the same identity differs (a 27-bit collision)
// goes on with sequential rounds over all the candidates.
template <bool kServe>
__device__ __forceinline__ int dedup_admit(const SearchParams &P, WaveLds &L, int lane, bool active, uint32_t cfg,
                                           int nw, int depth, uint32_t gen) {
    if (P.dd_buckets == 0) return active ? kDedupUninserted : kDedupDominated;
    // two 32-bit multiplicative lanes over the even and the odd key words
    // (each step a bijection of the lane state, so keys that differ in one
    // word get different 64-bit hashes), joined by the MurmurHash3 64-bit
    // finaliser
    uint32_t h1 = 0x9E3779B9u * (cfg + 1u), h2 = 0x85EBCA6Bu ^ cfg;
    for (int q = 0; q < nw; q += 2) {
        h1 = (h1 ^ L.ckey[q][lane]) * 0xCC9E2D51u;
        h1 ^= h1 >> 15;
        h2 = (h2 ^ (q + 1 < nw ? L.ckey[q + 1][lane] : 0u)) * 0x1B873593u;
        h2 ^= h2 >> 13;
    }
    const uint64_t h = fmix64(((uint64_t)h1 << 32) | h2);
    const unsigned long long ident = dedup_ident(gen, cfg, h);
    const unsigned long long bucket = __umul64hi(h, P.dd_buckets);
    unsigned long long *tags = P.dd_tag + bucket * kDedupBucket;
    const uint32_t *keys = P.dd_key + bucket * kDedupBucket * 16;
    const int width = (int)P.dd_width;
    const int nw2 = (nw + 1) >> 1;
    const unsigned long long epoch = gen;
    int res = active ? kDedupPending : kDedupDominated;
    {
        // Equal children in one chunk: sibling nodes of a batch reach common
        // grandchildren by commuting moves.  Of the lanes holding one key,
        // the one of least (depth, lane) looks it up and the others are
        // dominated by it here (its lookup leaves the key in the table at a
        // depth no greater than theirs, or finds it there, or admits it).
        // The lanes meet in a slot of L.twin chosen by the hash; a lane whose
        // slot another key holds is left to the table.
        const uint32_t me = ((uint32_t)depth << 6) | (uint32_t)lane;
        const int slot = (int)(h & (2 * kWave - 1));
        L.twin[lane] = ~0u;
        L.twin[lane + kWave] = ~0u;
        wave_sync();
        if (active) __hip_atomic_fetch_min(&L.twin[slot], me, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        wave_sync();
        const uint32_t held = L.twin[slot];
        if (active && held != me) {
            const int o = (int)(held & 63u);
            bool eq = true;
            for (int q = 0; q < nw; ++q) eq &= L.ckey[q][o] == L.ckey[q][lane];
            if (eq) res = kDedupDominated;
        }
        wave_sync();
    }
    int rounds = 0;
    bool slow = false;
    // the victim for this child: a free entry (serving: a zero tag; batch: an
    // entry of another epoch), else an entry of a configuration already done,
    // else the deepest entry not shallower than the child; never an entry of
    // the child's own identity
    auto free_victim = [&](const unsigned long long (&t)[kDedupBucket]) -> int {
        int v = -1;
#pragma unroll
        for (int q = kDedupBucket - 1; q >= 0; --q)
            if (q < width && (kServe ? t[q] == 0ULL : (t[q] >> 48) != epoch)) v = q;
        return v;
    };
    auto full_victim = [&](const unsigned long long (&t)[kDedupBucket], const unsigned long long (&f)[kDedupBucket],
                           uint32_t same_ident, bool &live) -> int {
        uint32_t done_mask = 0;
#pragma unroll
        for (int q = 0; q < kDedupBucket; ++q)
            if (q < width && (t[q] & kTagReady) && !((same_ident >> q) & 1u) && entry_done<kServe>(t[q], f[q]))
                done_mask |= 1u << q;
        if (done_mask) {
            live = false;
            return __ffs(done_mask) - 1;
        }
        int v = -1, vd = depth - 1;
#pragma unroll
        for (int q = 0; q < kDedupBucket; ++q) {
            const int dq = (int)(t[q] & 15ULL);
            if (q < width && (t[q] & kTagReady) && !((same_ident >> q) & 1u) && dq > vd) {
                vd = dq;
                v = q;
            }
        }
        live = true;
        return v;
    };
    while (true) {
        const unsigned long long pm = __ballot(res == kDedupPending);
        if (pm == 0) break;
        lds_add(&L.traffic[2], (unsigned long long)width * (unsigned long long)__popcll(pm));
        const bool pend = res == kDedupPending;
        unsigned long long t[kDedupBucket];
#pragma unroll
        for (int q = 0; q < kDedupBucket; ++q) t[q] = (pend && q < width) ? ld_relaxed(tags + q) : 0ULL;
        bool retry = false, busy = false;
        uint32_t cand = 0;   // ready entries with this identity
#pragma unroll
        for (int q = 0; q < kDedupBucket; ++q) {
            const bool same = pend && q < width && (t[q] & ~31ULL) == ident;
            busy |= same && !(t[q] & kTagReady);
            cand |= (same && (t[q] & kTagReady)) ? (1u << q) : 0u;
        }
        const uint32_t same_ident = cand;   // never evicted for this child
        const bool fast = pend && !slow, seq = pend && slow;

        // (2) the first candidate's key; a lane to insert picks a free entry
        // or loads the found flags of the bucket's configurations
        const bool cmp = fast && cand != 0;
        const int q1 = cmp ? __ffs(cand) - 1 : 0;
        bool ins = fast && cand == 0;
        if (ins && busy) {
            retry = true;   // the same identity is being written
            ins = false;
        }
        int victim = ins ? free_victim(t) : -1;
        const bool full = ins && victim < 0;
        // one register array for both: the key words (comparing lanes) or
        // the found words of the bucket's entries (lanes with a full bucket)
        unsigned long long kv[8];
#pragma unroll
        for (int w = 0; w < 8; ++w) kv[w] = 0ULL;
        if (cmp) {
            const unsigned long long *kw = reinterpret_cast<const unsigned long long *>(keys + q1 * 16);
#pragma unroll
            for (int w = 0; w < 8; ++w)
                if (w < nw2) kv[w] = ld_relaxed(kw + w);
        } else if (full) {
#pragma unroll
            for (int q = 0; q < kDedupBucket; ++q)
                if (q < width && (t[q] & kTagReady)) kv[q] = entry_done_word<kServe>(P, t[q]);
        }

        // (3) compare and re-read the tag; claim the victim
        const unsigned long long tq = pick_tag(t, q1);
        bool eq = false;
        unsigned long long tq2 = 0;
        if (cmp) {
            eq = true;
#pragma unroll
            for (int w = 0; w < 8; ++w)
                if (w < nw2) eq &= kv[w] == key_pair(L, lane, w, nw);
            count_active(&L.traffic[3], __popcll(__builtin_amdgcn_read_exec()));
            // The snapshot tag decides only if it still holds after the key
            // was read: an entry evicted and rewritten meanwhile (by a key of
            // the same 27-bit identity) changes the tag word, and the lookup
            // starts over instead of judging this key by the depth of the one
            // the snapshot described.  What remains is an ABA of the whole
            // tag word (same identity, same depth, ready again) inside the
            // window, in which case the entry holds this key at that depth or
            // a key written after it.
            if (eq) tq2 = ld_relaxed(tags + q1);
        }
        bool live = false;
        if (full) victim = full_victim(t, kv, same_ident, live);
        if (ins && victim < 0) res = kDedupUninserted;
        const bool claim = ins && victim >= 0;
        bool claimed = false;
        if (claim) claimed = cas64(tags + victim, pick_tag(t, victim), ident | (unsigned long long)depth);

        // (4) dominated or improved; write the claimed entry's key
        bool improve = false;
        if (cmp) {
            if (!eq) {
                slow = true;   // another key of this identity: sequential rounds
                retry = true;
            } else if (tq2 != tq) {
                retry = true;
            } else if ((int)(tq & 15ULL) <= depth) {
                res = kDedupDominated;
            } else {
                improve = true;
            }
        }
        bool improved = false;
        if (improve) improved = cas64(tags + q1, tq, (tq & ~15ULL) | (unsigned long long)depth);
        if (claimed) {
            // not ready: nobody else touches the entry until the key is
            // complete and the ready bit is published
            unsigned long long *kw =
                reinterpret_cast<unsigned long long *>(P.dd_key + (bucket * kDedupBucket + victim) * 16);
            for (int w = 0; w < nw2; ++w) st64(kw + w, key_pair(L, lane, w, nw));
        }
        complete_before();
        if (improve) {
            if (improved) res = kDedupImproved;
            else retry = true;
        }
        if (claimed) {
            st64(tags + victim, ident | kTagReady | (unsigned long long)depth);
            res = kDedupInserted;
            count_active(&L.traffic[4], __popcll(__builtin_amdgcn_read_exec()));
            count_active(&L.traffic[5], __popcll(__ballot(live)));
        } else if (claim) {
            retry = true;
        }

        // a lane with another key of its identity in the bucket: compare
        // every candidate, then insert
        if (seq) {
#pragma unroll 1
            while (cand && res == kDedupPending && !retry) {
                const int q = __ffs(cand) - 1;
                cand &= cand - 1;
                const unsigned long long tqs = pick_tag(t, q);
                const unsigned long long *kw = reinterpret_cast<const unsigned long long *>(keys + q * 16);
#pragma unroll
                for (int w = 0; w < 8; ++w) kv[w] = w < nw2 ? ld_relaxed(kw + w) : 0ULL;
                bool e = true;
#pragma unroll
                for (int w = 0; w < 8; ++w)
                    if (w < nw2) e &= kv[w] == key_pair(L, lane, w, nw);
                count_active(&L.traffic[3], __popcll(__builtin_amdgcn_read_exec()));
                if (!e) continue;
                if (ld_relaxed(tags + q) != tqs) {
                    retry = true;
                    continue;
                }
                if ((int)(tqs & 15ULL) <= depth) {
                    res = kDedupDominated;
                } else if (cas64(tags + q, tqs, (tqs & ~15ULL) | (unsigned long long)depth)) {
                    res = kDedupImproved;
                } else {
                    retry = true;
                }
            }
            if (res == kDedupPending && !retry) {
                if (busy) {
                    retry = true;
                } else {
                    int v = free_victim(t);
                    bool lv = false;
                    if (v < 0) {
#pragma unroll
                        for (int q = 0; q < kDedupBucket; ++q)
                            kv[q] = (q < width && (t[q] & kTagReady)) ? entry_done_word<kServe>(P, t[q]) : 0ULL;
                        v = full_victim(t, kv, same_ident, lv);
                    }
                    if (v < 0) {
                        res = kDedupUninserted;
                    } else if (cas64(tags + v, pick_tag(t, v), ident | (unsigned long long)depth)) {
                        unsigned long long *kw =
                            reinterpret_cast<unsigned long long *>(P.dd_key + (bucket * kDedupBucket + v) * 16);
                        for (int w = 0; w < nw2; ++w) st64(kw + w, key_pair(L, lane, w, nw));
                        complete_before();
                        st64(tags + v, ident | kTagReady | (unsigned long long)depth);
                        res = kDedupInserted;
                        count_active(&L.traffic[4], __popcll(__builtin_amdgcn_read_exec()));
                        count_active(&L.traffic[5], __popcll(__ballot(lv)));
                    } else {
                        retry = true;
                    }
                }
            }
        }
        if (retry && ++rounds >= kMaxDedupRounds) res = kDedupUninserted;
    }
    return res;
}

// Deduplicates and pushes the staged children (descending triple order).
// root_stage: children of the start node (not counted as admitted).
template <bool kLearned, bool kServe>
__device__ __forceinline__ void flush_stage(const SearchParams &P, WaveState &W, WaveLds &L, const ConfigDesc &C,
                                            int count, int n, int sw, int T, bool root_stage) {
    const uint32_t cfg = node_at(L, 0, sw)[0];   // every node of a batch has the same configuration
    for (int base = 0; base < count && !W.stop; base += kWave) {
        const int idx = base + W.lane;
        const bool act = idx < count;
        const uint32_t st = act ? L.stage[idx] : stage_entry(0, 1, 2, 0, 0);
        const int i = st & 63, j = (st >> 6) & 63, k = (st >> 12) & 63;
        const int ce = (int)((st >> 18) & 63u);
        const uint32_t *nd = node_at(L, st >> 24, sw);   // the child's parent
        const int d = (int)(nd[1] & 0xFF);
        const uint8_t *spi = reinterpret_cast<const uint8_t *>(nd + kNodeHeaderWords);
        const uint8_t *pi = spi + 4 * sw;
        const int t = lex_index(n, T, i, j, k);
        ChildMove m;
        m.a = pi[i]; m.b = pi[j]; m.c = pi[k];
        m.na = spi[m.c]; m.nb = spi[m.a]; m.nc = spi[m.b];
        const uint32_t mv1 = (uint32_t)m.a | ((uint32_t)m.b << 8) | ((uint32_t)m.c << 16);
        const uint32_t mv2 = m.na | (m.nb << 8) | (m.nc << 16);
        PHASE_T0();
        for (int q = 0; q < sw; ++q) L.ckey[q][W.lane] = child_spi_word(nd + kNodeHeaderWords, q, m);
        const int r = dedup_admit<kServe>(P, L, W.lane, act, cfg, sw, d + 1, kServe ? W.cur_gen : P.epoch);
        PHASE_ADD(L, kPhaseDedup);
        const bool admit = act && r != kDedupDominated;
        if (!root_stage) {
            const unsigned long long dom = __popcll(__ballot(act && r == kDedupDominated));
            const unsigned long long imp = __popcll(__ballot(act && r == kDedupImproved));
            const unsigned long long ins = __popcll(__ballot(act && r == kDedupInserted));
            const unsigned long long uni = __popcll(__ballot(act && r == kDedupUninserted));
            lds_add(&L.pend[kCntDominated], dom);
            lds_add(&L.pend[kCntImproved], imp);
            lds_add(&L.pend[kCntInserted], ins);
            lds_add(&L.pend[kCntUninserted], uni);
            lds_add(&L.pend[kCntAdmitted], imp + ins + uni);
        }
        unsigned long long pm = __ballot(admit);
        float score = 0.f;
        if (kLearned && pm) {
            // learned ordering: score the admitted children (MFMA MLP over
            // tiles of 16 of them, compacted in lane order)
            PHASE_T0();
            const int e = (int)((nd[1] >> 8) & 0xFF);
            uint32_t f[kFeatWords];
            const uint32_t *pinfo = L.posinfo[st >> 24];   // the child's parent
            move_features(pinfo[i], pinfo[j], pinfo[k], i, j, k, n, d, cfg_int(C.max_depth),
                          cfg_int(C.e_start), ce, ce - e, f);
            const unsigned long long below = (W.lane == 0) ? 0ULL : (pm & (~0ULL >> (64 - W.lane)));
            const int ci = __popcll(below);
            const int cnt = __popcll(pm);
            if (admit) L.rowsrc[ci] = (uint8_t)W.lane;
            wave_sync();
            for (int rt = 0; rt * 16 < cnt; ++rt) {
                const int row = rt * 16 + (W.lane & 15);
                const int src = row < cnt ? (int)L.rowsrc[row] : 0;
                uint32_t fw[kFeatWords];
#pragma unroll
                for (int w = 0; w < kFeatWords; ++w) {
                    const uint32_t v = (uint32_t)__shfl((int)f[w], src);
                    fw[w] = row < cnt ? v : 0u;
                }
                float out[4];
                mlp_tile_logits(P.mlp, fw, W.lane, out);
                const int local = ci - rt * 16;
                const int q = (local >> 2) & 3, r = local & 3;
                const float s0 = __shfl(out[0], q * 16), s1 = __shfl(out[1], q * 16);
                const float s2 = __shfl(out[2], q * 16), s3 = __shfl(out[3], q * 16);
                if (admit && local >= 0 && local < 16) score = r == 0 ? s0 : r == 1 ? s1 : r == 2 ? s2 : s3;
            }
            PHASE_ADD(L, kPhasePush);
        }
        if constexpr (kLearned) {
            // ascending score over the batch: the best-scored child is pushed
            // last and so taken first
            PHASE_T0();
            learned_push<kServe>(P, W, L, n, sw, T, admit, score, st, kWave);
            PHASE_ADD(L, kPhasePush);
        } else {
            PHASE_T0();
            // The push order is stage order: the batch's nodes are expanded
            // bottom node first (highest node index), each staging its
            // children in descending triple order, so the lowest admitted
            // lane is the next child to push and the top node's children end
            // on top.  The mask stays wave-uniform (scalar registers).
            auto pick = [&]() -> int {
                const int src = uniform(__ffsll((long long)pm) - 1);
                pm &= pm - 1;
                return src;
            };
            // two children per store when a node fits in half a wave: lanes
            // 0..31 write the first, lanes 32..63 the second
            const bool pairs = (int)P.node_words <= kWave / 2;
            // the stack bounds are wave-uniform; said so here, the push loop's
            // control stays in scalar registers
            W.hi = (uint32_t)uniform((int)W.hi);
            W.lo = (uint32_t)uniform((int)W.lo);
            while (pm) {
                const int src0 = pick();
                const bool two = pairs && pm != 0 && W.hi - W.lo + 2 <= P.stack_cap;
                const int src1 = two ? pick() : src0;
                const bool upper = two && W.lane >= kWave / 2;
                // the children's moves from their lanes' registers (no LDS reads)
                const uint32_t cst0 = (uint32_t)__builtin_amdgcn_readlane((int)st, src0);
                const uint32_t cst1 = (uint32_t)__builtin_amdgcn_readlane((int)st, src1);
                const int ct0 = __builtin_amdgcn_readlane(t, src0), ct1 = __builtin_amdgcn_readlane(t, src1);
                const uint32_t a0 = (uint32_t)__builtin_amdgcn_readlane((int)mv1, src0);
                const uint32_t a1 = (uint32_t)__builtin_amdgcn_readlane((int)mv1, src1);
                const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)mv2, src0);
                const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)mv2, src1);
                const uint32_t cst = upper ? cst1 : cst0, c1 = upper ? a1 : a0, c2 = upper ? b1 : b0;
                const int ct = upper ? ct1 : ct0;
                ChildMove cm;
                cm.a = (int)(c1 & 0xFFu); cm.b = (int)((c1 >> 8) & 0xFFu); cm.c = (int)(c1 >> 16);
                cm.na = c2 & 0xFFu; cm.nb = (c2 >> 8) & 0xFFu; cm.nc = c2 >> 16;
                if (two) {
                    const int wl = W.lane & (kWave / 2 - 1);
                    const uint32_t w = child_word(node_at(L, cst >> 24, sw), wl, n, sw, ct, (int)(cst & 63u),
                                                  (int)((cst >> 6) & 63u), (int)((cst >> 12) & 63u),
                                                  (int)((cst >> 18) & 63u), cm);
                    uint32_t *dst = stack_slot(P, W, W.hi + (upper ? 1u : 0u));
                    if (wl < (int)P.node_words) dst[wl] = w;
                    W.hi += 2;
                    lds_add(&L.traffic[1], 2);
                    lds_add(&L.live, 2ULL);
                } else {
                    push_child<kServe>(P, W, L, node_at(L, cst >> 24, sw), n, sw, ct, (int)(cst & 63u), (int)((cst >> 6) & 63u),
                               (int)((cst >> 12) & 63u), (int)((cst >> 18) & 63u), cm);
                }
                if (W.stop) break;
            }
            PHASE_ADD(L, kPhasePush);
        }
    }
}

template <bool kServe>
__device__ __forceinline__ void report_found(const SearchParams &P, WaveState &W, const uint32_t *nd, int cfg,
                                             int depth, int extra_t) {
    int won = 0;
    if (W.lane == 0) {
        int expected = 0;
        won = __hip_atomic_compare_exchange_strong(&P.results[cfg].found, &expected, 1, __ATOMIC_RELAXED,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1 : 0;
        if (won) {
            P.results[cfg].depth = depth;
            fetch_add(&P.ctrl[kCtlResolved].v, 1ULL);
        }
    }
    won = lane0(won);
    if (won) {
        const int pd = (int)(nd[1] & 0xFF);
        uint32_t mv = 0;
        if (W.lane < pd) {
            const uint32_t w = nd[2 + (W.lane >> 1)];
            mv = (W.lane & 1) ? (w >> 16) : (w & 0xFFFF);
        } else if (W.lane == pd && extra_t >= 0) {
            mv = (uint32_t)extra_t;
        }
        if (W.lane < depth) P.results[cfg].moves[W.lane] = mv;
        if (kServe) {
            // the verdict goes to the host now; the slot's dedup entries may go
            ServeReply *hr = P.host_reply + cfg;
            if (W.lane < depth) st_sys32(&hr->moves[W.lane], mv);
            // counters of the batches already ended (later ones are not in the reply)
            if (W.lane < kNumCounters) st_sys(&hr->counters[W.lane], ld_relaxed(&P.results[cfg].counters[W.lane]));
            if (W.lane == 0) {
                st_sys32(reinterpret_cast<uint32_t *>(&hr->found), 1u);
                st_sys32(reinterpret_cast<uint32_t *>(&hr->depth), (uint32_t)depth);
                __hip_atomic_fetch_or(&P.slot_state[cfg], kSlotDone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            complete_before();
            if (W.lane == 0) st_sys(&hr->seq, ld_relaxed(&P.req_index[cfg]) + 1);
        }
    }
}

// Serving: every node of the request in slot cfg is gone.  A request not
// found (or cancelled) gets its reply now, with its final counters; then the
// slot is marked done and handed back to the host.
__device__ __forceinline__ void serve_drained(const SearchParams &P, WaveState &W, int cfg) {
    ServeReply *hr = P.host_reply + cfg;
    complete_after();
    const int found = __hip_atomic_load(&P.results[cfg].found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long r = ld_relaxed(&P.req_index[cfg]);
    if (found != 1) {
        if (W.lane < kNumCounters) st_sys(&hr->counters[W.lane], ld_relaxed(&P.results[cfg].counters[W.lane]));
        if (W.lane == 0) {
            st_sys32(reinterpret_cast<uint32_t *>(&hr->found), (uint32_t)found);
            st_sys32(reinterpret_cast<uint32_t *>(&hr->depth), 0u);
        }
        complete_before();
        if (W.lane == 0) st_sys(&hr->seq, r + 1);
    }
    if (W.lane == 0)
        __hip_atomic_fetch_or(&P.slot_state[cfg], kSlotDone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    complete_before();
    if (W.lane == 0) st_sys(&hr->drained, r + 1);
}

// Serving: end of a batch (or root expansion) of the current request: its
// counters, then its pushes minus its takes; the subtraction that brings the
// outstanding count to zero reports the request drained.
__device__ __forceinline__ void serve_flush(const SearchParams &P, WaveState &W, WaveLds &L) {
    flush_counters<true>(P, W, L);
    const long long delta = (long long)L.live - (long long)L.takes;
    wave_sync();
    int drained = 0;
    if (W.lane == 0) {
        L.live = 0;
        L.takes = 0;
        complete_before();   // the counters land before the count can reach zero
        if (delta != 0) {
            const unsigned long long old = fetch_add(&P.results[W.cur_cfg].live, (unsigned long long)delta);
            drained = (old + (unsigned long long)delta == 0ULL) ? 1 : 0;
        }
    }
    wave_sync();
    if (lane0(drained)) serve_drained(P, W, W.cur_cfg);
}

// Children of an expansion that needs a gain of +2.  Such a child either takes
// its three symbols from one cycle, or two symbols from one even cycle and the
// third from another even cycle: three different cycles merge into one
// (delta <= 0), and two symbols of a cycle of length L with one of a cycle of
// length L' leave two cycles whose odd count can only rise by 2 when L and L'
// are both even.  So instead of all C(n,3) triples only
//   sum over cycles C(L,3)  +  sum over even cycles C(L,2) (E - L)
// triples are evaluated (E: symbols in even cycles), with the same parity
// delta as the full scan, and the result is the full scan's: the goal is the
// largest goal triple index, admitted children are appended to the stage in
// descending triple order (entries tagged with node index b).  Returns
// kLocalNo (nothing done) when the candidates are too many or not clearly
// fewer than C(n,3), kLocalFull when they might not fit behind the scount
// entries already staged; otherwise advances scount / sets goal_t as the scan
// would and returns kLocalDone.
enum { kLocalNo = 0, kLocalDone = 1, kLocalFull = 2 };

__device__ __forceinline__ int cycle_local_children(WaveState &W, WaveLds &L, const uint32_t *nd, int b, int pb,
                                                    uint32_t sinfo, int n, int sw, int T, int e, uint32_t gmask,
                                                    uint32_t amask, bool root_mode, int &scount, int &goal_t) {
    const int s = W.lane;
    const int root = (int)(sinfo & 63u), pos = (int)((sinfo >> 8) & 63u), len = (int)((sinfo >> 16) & 127u);
    const bool is_root = s < n && root == s;
    const bool big = is_root && len >= 3;            // cycles that hold a whole triple
    const bool ev = is_root && (len & 1) == 0;        // even cycles (length >= 2)
    const bool listed = is_root && len >= 2;          // cycles in the member table
    // inclusive prefix sums over the lanes (cycles in root order)
    const int ka = wave_scan_add(big ? choose3(len) : 0);
    const int mo = wave_scan_add(listed ? len : 0);
    const int eo = wave_scan_add(ev ? len : 0);
    const int KA = __builtin_amdgcn_readlane(ka, 63);
    const int E = __builtin_amdgcn_readlane(eo, 63);
    const int kb = wave_scan_add(ev ? choose2(len) * (E - len) : 0);
    const int KB = __builtin_amdgcn_readlane(kb, 63);
    const int K = KA + KB;
    if (K > kCycleLocalMax || 4 * K > T) return kLocalNo;
    if (scount + K > kStage) return kLocalFull;   // the caller flushes the stage first
    const unsigned long long roots_a = __ballot(big), roots_e = __ballot(ev);
    const int moff = mo - (listed ? len : 0);
    const int eoff = eo - (ev ? len : 0);
    // per cycle, in its root lane: first candidate | length << 16 | member
    // offset << 24 (candidates of a cycle are consecutive, cycles in root order)
    const uint32_t acw = (uint32_t)(ka - (big ? choose3(len) : 0)) | ((uint32_t)len << 16) | ((uint32_t)moff << 24);
    const uint32_t ecw = (uint32_t)(kb - (ev ? choose2(len) * (E - len) : 0)) | ((uint32_t)len << 16) |
                         ((uint32_t)eoff << 24);
    // member entries: the symbol's cycle info with its position in pi in
    // bits 24..29, so that one LDS read yields both and sorting three entries
    // sorts them by position
    const uint8_t *pi = reinterpret_cast<const uint8_t *>(nd + kNodeHeaderWords) + 4 * sw;
    // the position of symbol s in pi: lane s sends s to lane pi[s]
    const int pos_in_pi = __builtin_amdgcn_ds_permute((s < n ? (int)pi[s] : s) << 2, s);
    const int offs = __shfl(moff | (eoff << 16), root);
    const int my_moff = offs & 0xFFFF, my_eoff = offs >> 16;
    const uint32_t entry = sinfo | ((uint32_t)pos_in_pi << 24);
    if (s < n && len >= 2) L.memb[my_moff + pos] = entry;
    if (s < n && (len & 1) == 0) L.ememb[my_eoff + pos] = entry;
    wave_sync();
    uint32_t *tmp_k = &L.ckey[0][0];
    uint32_t *tmp_e = &L.ckey[8][0];
    int best = -1, cnt = 0;
    for (int g0 = 0; g0 < K; g0 += kWave) {
        const int g = g0 + s;
        const bool valid = g < K;
        uint32_t ma, mb, mc;
        if (!valid) {
            // three distinct positions, result masked below
            ma = L.posinfo[pb][0];
            mb = L.posinfo[pb][1] | (1u << 24);
            mc = L.posinfo[pb][2] | (2u << 24);
        } else if (g < KA) {
            // three members of one cycle: the cycle holding candidate g (the
            // last one starting at or before it), then its colex triple
            uint32_t cw = 0;
            for (unsigned long long rm = roots_a; rm; rm &= rm - 1) {
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)acw, __ffsll((long long)rm) - 1);
                cw = ((int)(w & 0xFFFFu) <= g) ? w : cw;
            }
            const uint32_t tri = W.colex3[g - (int)(cw & 0xFFFFu)];
            const int co = (int)(cw >> 24);
            ma = L.memb[co + (int)(tri & 0xFFu)];
            mb = L.memb[co + (int)((tri >> 8) & 0xFFu)];
            mc = L.memb[co + (int)(tri >> 16)];
        } else {
            // two members of one even cycle, the third from another even cycle
            const int gg = g - KA;
            uint32_t cw = 0;
            for (unsigned long long rm = roots_e; rm; rm &= rm - 1) {
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)ecw, __ffsll((long long)rm) - 1);
                cw = ((int)(w & 0xFFFFu) <= gg) ? w : cw;
            }
            const int cl = (int)((cw >> 16) & 0xFFu), co = (int)(cw >> 24);
            const int r = gg - (int)(cw & 0xFFFFu);
            const int others = E - cl;
            const int pair = r / others, zi = r - pair * others;
            const uint32_t pw = W.colex[pair];
            ma = L.ememb[co + (int)(pw & 0xFFu)];
            mb = L.ememb[co + (int)(pw >> 8)];
            mc = L.ememb[zi < co ? zi : zi + cl];
        }
        // sorted by position (distinct positions in bits 24..29)
        const uint32_t lo = min(ma, min(mb, mc)), hi = max(ma, max(mb, mc));
        const uint32_t mid = ma ^ mb ^ mc ^ lo ^ hi;
        const int i = (int)(lo >> 24), j = (int)(mid >> 24), k = (int)(hi >> 24);
        const int delta = parity_delta(lo, mid, hi);
        const uint32_t di = (uint32_t)(delta + 2) >> 1;
        const bool goal = valid && !root_mode && ((gmask >> di) & 1u);
        const bool adm = valid && ((amask >> di) & 1u);
        // i << 12 | j << 6 | k grows with the lexicographic triple index
        const int key = (i << 12) | (j << 6) | k;
        if (goal) best = max(best, key);
        const unsigned long long am = __ballot(adm);
        const unsigned long long below = (s == 0) ? 0ULL : (am & (~0ULL >> (64 - s)));
        if (adm) {
            tmp_k[cnt + __popcll(below)] = (uint32_t)key;
            tmp_e[cnt + __popcll(below)] = stage_entry(i, j, k, e + delta, b);
        }
        cnt += __popcll(am);
    }
    best = wave_max(best);
    if (best >= 0) {
        goal_t = lex_index(n, T, best >> 12, (best >> 6) & 63, best & 63);
        return kLocalDone;
    }
    wave_sync();
    // descending triple order: the rank of an entry is the number of larger
    // keys (all distinct); up to 64 entries rank against registers
    if (cnt <= kWave) {
        const uint32_t mk = s < cnt ? tmp_k[s] : 0u;
        const uint32_t me = s < cnt ? tmp_e[s] : 0u;
        int rank = 0;
        for (int q = 0; q < cnt; ++q) rank += (uint32_t)__builtin_amdgcn_readlane((int)mk, q) > mk ? 1 : 0;
        wave_sync();
        if (s < cnt) L.stage[scount + rank] = me;
    } else {
        for (int idx = s; idx < cnt; idx += kWave) {
            const uint32_t kk = tmp_k[idx];
            int rank = 0;
            for (int q = 0; q < cnt; ++q) rank += tmp_k[q] > kk ? 1 : 0;
            L.stage[scount + rank] = tmp_e[idx];
        }
    }
    if (root_mode) lds_add(&L.cnt[kCntRootsLive], (unsigned long long)cnt);
    wave_sync();
    scount += cnt;
    return kLocalDone;
}

// The deltas that make a goal child (gmask) and an admissible one (amask) of
// a node at depth d with e odd cycles: bits over delta in {-2, 0, +2} (both
// predicates are monotone in the odd-cycle count and a move changes it by -2,
// 0 or +2).  Root mode admits goals and live roots alike.
__device__ __forceinline__ void child_masks(const ConfigDesc &C, int d, int e, bool root_mode, uint32_t &gmask,
                                            uint32_t &amask) {
    const int D = cfg_int(C.max_depth);
    const uint64_t acc = cfg_bits(C.accept[d + 1]);
    const uint64_t rch = (d + 1 < D) ? cfg_bits(C.reach[d + 1]) : 0ULL;
    gmask = 0;
    amask = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int ev = e + 2 * q - 2;
        if (ev >= 0 && ev < 64) {
            gmask |= (uint32_t)((acc >> ev) & 1ULL) << q;
            amask |= (uint32_t)((rch >> ev) & 1ULL) << q;
        }
    }
    if (root_mode) {
        amask |= gmask;
        gmask = 0;
    }
}

// Two nodes of a batch (nodes ba and bb of the pool, n <= 32) expanded together
// when only delta = +2 qualifies for both: lanes 0..31 hold the symbols of
// node ba, lanes 32..63 those of node bb, so one pass of pointer jumping
// analyses both and one chunk loop enumerates the cycle-local children of
// both (node ba's candidates first).  The result is that of
// cycle_local_children on ba and then bb: ba's admitted children staged
// before bb's, each node's in descending triple order; a goal of ba ends the
// enumeration before bb's children count, a goal of bb drops ba's staged
// children.  Returns kLocalNo (nothing done: an even cycle, too many
// candidates for the pair or for either node) or kLocalFull (nothing done:
// no room behind scount), else kLocalDone with goal_b / goal_t set on a goal.
// With kPosinfo (the learned ordering) both nodes' cycle info per position
// goes to L.posinfo[ba] and L.posinfo[bb] as build_posinfo writes it (roots
// are lane indices there, which the features compare only for equality).
template <bool kPosinfo>
__device__ __forceinline__ int pair_local_children(WaveState &W, WaveLds &L, int ba, int bb, int n, int sw, int T,
                                                   int ea, int eb, uint32_t ga, uint32_t aa, uint32_t gb,
                                                   uint32_t ab, int &scount, int &goal_b, int &goal_t) {
    const int l = W.lane, half = l >> 5, s = l & 31, base = l & 32;
    // cycle analysis of both nodes (build_posinfo's pointer jumping, lane
    // indices as symbols: a cycle of node bb lies in lanes 32..63)
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(node_at(L, half ? bb : ba, sw) + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    const int nxt = (s < n) ? base + (int)spi[s] : l;
    int w = l, p = nxt;
    for (int r = 0; r < 5; ++r) {
        const int wo = __shfl(w, p);
        if (__ballot((wo & 0xFF) != (w & 0xFF)) == 0) break;
        const int pp = __shfl(p, p);
        if ((wo & 0xFF) < (w & 0xFF)) w = wo + (1 << (8 + r));
        p = pp;
    }
    const int root = w & 0xFF;
    const int f = w >> 8;
    const int f_next = __shfl(f, nxt);
    const int len = __shfl(f_next + 1, root);
    const int pos = (f == 0) ? 0 : len - f;
    const bool is_root = s < n && root == l;
    if (__ballot(is_root && (len & 1) == 0) != 0) return kLocalNo;   // even cycles: one node at a time
    if (kPosinfo) {
        // posinfo[p] = info(pi[p]): position p of each node in its half's lane p
        const uint32_t at = (uint32_t)__shfl((int)cycle_info(root, pos, len), s < n ? base + (int)pi[s] : l);
        if (s < n) L.posinfo[half ? bb : ba][s] = at;
    }
    const bool big = is_root && len >= 3;
    const bool listed = is_root && len >= 2;
    // candidates and members over both nodes (node ba's first)
    const int ka = wave_scan_add(big ? choose3(len) : 0);
    const int mo = wave_scan_add(listed ? len : 0);
    const int KA = __builtin_amdgcn_readlane(ka, 31);
    const int K = __builtin_amdgcn_readlane(ka, 63);
    const int KB = K - KA;
    if (K > kCycleLocalMax || 4 * KA > T || 4 * KB > T) return kLocalNo;
    if (scount + K > kStage) return kLocalFull;
    const unsigned long long roots_a = __ballot(big);
    const int moff = mo - (listed ? len : 0);
    const uint32_t acw = (uint32_t)(ka - (big ? choose3(len) : 0)) | ((uint32_t)len << 16) | ((uint32_t)moff << 24);
    // member entries: cycle info (root lane, position, length) with the
    // symbol's position in its node's pi in bits 24..29
    const int pos_in_pi = __builtin_amdgcn_ds_permute((s < n ? base + (int)pi[s] : l) << 2, s);
    const int my_moff = __shfl(moff, root);
    if (s < n && len >= 2) L.memb[my_moff + pos] = cycle_info(root, pos, len) | ((uint32_t)pos_in_pi << 24);
    wave_sync();
    uint32_t *tmp_k = &L.ckey[0][0];
    uint32_t *tmp_e = &L.ckey[8][0];
    int best = -1, cnt = 0;
    for (int g0 = 0; g0 < K; g0 += kWave) {
        const int g = g0 + l;
        const bool valid = g < K;
        const bool inb = g >= KA;
        uint32_t ma, mb, mc;
        if (!valid) {
            // three fixed points at distinct positions, result masked below
            ma = cycle_info(0, 0, 1);
            mb = cycle_info(1, 0, 1) | (1u << 24);
            mc = cycle_info(2, 0, 1) | (2u << 24);
        } else {
            // the cycle holding candidate g: the last one starting at or
            // before it (node bb's cycles start at KA or later)
            uint32_t cw = 0;
            for (unsigned long long rm = roots_a; rm; rm &= rm - 1) {
                const uint32_t cv = (uint32_t)__builtin_amdgcn_readlane((int)acw, __ffsll((long long)rm) - 1);
                cw = ((int)(cv & 0xFFFFu) <= g) ? cv : cw;
            }
            const uint32_t tri = W.colex3[g - (int)(cw & 0xFFFFu)];
            const int co = (int)(cw >> 24);
            ma = L.memb[co + (int)(tri & 0xFFu)];
            mb = L.memb[co + (int)((tri >> 8) & 0xFFu)];
            mc = L.memb[co + (int)(tri >> 16)];
        }
        const uint32_t lo = min(ma, min(mb, mc)), hi = max(ma, max(mb, mc));
        const uint32_t mid = ma ^ mb ^ mc ^ lo ^ hi;
        const int i = (int)(lo >> 24), j = (int)(mid >> 24), k = (int)(hi >> 24);
        const int delta = parity_delta(lo, mid, hi);
        const uint32_t di = (uint32_t)(delta + 2) >> 1;
        const bool goal = valid && (((inb ? gb : ga) >> di) & 1u);
        const bool adm = valid && (((inb ? ab : aa) >> di) & 1u);
        // descending order puts node ba's entries first
        const int key = (inb ? 0 : (1 << 18)) | (i << 12) | (j << 6) | k;
        if (goal) best = max(best, key);
        const unsigned long long am = __ballot(adm);
        const unsigned long long below = (l == 0) ? 0ULL : (am & (~0ULL >> (64 - l)));
        if (adm) {
            tmp_k[cnt + __popcll(below)] = (uint32_t)key;
            tmp_e[cnt + __popcll(below)] = stage_entry(i, j, k, (inb ? eb : ea) + delta, inb ? bb : ba);
        }
        cnt += __popcll(am);
    }
    best = wave_max(best);
    if (best >= 0) {
        goal_b = (best >> 18) ? ba : bb;
        goal_t = lex_index(n, T, (best >> 12) & 63, (best >> 6) & 63, best & 63);
        return kLocalDone;
    }
    wave_sync();
    if (cnt <= kWave) {
        const uint32_t mk = l < cnt ? tmp_k[l] : 0u;
        const uint32_t me = l < cnt ? tmp_e[l] : 0u;
        int rank = 0;
        for (int q = 0; q < cnt; ++q) rank += (uint32_t)__builtin_amdgcn_readlane((int)mk, q) > mk ? 1 : 0;
        wave_sync();
        if (l < cnt) L.stage[scount + rank] = me;
    } else {
        for (int idx = l; idx < cnt; idx += kWave) {
            const uint32_t kk = tmp_k[idx];
            int rank = 0;
            for (int q = 0; q < cnt; ++q) rank += tmp_k[q] > kk ? 1 : 0;
            L.stage[scount + rank] = tmp_e[idx];
        }
    }
    wave_sync();
    scount += cnt;
    return kLocalDone;
}

// Evaluates every child of node nd (index b of the batch in L.node) and
// stages the admissible ones behind the scount entries already staged,
// flushing the stage only when it would overflow.  Normal mode: returns the
// goal triple (terminal expansion) or -1.  Root mode: stages every live root.
template <bool kLearned, bool kServe>
__device__ __forceinline__ int expand_stage(const SearchParams &P, WaveState &W, WaveLds &L, const ConfigDesc &C,
                                            const uint32_t *nd, int b, bool root_mode, int &scount) {
    const int n = cfg_int(C.n);
    const int sw = (int)P.spi_words;
    const int T = cfg_int(C.triple_count);
    const int D = cfg_int(C.max_depth);
    const int d = (int)(nd[1] & 0xFF);
    const int e = (int)((nd[1] >> 8) & 0xFF);
    // the learned ordering scores children at the flush from their parent's
    // cycle info, so each node of its batches keeps its own
    const int pb = kLearned ? b : 0;
    uint32_t sinfo;
    {
        PHASE_T0();
        sinfo = build_posinfo(L, nd, pb, n, sw, W.lane);
        PHASE_ADD(L, kPhaseAnalyse);
    }
    // Both predicates are monotone in the odd-cycle count and a move changes it
    // by -2, 0 or +2, so each expansion reduces to two 3-bit masks over
    // delta in {-2, 0, +2}: which deltas make a goal child, which make an
    // admissible one.
    const uint64_t acc = cfg_bits(C.accept[d + 1]);
    const uint64_t rch = (d + 1 < D) ? cfg_bits(C.reach[d + 1]) : 0ULL;
    uint32_t gmask = 0, amask = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int ev = e + 2 * q - 2;
        if (ev >= 0 && ev < 64) {
            gmask |= (uint32_t)((acc >> ev) & 1ULL) << q;
            amask |= (uint32_t)((rch >> ev) & 1ULL) << q;
        }
    }
    if (root_mode) {
        amask |= gmask;   // live roots: goals or can_still_reach at depth 1
        gmask = 0;
    }
    if (gmask == 0 && amask == 0) return -1;   // no child can be a goal or admitted
    int goal_t = -1;
    PHASE_T0();
    // When only delta = +2 can qualify, a child must hit one cycle twice (three
    // different cycles merge, delta <= 0): chunks without such a child are
    // skipped wave-wide.
    const bool plus2_only = (gmask | amask) == 4u;
    // Visits the chunk whose lane l holds the child at descending offset
    // obase + l (lexicographic index T-1-obase-l); true ends the scan.
    auto visit = [&](bool valid, int i, int j, int k, int obase) -> int {
        const uint32_t ia = L.posinfo[pb][i], ib = L.posinfo[pb][j], ic = L.posinfo[pb][k];
        int delta;
        if (plus2_only) {
            const bool shared = hits_a_cycle_twice(ia, ib, ic);
            if (__ballot(valid && shared) == 0) return kScanGo;
            delta = shared ? parity_delta(ia, ib, ic) : -2;
        } else {
            delta = child_delta(ia, ib, ic);
        }
        const uint32_t di = (uint32_t)(delta + 2) >> 1;
        const bool goal = valid && ((gmask >> di) & 1u);
        const bool adm = valid && ((amask >> di) & 1u);
        if (!root_mode) {
            const unsigned long long gm = __ballot(goal);
            if (gm) {
                goal_t = T - 1 - (obase + __ffsll((long long)gm) - 1);
                return kScanStop;
            }
        }
        const unsigned long long am = __ballot(adm);
        if (am) {
            const unsigned long long below = (W.lane == 0) ? 0ULL : (am & (~0ULL >> (64 - W.lane)));
            if (adm) L.stage[scount + __popcll(below)] = stage_entry(i, j, k, e + delta, b);
            scount += __popcll(am);
            if (root_mode) lds_add(&L.cnt[kCntRootsLive], __popcll(am));
            if (scount > kStage - kWave) return kScanPause;
        }
        return kScanGo;
    };
    int local = kLocalNo;
    if (plus2_only) {
        local = cycle_local_children(W, L, nd, b, pb, sinfo, n, sw, T, e, gmask, amask, root_mode, scount, goal_t);
        if (local == kLocalFull) {
            // the children of earlier nodes of the batch leave no room
            wave_sync();
            flush_stage<kLearned, kServe>(P, W, L, C, scount, n, sw, T, root_mode);
            scount = 0;
            if (W.stop) return -1;
            local = cycle_local_children(W, L, nd, b, pb, sinfo, n, sw, T, e, gmask, amask, root_mode, scount, goal_t);
        }
    }
    if (local == kLocalNo) {
        if (scount > kStage - kWave) {
            // a chunk of the scan may stage up to kWave children
            wave_sync();
            flush_stage<kLearned, kServe>(P, W, L, C, scount, n, sw, T, root_mode);
            scount = 0;
            if (W.stop) return -1;
        }
        int mb = -1, qb = 0;
        while (scan_children(W.colex, n, T, W.lane, mb, qb, visit) == kScanPause) {
            // stage full: flush, then resume the scan at the next chunk
            wave_sync();
            flush_stage<kLearned, kServe>(P, W, L, C, scount, n, sw, T, root_mode);
            scount = 0;
            if (W.stop) return -1;
        }
    }
    if (W.stop) return -1;
    PHASE_ADD(L, kPhaseScan);
    return goal_t;
}

// A batch launch's descriptors do not change while it runs: the wave keeps
// its current configuration's descriptor in LDS, so the per-node predicate
// tables and sizes are LDS reads rather than vector loads that would also
// wait for the wave's outstanding stores.
__device__ __forceinline__ void switch_config(const SearchParams &P, WaveState &W, WaveLds &L, ConfigDesc &copy,
                                              int cfg) {
    if (cfg != W.cur_cfg) {
        flush_counters<false>(P, W, L);
        W.cur_cfg = cfg;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(P.cfgs + cfg);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&copy);
        constexpr int kWords = (int)(sizeof(ConfigDesc) / 4);
        for (int w = W.lane; w < kWords; w += kWave) dst[w] = src[w];
        wave_sync();
    }
}

__device__ __forceinline__ bool config_found(const SearchParams &P, int cfg) {
    return __hip_atomic_load(&P.results[cfg].found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

template <bool kServe>
__device__ __forceinline__ bool should_stop(const SearchParams &P) {
    int s = 0;
    if ((threadIdx.x & 63) == 0) {
        s = ctl_load(P, kCtlStop) != 0;
        if (P.cancel_flag)
            s |= __hip_atomic_load(P.cancel_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        // a serving session has no fixed set of configurations to resolve
        if (!kServe) s |= ctl_load(P, kCtlResolved) >= (unsigned long long)P.num_configs;
    }
    return lane0(s) != 0;
}

// Serving: the slot's descriptor into the wave's LDS copy, read with
// device-coherent loads (the feeder rewrites a slot's descriptor for each
// new request, so no cached copy of an earlier one may be used).
__device__ __forceinline__ void serve_load_config(const SearchParams &P, const WaveState &W, ConfigDesc &copy,
                                                  int cfg) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(P.cfgs + cfg);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&copy);
    constexpr int kWords = (int)(sizeof(ConfigDesc) / 4);
    for (int w = W.lane; w < kWords; w += kWave) dst[w] = ld32(src + w);
    wave_sync();
}

// Serving: the state word of slot cfg; a wave whose LDS copy belongs to
// another configuration or generation reloads it.  Returns true when the
// request is done (its remaining nodes are dropped).
__device__ __forceinline__ bool serve_switch(const SearchParams &P, WaveState &W, WaveLds &L, ConfigDesc &copy,
                                             int cfg) {
    unsigned long long st = 0;
    if (W.lane == 0) st = ld_relaxed(&P.slot_state[cfg]);
    const uint32_t lo = (uint32_t)lane0((int)(uint32_t)st);
    const uint32_t gen = lo >> 2;
    if (cfg != W.cur_cfg || gen != W.cur_gen) {
        flush_counters<true>(P, W, L);
        W.cur_cfg = cfg;
        W.cur_gen = gen;
        serve_load_config(P, W, copy, cfg);
    }
    return (lo & (uint32_t)kSlotDone) != 0;
}

// Serving: claims the next request the feeder has published, if any.
// Returns its slot or -1 (wave-uniform).
__device__ __forceinline__ int serve_claim(const SearchParams &P, const WaveState &W) {
    int slot = -1;
    if (W.lane == 0) {
        const unsigned long long pub = ctl_load(P, kCtlPublished);
        complete_after();
        const unsigned long long next = ctl_load(P, kCtlNextConfig);
        if (next < pub && cas64(&P.ctrl[kCtlNextConfig].v, next, next + 1)) {
            complete_after();
            slot = (int)ld32(&P.ring_slot[next % P.ring]);
        }
    }
    return lane0(slot);
}

// Serving: the feeder wave.  Copies each request the host publishes in its
// ring (in ring order) into the device slot the host chose, resets the
// slot's result and outstanding count, starts a new generation of the slot,
// then publishes it to the search waves.  Ends on the host's stop word (and
// then stops the search waves), on any other stop, or at the watchdog.
__device__ __forceinline__ void serve_feeder(const SearchParams &P, WaveState &W) {
    unsigned long long next = 0, cancels = 0;
    int backoff = 1;
    constexpr int kWords = (int)(sizeof(ConfigDesc) / 4);
    while (true) {
        unsigned long long seq = 0, cancel_pub = 0;
        int host_stop = 0, dev_stop = 0;
        if (W.lane == 0) {
            host_stop = ld_sys(&P.host_ctl[kServeHostStop]) != 0;
            dev_stop = ctl_load(P, kCtlStop) != 0;
            seq = ld_sys(&P.host_req[next % P.ring].seq);
            cancel_pub = ld_sys(&P.host_ctl[kServeHostCancels]);
            // cancellations: entry = request << 16 | slot; a request still in
            // its slot is marked cancelled (found = 2) and done, so its nodes
            // are dropped and the wave that drains it replies "cancelled"
            for (; cancels < cancel_pub; ++cancels) {
                complete_after();
                const unsigned long long e =
                    ld_sys(&P.host_ctl[kServeHostCancelRing + (cancels % kServeCancelEntries)]);
                const int slot = (int)(e & 0xFFFFULL);
                if (slot < P.num_configs && ld_relaxed(&P.req_index[slot]) == (e >> 16)) {
                    int expected = 0;
                    __hip_atomic_compare_exchange_strong(&P.results[slot].found, &expected, 2, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_or(&P.slot_state[slot], kSlotDone, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (lane0(host_stop)) {
            if (W.lane == 0) __hip_atomic_store(&P.ctrl[kCtlStop].v, 3ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        if (lane0(dev_stop)) break;
        if (past_deadline(W)) {
            set_watchdog(P, W);
            break;
        }
        seq = lane0_64(seq);
        // at most one ring of published requests not yet claimed (the
        // claimers read their slot from ring_slot)
        const bool ahead = next - ctl_load_uniform(P, kCtlNextConfig) >= (unsigned long long)P.ring;
        if (seq != next + 1 || ahead) {
            for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(32);
            if (backoff < 16) backoff *= 2;
            continue;
        }
        backoff = 1;
        complete_after();
        const ServeRequest *rq = P.host_req + next % P.ring;
        const int slot = uniform((int)ld_sys32(&rq->slot));
        const uint32_t *src = reinterpret_cast<const uint32_t *>(&rq->desc);
        // the session's descriptor slots are written only by this wave
        uint32_t *dst = reinterpret_cast<uint32_t *>(const_cast<ConfigDesc *>(P.cfgs + slot));
        for (int w = W.lane; w < kWords; w += kWave) st32(dst + w, ld_sys32(src + w));
        ConfigResult &res = P.results[slot];
        if (W.lane < kNumCounters) st64(&res.counters[W.lane], 0ULL);
        if (W.lane < 16) st32(&res.moves[W.lane], 0u);
        if (W.lane == 0) {
            st32(reinterpret_cast<uint32_t *>(&res.found), 0u);
            st32(reinterpret_cast<uint32_t *>(&res.depth), 0u);
            st64(&res.live, 0ULL);
            st64(&P.req_index[slot], next);
            st32(&P.ring_slot[next % P.ring], (uint32_t)slot);
            // next generation of the slot (1..65535): the previous request's
            // dedup entries no longer match and may be evicted
            const unsigned long long st = ld_relaxed(&P.slot_state[slot]);
            unsigned long long gen = ((st >> 2) + 1) & 0xFFFFULL;
            if (gen == 0) gen = 1;
            st64(&P.slot_state[slot], gen << 2);
        }
        complete_before();
        if (W.lane == 0) {
            st64(&P.ctrl[kCtlPublished].v, next + 1);
            st_sys(&P.host_ctl[kServeHostConsumed], next + 1);
        }
        ++next;
    }
}

}  // namespace

template <bool kLearned, bool kServe>
__device__ __forceinline__ void search_body(const SearchParams &P) {
    __shared__ WaveLds lds[kWavesPerBlock];
    __shared__ uint16_t colex[kColexPairs];
    __shared__ uint32_t colex3[kCycleLocalMax];
    // each wave's copy of the descriptor of its current configuration
    // (serving: of its current request)
    __shared__ ConfigDesc cfg_copy[kWavesPerBlock];
    fill_colex(colex, colex3);
    const int wib = threadIdx.x >> 6;
    WaveLds &L = lds[wib];
    WaveState W;
    W.lane = threadIdx.x & 63;
    // a serving launch: block 0 (dispatched first, so it is resident whatever
    // else runs) holds the feeder wave; the search waves follow
    W.wave = (int)(kServe ? blockIdx.x - 1 : blockIdx.x) * kWavesPerBlock + wib;
    if (kServe && blockIdx.x == 0) {
        if (wib != 0) return;
        W.deadline = P.timeout_ticks ? __builtin_amdgcn_s_memrealtime() + P.timeout_ticks : 0ULL;
        W.stop = false;
        serve_feeder(P, W);
        return;
    }
    if (W.wave >= P.num_waves) return;
    W.colex = colex;
    W.colex3 = colex3;
    W.lo = W.hi = 0;
    W.cur_cfg = -1;
    W.cur_gen = 0;
    W.configs_left = true;
    W.stop = false;
    W.deadline = P.timeout_ticks ? __builtin_amdgcn_s_memrealtime() + P.timeout_ticks : 0ULL;
    if (W.lane < kNumCounters) L.cnt[W.lane] = 0;
    if (W.lane == 0) L.live = 0;
    if (W.lane == 0) L.takes = 0;
    if (W.lane < 6) L.traffic[W.lane] = 0;
    if (W.lane < kNumPhases) L.phase[W.lane] = 0;
    wave_sync();
    if (W.lane == 0) fetch_add(&P.ctrl[kCtlRegistered].v, 1ULL);
    bool idle = false;
    unsigned iter = 0;
    unsigned polls = 0, idle_polls = 0;
    int backoff = 1;

    while (true) {
        W.hi = (uint32_t)uniform((int)W.hi);
        W.lo = (uint32_t)uniform((int)W.lo);
        if (W.stop || ((++polls & 31) == 0 && should_stop<kServe>(P))) break;
        if ((polls & 63) == 0 && past_deadline(W)) {
            set_watchdog(P, W);
            break;
        }
        bool have = false;
        int batch = 1;
        if (W.hi > W.lo) {
            PHASE_T0();
            // up to P.batch nodes from the top of the stack (loads in flight
            // together), of which the leading ones of one configuration are
            // taken
            const uint32_t avail = W.hi - W.lo;
            // A stack more than half full takes one node.
            const uint32_t most = (avail * 2 > P.stack_cap) ? 1u : (uint32_t)P.batch;
            const int want = (int)(avail < most ? avail : most);
            uint32_t v[kBatchMax];
#pragma unroll
            for (int b = 0; b < kBatchMax; ++b) {
                v[b] = 0;
                if (b < want && W.lane < (int)P.node_words) v[b] = stack_slot(P, W, W.hi - 1 - b)[W.lane];
            }
#pragma unroll
            for (int b = 0; b < kBatchMax; ++b)
                if (b < want && W.lane < (int)P.node_words) node_at(L, b, (int)P.spi_words)[W.lane] = v[b];
            wave_sync();
            const bool same = W.lane < want && node_at(L, W.lane < want ? W.lane : 0, (int)P.spi_words)[0] == node_at(L, 0, (int)P.spi_words)[0];
            batch = __ffsll((long long)~__ballot(same)) - 1;
            W.hi -= (uint32_t)batch;
            lds_add(&L.traffic[0], (unsigned long long)batch);
            have = true;
            PHASE_ADD(L, kPhasePop);
        } else if (kServe ? (ctl_load_uniform(P, kCtlPublished) > ctl_load_uniform(P, kCtlNextConfig))
                          : W.configs_left) {
            int c = 0;
            if constexpr (kServe) {
                if (idle) {
                    // leave the idle set first; a wave already claimed by a
                    // donor keeps its place and takes the nodes on their way
                    int mine = 0;
                    if (W.lane == 0) {
                        const unsigned long long bit = 1ULL << (W.wave & 63);
                        mine = (__hip_atomic_fetch_and(idle_word(P, W.wave >> 6), ~bit, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) & bit) ? 1 : 0;
                        if (mine) fetch_add(&P.ctrl[kCtlIdle].v, ~0ULL);
                    }
                    if (!lane0(mine)) {
                        // a donor has claimed this wave: its nodes are on the
                        // way to the inbox (the donor took the wave out of the
                        // idle count)
                        while (!receive_nodes(P, W)) {
                            if (past_deadline(W)) {
                                set_watchdog(P, W);
                                break;
                            }
                            __builtin_amdgcn_s_sleep(2);
                        }
                        idle = false;
                        continue;
                    }
                    idle = false;
                }
                c = serve_claim(P, W);
                if (c < 0) continue;
                serve_switch(P, W, L, cfg_copy[wib], c);
                // the start node counts as pushed before its root expansion
                // takes it (flushed with the roots at the end)
                if (W.lane == 0) fetch_add(&P.results[c].live, 1ULL);
                if (W.lane == 0) L.takes = 1;
            } else {
                if (W.lane == 0) c = (int)fetch_add(&P.ctrl[kCtlNextConfig].v, 1ULL);
                c = lane0(c);
                if (c >= P.num_configs) {
                    W.configs_left = false;
                    continue;
                }
                // start node of configuration c, expanded in root mode
                switch_config(P, W, L, cfg_copy[wib], c);
            }
            const ConfigDesc &C = cfg_copy[wib];
            const int sw = (int)P.spi_words;
            if (W.lane < (int)P.node_words) {
                uint32_t w = 0;
                if (W.lane == 0) w = (uint32_t)c;
                else if (W.lane == 1) w = C.e_start << 8;
                else if (W.lane >= kNodeHeaderWords && W.lane < kNodeHeaderWords + sw)
                    w = reinterpret_cast<const uint32_t *>(C.spi0)[W.lane - kNodeHeaderWords];
                else if (W.lane >= kNodeHeaderWords + sw && W.lane < kNodeHeaderWords + 2 * sw)
                    w = reinterpret_cast<const uint32_t *>(C.pi0)[W.lane - kNodeHeaderWords - sw];
                node_at(L, 0, sw)[W.lane] = w;
            }
            wave_sync();
            if (cfg_int(C.triple_count) > 0) {
                PHASE_T0();
                int scount = 0;
                if (kLearned && W.lane == 0) L.ncarry = 0;
                wave_sync();
                expand_stage<kLearned, kServe>(P, W, L, C, node_at(L, 0, sw), 0, true, scount);
                if (scount && !W.stop) {
                    wave_sync();
                    flush_stage<kLearned, kServe>(P, W, L, C, scount, cfg_int(C.n), (int)P.spi_words, cfg_int(C.triple_count), true);
                }
                if (kLearned && !W.stop)
                    learned_push<kServe>(P, W, L, cfg_int(C.n), (int)P.spi_words, cfg_int(C.triple_count), false, 0.f, 0u, 0);
                PHASE_ADD(L, kPhaseRoot);
            }
            if (kServe && !W.stop) serve_flush(P, W, L);
            continue;
        } else if (P.donate) {
            PHASE_T0();
            if (!idle) {
                // count first, then advertise: a claim always finds the wave
                // already counted idle
                if (W.lane == 0) {
                    fetch_add(&P.ctrl[kCtlIdle].v, 1ULL);
                    __hip_atomic_fetch_or(idle_word(P, W.wave >> 6), 1ULL << (W.wave & 63), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                }
                idle = true;
                backoff = 1;
            }
            if (receive_nodes(P, W)) {
                // the claiming wave already took this wave out of the idle count
                idle = false;
                PHASE_ADD(L, kPhaseShare);
                continue;
            }
            // every 4th empty poll: drain the spill pool, or end the launch.
            // The launch is over when the pool is empty at the first read
            // (head == tail), every registered wave is idle at the second,
            // and neither head nor tail moved by the third: work leaves a
            // busy wave only through an inbox (the claim takes the receiver
            // out of the idle count first) or the pool (the tail moves), and
            // enters an idle wave from the pool only after the taker left
            // the idle count (then the head moves).
            if ((++idle_polls & 3) == 0) {
                int work = 0, done = 0;
                if (W.lane == 0) {
                    const unsigned long long h0 = ctl_load(P, kCtlPoolHead);
                    complete_after();
                    const unsigned long long t0 = ctl_load(P, kCtlPoolTail);
                    if (h0 < t0) {
                        work = 1;
                    } else {
                        complete_after();
                        const unsigned long long reg = ctl_load(P, kCtlRegistered);
                        complete_after();
                        const unsigned long long idl = ctl_load(P, kCtlIdle);
                        complete_after();
                        const unsigned long long h2 = ctl_load(P, kCtlPoolHead);
                        complete_after();
                        const unsigned long long t2 = ctl_load(P, kCtlPoolTail);
                        done = (!kServe && idl == reg && h2 == h0 && t2 == t0) ? 1 : 0;
                    }
                    if (work) {
                        // leave the idle set before taking: clearing the own
                        // bit decides between this take and a claim already
                        // made (whose nodes then arrive in the inbox)
                        const unsigned long long bit = 1ULL << (W.wave & 63);
                        work = (__hip_atomic_fetch_and(idle_word(P, W.wave >> 6), ~bit, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) & bit) ? 1 : 0;
                        if (work) fetch_add(&P.ctrl[kCtlIdle].v, ~0ULL);
                    }
                }
                if (lane0(done)) break;
                if (lane0(work)) {
                    idle = false;   // a failed take re-enters the idle set
                    have = take_one(P, W, L);
                    PHASE_ADD(L, kPhaseShare);
                    if (!have) continue;
                } else if (should_stop<kServe>(P)) {
                    break;
                }
            }
            if (!have) {
                for (int z = 0; z < backoff; ++z) __builtin_amdgcn_s_sleep(127);
                if (backoff < (int)P.idle_backoff) backoff *= 2;
                PHASE_ADD(L, kPhaseShare);
                continue;
            }
        } else {
            break;
        }
        if (!have) continue;

        // ---- expand the nodes 0..batch of the pool (one configuration): their
        // children are staged together and deduplicated in one pass
        const int cfg = uniform((int)node_at(L, 0, (int)P.spi_words)[0]);
        const ConfigDesc &C = cfg_copy[wib];
        {
            PHASE_T0();
            bool skip;
            if constexpr (kServe) {
                skip = serve_switch(P, W, L, cfg_copy[wib], cfg);
                lds_add(&L.takes, (unsigned long long)batch);   // the nodes just taken
            } else {
                switch_config(P, W, L, cfg_copy[wib], cfg);
                lds_add(&L.live, (unsigned long long)(-(long long)batch));   // the nodes just taken
                skip = config_found(P, cfg);
            }
            PHASE_ADD(L, kPhaseCheck);
            if (skip) {
                if (kServe) serve_flush(P, W, L);   // dropped: taken without children
                continue;
            }
        }
        if (W.lane < kNumCounters) L.pend[W.lane] = 0;
        if (kLearned && W.lane == 0) L.ncarry = 0;
        wave_sync();
        int scount = 0, expanded = 0, rest = 0;
        bool terminal = false;
        const uint32_t hi0 = W.hi;   // children flushed during the batch land above
        // Bottom node first.  Expanding several nodes before their children
        // are pushed reaches states at smaller depths first (fewer improved
        // re-expansions) and finds goals among siblings earlier; it also
        // leaves more nodes on the stack than DFS, so once the batch has
        // produced half a stage of children the nodes above go back onto the
        // stack, on top.
        for (int b = batch - 1; b >= 0 && !W.stop; --b) {
            rest = b;
            const uint32_t *nd = node_at(L, b, (int)P.spi_words);
            const int d = uniform((int)(nd[1] & 0xFF));
            const int e = uniform((int)((nd[1] >> 8) & 0xFF));
            if ((cfg_bits(C.accept[d]) >> e) & 1ULL) {
                report_found<kServe>(P, W, nd, cfg, d, -1);
                terminal = true;
                break;
            }
            if (d >= cfg_int(C.max_depth) || !((cfg_bits(C.reach[d]) >> e) & 1ULL)) continue;
            lds_add(&L.cnt[kCntExpansions], 1);
            lds_add(&L.cnt[kCntGenerated], (unsigned long long)cfg_int(C.triple_count));
            ++expanded;
            if (b >= 1 && cfg_int(C.n) <= kWave / 2) {
                // the next node too when it is expandable and both take the
                // +2-only enumeration: one pass for the pair
                const uint32_t *nd2 = node_at(L, b - 1, (int)P.spi_words);
                const int d2 = uniform((int)(nd2[1] & 0xFF));
                const int e2 = uniform((int)((nd2[1] >> 8) & 0xFF));
                uint32_t ga, aa, gb, ab;
                child_masks(C, d, e, false, ga, aa);
                child_masks(C, d2, e2, false, gb, ab);
                if ((ga | aa) == 4u && (gb | ab) == 4u && !((cfg_bits(C.accept[d2]) >> e2) & 1ULL) &&
                    d2 < cfg_int(C.max_depth) && ((cfg_bits(C.reach[d2]) >> e2) & 1ULL)) {
                    PHASE_T0();
                    int goal_b = -1, goal_t = -1;
                    const int r = pair_local_children<kLearned>(W, L, b, b - 1, cfg_int(C.n), (int)P.spi_words,
                                                      cfg_int(C.triple_count), e, e2, ga, aa, gb, ab, scount,
                                                      goal_b, goal_t);
                    PHASE_ADD(L, kPhaseScan);
                    if (r == kLocalDone) {
                        lds_add(&L.cnt[kCntExpansions], 1);
                        lds_add(&L.cnt[kCntGenerated], (unsigned long long)cfg_int(C.triple_count));
                        ++expanded;
                        --b;
                        rest = b;
                        if (goal_t >= 0) {
                            report_found<kServe>(P, W, node_at(L, goal_b, (int)P.spi_words), cfg, (goal_b == b ? d2 : d) + 1, goal_t);
                            terminal = true;
                            break;
                        }
                        if (scount + (int)(W.hi - hi0) >= kStage / 2) break;
                        continue;
                    }
                }
            }
            const int goal = expand_stage<kLearned, kServe>(P, W, L, C, nd, b, false, scount);
            if (goal >= 0) {
                report_found<kServe>(P, W, nd, cfg, d + 1, goal);
                terminal = true;   // the configuration is found: staged children are dropped
                break;
            }
            // half a stage of children (staged, or pushed by a flush inside
            // a wide expansion) ends the batch
            if (scount + (int)(W.hi - hi0) + (kLearned ? (int)L.ncarry : 0) >= kStage / 2) break;
        }
        // a terminal expansion's dedup side effects are not counted
        if (!terminal && !W.stop) {
            if (scount) {
                wave_sync();
                flush_stage<kLearned, kServe>(P, W, L, C, scount, cfg_int(C.n), (int)P.spi_words, cfg_int(C.triple_count), false);
            }
            if (kLearned && !W.stop)
                learned_push<kServe>(P, W, L, cfg_int(C.n), (int)P.spi_words, cfg_int(C.triple_count), false, 0.f, 0u, 0);
            // nodes of the batch not expanded return on top, the top one last
            for (int b = rest - 1; b >= 0 && !W.stop; --b) push_node<kServe>(P, W, L, node_at(L, b, (int)P.spi_words));
            wave_sync();
            if (W.lane < kNumCounters) L.cnt[W.lane] += L.pend[W.lane];
            wave_sync();
        }
        // serving: the batch's counters and its pushes minus its takes (a
        // terminal batch's unexpanded nodes and staged children are dropped)
        if (kServe && !W.stop) serve_flush(P, W, L);

        // donation: every share_mask + 1 expansions, a wave with at least two
        // nodes hands up to half of them (at most share_max) to an idle wave
        const unsigned iter0 = iter;
        iter += (unsigned)expanded;
        if (P.donate && ((iter0 ^ iter) & ~P.share_mask) != 0 && W.hi - W.lo >= 2) {
            PHASE_T0();
            uint32_t give = (W.hi - W.lo) / 2;
            if (give > P.share_max) give = P.share_max;
            share_nodes(P, W, give, iter);
            PHASE_ADD(L, kPhaseShare);
        }
    }
    flush_counters<kServe>(P, W, L);
#ifdef TDP_PHASE_TIMING
    if (W.lane < kNumPhases) fetch_add(&P.ctrl[kCtlPhase0 + W.lane].v, L.phase[W.lane]);
#endif
    if (W.lane == 0) {
        fetch_add(&P.ctrl[kCtlPops].v, L.traffic[0]);
        fetch_add(&P.ctrl[kCtlPushes].v, L.traffic[1]);
        fetch_add(&P.ctrl[kCtlProbes].v, L.traffic[2]);
        fetch_add(&P.ctrl[kCtlKeyReads].v, L.traffic[3]);
        fetch_add(&P.ctrl[kCtlKeyWrites].v, L.traffic[4]);
        fetch_add(&P.ctrl[kCtlEvictions].v, L.traffic[5]);
    }
}

// Everything is inlined into the kernels: a call would pass the by-value
// parameter block through scratch memory and LDS references as 64-bit flat
// pointers.  Three waves per SIMD (<= 168 VGPRs) costs the plain kernels a
// few spilled registers; four would spill heavily.  The learned-ordering
// variant is a kernel of its own so that the plain search carries no MFMA
// code; it runs at two waves per SIMD (<= 256 VGPRs), where its scoring
// state (MLP fragments, features, the sort's keys) needs no spills.
#ifndef TDP_WAVES_PER_EU
#define TDP_WAVES_PER_EU 3
#endif
#ifndef TDP_LEARNED_WAVES_PER_EU
#define TDP_LEARNED_WAVES_PER_EU 2
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TDP_WAVES_PER_EU))) void search_kernel(SearchParams P) {
    search_body<false, false>(P);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TDP_LEARNED_WAVES_PER_EU))) void search_learned_kernel(SearchParams P) {
    search_body<true, false>(P);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TDP_WAVES_PER_EU))) void serve_kernel(SearchParams P) {
    search_body<false, true>(P);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TDP_LEARNED_WAVES_PER_EU))) void serve_learned_kernel(SearchParams P) {
    search_body<true, true>(P);
}

// Reset of the launch state (replaces the per-search memsets of
// sorting_search.cu:744-777 with one small kernel).
__global__ void reset_kernel(SearchParams P) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < (size_t)kNumControl) P.ctrl[tid].v = 0;
    if (tid < (size_t)P.num_configs) {
        ConfigResult &r = P.results[tid];
        r.found = 0;
        r.depth = 0;
        for (int m = 0; m < 16; ++m) r.moves[m] = 0;
        for (int c = 0; c < kNumCounters; ++c) r.counters[c] = 0;
        r.live = 0;
    }
    if (tid < (size_t)P.num_waves && P.inbox) P.inbox[tid * 16] = 0;
    if (tid < (size_t)P.idle_words && P.idle_bits) P.idle_bits[tid * 16] = 0;
    if (tid < (size_t)P.pool_cap && P.pool_seq) P.pool_seq[tid] = tid;
}

// Debug/validation kernel: odd-cycle count of every child of one node,
// computed with the same LDS cycle analysis and parity delta as the search.
__global__ __launch_bounds__(64) void eval_children_kernel(const uint8_t *pi, const uint8_t *spi, int n, int T,
                                                           int e, int32_t *out, int e0, int d, int D,
                                                           uint32_t *features) {
    __shared__ WaveLds L;
    __shared__ uint16_t colex[kColexPairs];
    fill_colex(colex);
    const int lane = threadIdx.x;
    const int sw = spi_words_for(kMaxN);
    uint32_t w = 0;
    if (lane < sw) {
        for (int b = 0; b < 4; ++b) {
            const int s = 4 * lane + b;
            if (s < n) w |= (uint32_t)spi[s] << (8 * b);
        }
        node_at(L, 0, sw)[kNodeHeaderWords + lane] = w;
        w = 0;
        for (int b = 0; b < 4; ++b) {
            const int s = 4 * lane + b;
            if (s < n) w |= (uint32_t)pi[s] << (8 * b);
        }
        node_at(L, 0, sw)[kNodeHeaderWords + sw + lane] = w;
    }
    wave_sync();
    build_posinfo(L, node_at(L, 0, sw), 0, n, sw, lane);
    // same enumeration and delta as the search scan, written at t = T-1-o
    int mb = -1, qb = 0;
    scan_children(colex, n, T, lane, mb, qb, [&](bool valid, int i, int j, int k, int obase) -> int {
        const int delta = child_delta(L.posinfo[0][i], L.posinfo[0][j], L.posinfo[0][k]);
        const int t = T - 1 - (obase + lane);
        if (valid && out) out[t] = e + delta;
        if (valid && features) {
            // the learned ordering's features of this move, as the search computes them
            uint32_t f[kFeatWords];
            move_features(L.posinfo[0][i], L.posinfo[0][j], L.posinfo[0][k], i, j, k, n, d, D, e0, e + delta, delta,
                          f);
            for (int w = 0; w < kFeatWords; ++w) features[(size_t)t * kFeatWords + w] = f[w];
        }
        return kScanGo;
    });
}

// Validation: the cycle-local enumeration of two nodes (index 1 = a, 0 = b
// of a batch, same n <= 32, +2-only masks) by the paired pass and by one
// pass per node, as the batch loop would run them.  out[0..4) and
// out[4..4 + kStage) of each record: status, goal node, goal triple, staged
// count, then the stage entries; record 0 is the pair, record 1 the
// sequential passes.
__global__ __launch_bounds__(64) void eval_pair_kernel(const uint8_t *pia, const uint8_t *spia, const uint8_t *pib,
                                                      const uint8_t *spib, int n, int ea, int eb, uint32_t ga,
                                                      uint32_t aa, uint32_t gb, uint32_t ab, int32_t *out) {
    __shared__ WaveLds L;
    __shared__ uint16_t colex[kColexPairs];
    __shared__ uint32_t colex3[kCycleLocalMax];
    fill_colex(colex, colex3);
    WaveState W;
    W.lane = threadIdx.x;
    W.colex = colex;
    W.colex3 = colex3;
    const int sw = spi_words_for(kMaxN);
    const int T = n * (n - 1) * (n - 2) / 6;
    for (int b = 0; b < 2; ++b) {
        const uint8_t *pi = b ? pia : pib, *spi = b ? spia : spib;
        if (W.lane < sw) {
            uint32_t w = 0, v = 0;
            for (int q = 0; q < 4; ++q) {
                const int s = 4 * W.lane + q;
                if (s < n) w |= (uint32_t)spi[s] << (8 * q);
                if (s < n) v |= (uint32_t)pi[s] << (8 * q);
            }
            node_at(L, b, sw)[kNodeHeaderWords + W.lane] = w;
            node_at(L, b, sw)[kNodeHeaderWords + sw + W.lane] = v;
        }
    }
    wave_sync();
    for (int rec = 0; rec < 2; ++rec) {
        int32_t *o = out + rec * (4 + kStage);
        int scount = 0, goal_b = -1, goal_t = -1, st;
        if (rec == 0) {
            st = pair_local_children<false>(W, L, 1, 0, n, sw, T, ea, eb, ga, aa, gb, ab, scount, goal_b, goal_t);
        } else {
            uint32_t si = build_posinfo(L, node_at(L, 1, sw), 0, n, sw, W.lane);
            st = cycle_local_children(W, L, node_at(L, 1, sw), 1, 0, si, n, sw, T, ea, ga, aa, false, scount, goal_t);
            if (st == kLocalDone && goal_t >= 0) {
                goal_b = 1;
            } else if (st == kLocalDone) {
                si = build_posinfo(L, node_at(L, 0, sw), 0, n, sw, W.lane);
                st = cycle_local_children(W, L, node_at(L, 0, sw), 0, 0, si, n, sw, T, eb, gb, ab, false, scount, goal_t);
                if (goal_t >= 0) goal_b = 0;
            }
        }
        wave_sync();
        if (W.lane == 0) {
            o[0] = st;
            o[1] = goal_b;
            o[2] = goal_t;
            o[3] = scount;
        }
        for (int q = W.lane; q < scount; q += kWave) o[4 + q] = (int32_t)L.stage[q];
        wave_sync();
    }
}

}  // namespace tdp
