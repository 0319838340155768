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

namespace tdp {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kStage = 512;
constexpr int kMaxProbes = 4096;

enum DedupOutcome { kDedupPending = -1, kDedupDominated = 0, kDedupImproved = 1,
                    kDedupInserted = 2, kDedupUninserted = 3 };

struct WaveLds {
    uint32_t node[48];
    uint32_t posinfo[64];
    uint32_t scratch[64];
    uint32_t stage[kStage];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_acquire(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long fetch_add(unsigned long long *p, unsigned long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas64(unsigned long long *p, unsigned long long expected,
                                      unsigned long long desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long ctl_load(const SearchParams &P, int idx) {
    return ld_relaxed(&P.ctrl[idx].v);
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
    if ((m.a >> 2) == q) w = put_byte(w, m.a, m.na);
    if ((m.b >> 2) == q) w = put_byte(w, m.b, m.nb);
    if ((m.c >> 2) == q) w = put_byte(w, m.c, m.nc);
    return w;
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
__device__ void build_posinfo(WaveLds &L, int n, int sw, int lane) {
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(L.node + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    const int s = lane;
    const int nxt = (s < n) ? (int)spi[s] : s;
    int m = s, p = nxt;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const int mo = __shfl(m, p);
        const int pp = __shfl(p, p);
        m = min(m, mo);
        p = pp;
    }
    const int root = m;
    int f = (s == root) ? 0 : 1;
    int q = (s == root) ? s : nxt;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const int fq = __shfl(f, q);
        const int qq = __shfl(q, q);
        f += fq;
        q = qq;
    }
    if (nxt == root) L.scratch[root] = (uint32_t)(f + 1);
    wave_sync();
    const int len = (int)L.scratch[root];
    const int pos = (f == 0) ? 0 : len - f;
    const uint32_t info = cycle_info(root, pos, len);
    const int x = (s < n) ? (int)pi[s] : s;
    L.posinfo[lane] = (uint32_t)__shfl((int)info, x);
    wave_sync();
}

struct Counters {
    unsigned long long v[kNumCounters];
};

struct WaveState {
    int wave, lane;
    uint32_t lo, hi;               // private stack window [lo, hi), wraps mod 2^32
    int cur_cfg;
    Counters cnt;
    bool configs_left;
    bool stop;
    unsigned long long deadline;   // s_memrealtime ticks (100 MHz); 0 = none
};

__device__ __forceinline__ bool past_deadline(const WaveState &W) {
    return W.deadline != 0 && __builtin_amdgcn_s_memrealtime() > W.deadline;
}

__device__ void set_watchdog(const SearchParams &P, WaveState &W) {
    if (W.lane == 0) __hip_atomic_store(&P.ctrl[kCtlStop].v, 2ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    W.stop = true;
}

__device__ void flush_counters(const SearchParams &P, WaveState &W) {
    if (W.cur_cfg >= 0) {
        if (W.lane < kNumCounters) {
            unsigned long long v = 0;
#pragma unroll
            for (int c = 0; c < kNumCounters; ++c) v = (W.lane == c) ? W.cnt.v[c] : v;
            if (v) fetch_add(&P.results[W.cur_cfg].counters[W.lane], v);
        }
    }
#pragma unroll
    for (int c = 0; c < kNumCounters; ++c) W.cnt.v[c] = 0;
}

__device__ __forceinline__ uint32_t *stack_slot(const SearchParams &P, const WaveState &W, uint32_t idx) {
    return P.stacks + ((size_t)W.wave * P.stack_cap + (size_t)(idx & (P.stack_cap - 1))) * P.node_words;
}

__device__ void set_overflow(const SearchParams &P, WaveState &W) {
    if (W.lane == 0) __hip_atomic_store(&P.ctrl[kCtlStop].v, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    W.stop = true;
}

// Moves the oldest node of the private stack into the shared pool.  Returns
// false when the pool is full.
__device__ bool donate_one(const SearchParams &P, WaveState &W) {
    unsigned long long ticket = 0;
    int ok = 0;
    if (W.lane == 0) {
        for (int attempt = 0; attempt < 64; ++attempt) {
            const unsigned long long t = ctl_load(P, kCtlPoolTail);
            const unsigned long long h = ctl_load(P, kCtlPoolHead);
            if (t - h >= P.pool_cap) break;
            if (cas64(&P.ctrl[kCtlPoolTail].v, t, t + 1)) { ticket = t; ok = 1; break; }
        }
    }
    ok = __shfl(ok, 0);
    if (!ok) return false;
    ticket = ((unsigned long long)(unsigned)__shfl((int)(ticket >> 32), 0) << 32) |
             (unsigned)__shfl((int)(ticket & 0xFFFFFFFFu), 0);
    const size_t slot = (size_t)(ticket & (P.pool_cap - 1));
    int late = 0;
    if (W.lane == 0) {
        while (ld_acquire(&P.pool_seq[slot]) != ticket) {
            if (past_deadline(W)) { late = 1; break; }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    if (__shfl(late, 0)) { set_watchdog(P, W); return false; }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t *src = stack_slot(P, W, W.lo);
    uint32_t *dst = P.pool_nodes + slot * P.node_words;
    if (W.lane < (int)P.node_words) dst[W.lane] = src[W.lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (W.lane == 0) st_release(&P.pool_seq[slot], ticket + 1);
    W.lo++;
    return true;
}

// Takes one node from the shared pool into L.node.
__device__ bool take_one(const SearchParams &P, WaveState &W, WaveLds &L) {
    unsigned long long ticket = 0;
    int ok = 0;
    if (W.lane == 0) {
        for (int attempt = 0; attempt < 8; ++attempt) {
            const unsigned long long h = ctl_load(P, kCtlPoolHead);
            const unsigned long long t = ctl_load(P, kCtlPoolTail);
            if (h >= t) break;
            if (cas64(&P.ctrl[kCtlPoolHead].v, h, h + 1)) { ticket = h; ok = 1; break; }
        }
    }
    ok = __shfl(ok, 0);
    if (!ok) return false;
    ticket = ((unsigned long long)(unsigned)__shfl((int)(ticket >> 32), 0) << 32) |
             (unsigned)__shfl((int)(ticket & 0xFFFFFFFFu), 0);
    const size_t slot = (size_t)(ticket & (P.pool_cap - 1));
    int late = 0;
    if (W.lane == 0) {
        while (ld_acquire(&P.pool_seq[slot]) != ticket + 1) {
            if (past_deadline(W)) { late = 1; break; }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    if (__shfl(late, 0)) { set_watchdog(P, W); return false; }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t *src = P.pool_nodes + slot * P.node_words;
    if (W.lane < (int)P.node_words) L.node[W.lane] = src[W.lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (W.lane == 0) {
        st_release(&P.pool_seq[slot], ticket + P.pool_cap);
        fetch_add(&P.ctrl[kCtlSteals].v, 1ULL);
    }
    wave_sync();
    return true;
}

// Ensures room for one push on the private stack.
__device__ bool make_room(const SearchParams &P, WaveState &W) {
    if (W.hi - W.lo < P.stack_cap) return true;
    if (P.donate) {
        for (int r = 0; r < 32 && W.hi - W.lo > 1; ++r)
            if (!donate_one(P, W)) break;
        if (W.stop) return false;
        if (W.hi - W.lo < P.stack_cap) return true;
    }
    set_overflow(P, W);
    return false;
}

// Pushes the child (t, i, j, k, even) of the node in L.node.
__device__ void push_child(const SearchParams &P, WaveState &W, WaveLds &L, int n, int sw,
                           int t, int i, int j, int k, int even) {
    if (!make_room(P, W)) return;
    const uint32_t *pw = L.node;
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(pw + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    const int d = (int)(pw[1] & 0xFF);
    ChildMove m;
    m.a = pi[i]; m.b = pi[j]; m.c = pi[k];
    m.na = spi[m.c]; m.nb = spi[m.a]; m.nc = spi[m.b];
    uint32_t *dst = stack_slot(P, W, W.hi);
    const int lane = W.lane;
    if (lane < (int)P.node_words) {
        uint32_t w = 0;
        if (lane == 0) {
            w = pw[0];
        } else if (lane == 1) {
            w = (uint32_t)(d + 1) | ((uint32_t)even << 8);
        } else if (lane < kNodeHeaderWords) {
            w = pw[lane];
            const int m0 = 2 * (lane - 2);
            if (m0 == d) w = (w & 0xFFFF0000u) | (uint32_t)t;
            else if (m0 + 1 == d) w = (w & 0xFFFFu) | ((uint32_t)t << 16);
        } else if (lane < kNodeHeaderWords + sw) {
            w = child_spi_word(pw + kNodeHeaderWords, lane - kNodeHeaderWords, m);
        } else if (lane < kNodeHeaderWords + 2 * sw) {
            const int q = lane - kNodeHeaderWords - sw;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int p = 4 * q + b;
                if (p < n) w |= (uint32_t)pi[transposed_src_index(p, i, j, k)] << (8 * b);
            }
        }
        dst[lane] = w;
    }
    W.hi++;
}

// Lane-parallel exact dedup (sorting_search.cu:261-298 semantics with an
// exact table): returns a DedupOutcome for active lanes.
__device__ int dedup_admit(const SearchParams &P, bool active, uint32_t cfg, const uint32_t *pw,
                           int nw, const ChildMove &m, int depth) {
    if (P.dd_cap == 0) return active ? kDedupUninserted : kDedupPending;
    uint64_t h = 0x9E3779B97F4A7C15ULL * (uint64_t)(cfg + 1);
    for (int q = 0; q < nw; ++q) {
        h ^= child_spi_word(pw, q, m);
        h *= 0xff51afd7ed558ccdULL;
        h ^= h >> 29;
    }
    h = fmix64(h);
    const unsigned long long epoch = P.epoch;
    const unsigned long long mytag = (epoch << 48) | (h & 0xFFFFFFFFFFFFULL);
    const uint32_t bword = ((uint32_t)epoch << 16) | (uint32_t)depth;
    unsigned long long slot = __umul64hi(h, P.dd_cap);
    int probes = 0;
    int res = active ? kDedupPending : kDedupDominated;
    while (__ballot(res == kDedupPending) != 0) {
        if (res == kDedupPending) {
            const unsigned long long tg = ld_relaxed(&P.dd_tag[slot]);
            bool advance = false;
            if ((tg >> 48) != epoch) {
                if (cas64(&P.dd_tag[slot], tg, mytag)) {
                    P.dd_cfg[slot] = cfg;
                    uint32_t *kw = P.dd_key + slot * 16;
                    for (int q = 0; q < nw; ++q) kw[q] = child_spi_word(pw, q, m);
                    __hip_atomic_store(&P.dd_best[slot], bword, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    res = kDedupInserted;
                }
            } else if (tg == mytag) {
                const uint32_t bw = __hip_atomic_load(&P.dd_best[slot], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                if ((bw >> 16) == (uint32_t)epoch) {
                    bool eq = P.dd_cfg[slot] == cfg;
                    const uint32_t *kw = P.dd_key + slot * 16;
                    for (int q = 0; q < nw && eq; ++q) eq = kw[q] == child_spi_word(pw, q, m);
                    if (eq) {
                        const uint32_t old = __hip_atomic_fetch_min(&P.dd_best[slot], bword, __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT);
                        res = ((old & 0xFFFFu) <= (uint32_t)depth) ? kDedupDominated : kDedupImproved;
                    } else {
                        advance = true;
                    }
                }
            } else {
                advance = true;
            }
            if (advance) {
                slot = (slot + 1 == P.dd_cap) ? 0 : slot + 1;
                if (++probes >= kMaxProbes) res = kDedupUninserted;
            }
        }
    }
    return res;
}

// Deduplicates and pushes the staged children (descending triple order).
// root_stage: children of the start node (not counted as admitted).
__device__ void flush_stage(const SearchParams &P, WaveState &W, WaveLds &L, int count, int n, int sw,
                            const uint32_t *tri, Counters &pending, bool root_stage) {
    const uint32_t cfg = L.node[0];
    const int d = (int)(L.node[1] & 0xFF);
    const uint8_t *spi = reinterpret_cast<const uint8_t *>(L.node + kNodeHeaderWords);
    const uint8_t *pi = spi + 4 * sw;
    for (int base = 0; base < count && !W.stop; base += kWave) {
        const int idx = base + W.lane;
        const bool act = idx < count;
        const uint32_t st = act ? L.stage[idx] : 0;
        const int t = (int)(st & 0xFFFF);
        const int ce = (int)(st >> 16);
        const uint32_t tp = act ? tri[t] : 0;
        const int i = tp & 0xFF, j = (tp >> 8) & 0xFF, k = (tp >> 16) & 0xFF;
        ChildMove m;
        m.a = pi[i]; m.b = pi[j]; m.c = pi[k];
        m.na = spi[m.c]; m.nb = spi[m.a]; m.nc = spi[m.b];
        const int r = dedup_admit(P, act, cfg, L.node + kNodeHeaderWords, sw, m, d + 1);
        const bool admit = act && r != kDedupDominated;
        if (!root_stage) {
            const unsigned long long dom = __popcll(__ballot(act && r == kDedupDominated));
            const unsigned long long imp = __popcll(__ballot(act && r == kDedupImproved));
            const unsigned long long ins = __popcll(__ballot(act && r == kDedupInserted));
            const unsigned long long uni = __popcll(__ballot(act && r == kDedupUninserted));
            pending.v[kCntDominated] += dom;
            pending.v[kCntImproved] += imp;
            pending.v[kCntInserted] += ins;
            pending.v[kCntUninserted] += uni;
            pending.v[kCntAdmitted] += imp + ins + uni;
        }
        unsigned long long pm = __ballot(admit);
        while (pm) {
            const int src = __ffsll((long long)pm) - 1;
            pm &= pm - 1;
            const int ct = __shfl(t, src), ci = __shfl(i, src), cj = __shfl(j, src), ck = __shfl(k, src);
            const int cce = __shfl(ce, src);
            push_child(P, W, L, n, sw, uniform(ct), uniform(ci), uniform(cj), uniform(ck), uniform(cce));
            if (W.stop) break;
        }
    }
}

__device__ void report_found(const SearchParams &P, WaveState &W, WaveLds &L, int cfg, int depth,
                             int extra_t) {
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
    won = __shfl(won, 0);
    if (won) {
        const int pd = (int)(L.node[1] & 0xFF);
        if (W.lane < pd) {
            const uint32_t w = L.node[2 + (W.lane >> 1)];
            P.results[cfg].moves[W.lane] = (W.lane & 1) ? (w >> 16) : (w & 0xFFFF);
        } else if (W.lane == pd && extra_t >= 0) {
            P.results[cfg].moves[W.lane] = (uint32_t)extra_t;
        }
    }
}

// Evaluates every child of L.node.  Normal mode: returns the goal triple
// (terminal expansion) or -1 after staging/pushing admissible children.
// Root mode: stages and pushes every live root.
__device__ int expand(const SearchParams &P, WaveState &W, WaveLds &L, const ConfigDesc &C,
                      bool root_mode) {
    const int n = (int)C.n;
    const int sw = (int)P.spi_words;
    const int T = (int)C.triple_count;
    const int D = (int)C.max_depth;
    const int d = (int)(L.node[1] & 0xFF);
    const int e = (int)((L.node[1] >> 8) & 0xFF);
    const uint32_t *tri = P.triples + C.triple_offset;
    build_posinfo(L, n, sw, W.lane);
    const uint64_t acc = C.accept[d + 1];
    const uint64_t rch = (d + 1 < D) ? C.reach[d + 1] : 0ULL;
    Counters pending;
#pragma unroll
    for (int c = 0; c < kNumCounters; ++c) pending.v[c] = 0;
    int scount = 0;
    int goal_t = -1;
    for (int base = T - 1; base >= 0; base -= kWave) {
        const int t = base - W.lane;
        const bool valid = t >= 0;
        const uint32_t tp = valid ? tri[t] : 0;
        const int ce = e + parity_delta(L.posinfo[tp & 0xFF], L.posinfo[(tp >> 8) & 0xFF],
                                        L.posinfo[(tp >> 16) & 0xFF]);
        const bool goal = valid && ((acc >> ce) & 1ULL);
        const bool adm = valid && ((rch >> ce) & 1ULL);
        if (!root_mode) {
            const unsigned long long gm = __ballot(goal);
            if (gm) {
                goal_t = base - (__ffsll((long long)gm) - 1);
                break;
            }
        }
        const unsigned long long am = __ballot(root_mode ? (goal || adm) : adm);
        if (am) {
            const unsigned long long below = (W.lane == 0) ? 0ULL : (am & (~0ULL >> (64 - W.lane)));
            if ((root_mode ? (goal || adm) : adm))
                L.stage[scount + __popcll(below)] = (uint32_t)t | ((uint32_t)ce << 16);
            scount += __popcll(am);
            if (root_mode) W.cnt.v[kCntRootsLive] += __popcll(am);
            if (scount > kStage - kWave) {
                wave_sync();
                flush_stage(P, W, L, scount, n, sw, tri, pending, root_mode);
                scount = 0;
                if (W.stop) return -1;
            }
        }
    }
    if (goal_t >= 0) return goal_t;
    if (scount) {
        wave_sync();
        flush_stage(P, W, L, scount, n, sw, tri, pending, root_mode);
    }
    if (!root_mode) {
#pragma unroll
        for (int c = 0; c < kNumCounters; ++c) W.cnt.v[c] += pending.v[c];
    }
    return -1;
}

__device__ void switch_config(const SearchParams &P, WaveState &W, int cfg) {
    if (cfg != W.cur_cfg) {
        flush_counters(P, W);
        W.cur_cfg = cfg;
    }
}

__device__ bool config_found(const SearchParams &P, int cfg) {
    return __hip_atomic_load(&P.results[cfg].found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ bool should_stop(const SearchParams &P) {
    int s = 0;
    if ((threadIdx.x & 63) == 0) {
        s = ctl_load(P, kCtlStop) != 0;
        if (P.cancel_flag)
            s |= __hip_atomic_load(P.cancel_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        s |= ctl_load(P, kCtlResolved) >= (unsigned long long)P.num_configs;
    }
    return __shfl(s, 0) != 0;
}

}  // namespace

__global__ __launch_bounds__(256) void search_kernel(SearchParams P) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const int wib = threadIdx.x >> 6;
    WaveLds &L = lds[wib];
    WaveState W;
    W.lane = threadIdx.x & 63;
    W.wave = blockIdx.x * kWavesPerBlock + wib;
    if (W.wave >= P.num_waves) return;
    W.lo = W.hi = 0;
    W.cur_cfg = -1;
    W.configs_left = true;
    W.stop = false;
    W.deadline = P.timeout_ticks ? __builtin_amdgcn_s_memrealtime() + P.timeout_ticks : 0ULL;
#pragma unroll
    for (int c = 0; c < kNumCounters; ++c) W.cnt.v[c] = 0;
    if (W.lane == 0) fetch_add(&P.ctrl[kCtlRegistered].v, 1ULL);
    bool idle = false;
    unsigned iter = 0;

    while (true) {
        if (should_stop(P) || W.stop) break;
        if (past_deadline(W)) {
            set_watchdog(P, W);
            break;
        }
        bool have = false;
        if (W.hi > W.lo) {
            W.hi--;
            const uint32_t *src = stack_slot(P, W, W.hi);
            if (W.lane < (int)P.node_words) L.node[W.lane] = src[W.lane];
            wave_sync();
            have = true;
        } else if (W.configs_left) {
            int c = 0;
            if (W.lane == 0) c = (int)fetch_add(&P.ctrl[kCtlNextConfig].v, 1ULL);
            c = uniform(__shfl(c, 0));
            if (c >= P.num_configs) {
                W.configs_left = false;
                continue;
            }
            // start node of configuration c, expanded in root mode
            switch_config(P, W, c);
            const ConfigDesc &C = P.cfgs[c];
            const int sw = (int)P.spi_words;
            if (W.lane < (int)P.node_words) {
                uint32_t w = 0;
                if (W.lane == 0) w = (uint32_t)c;
                else if (W.lane == 1) w = C.e_start << 8;
                else if (W.lane >= kNodeHeaderWords && W.lane < kNodeHeaderWords + sw)
                    w = reinterpret_cast<const uint32_t *>(C.spi0)[W.lane - kNodeHeaderWords];
                else if (W.lane >= kNodeHeaderWords + sw && W.lane < kNodeHeaderWords + 2 * sw)
                    w = reinterpret_cast<const uint32_t *>(C.pi0)[W.lane - kNodeHeaderWords - sw];
                L.node[W.lane] = w;
            }
            wave_sync();
            if (C.triple_count > 0) expand(P, W, L, C, true);
            continue;
        } else if (P.donate) {
            if (!idle) {
                if (W.lane == 0) fetch_add(&P.ctrl[kCtlIdle].v, 1ULL);
                idle = true;
            }
            const unsigned long long takes0 = ctl_load(P, kCtlTakes);
            const unsigned long long head = ctl_load(P, kCtlPoolHead);
            const unsigned long long tail = ctl_load(P, kCtlPoolTail);
            if (head < tail) {
                if (W.lane == 0) fetch_add(&P.ctrl[kCtlIdle].v, ~0ULL);
                idle = false;
                if (take_one(P, W, L)) {
                    if (W.lane == 0) fetch_add(&P.ctrl[kCtlTakes].v, 1ULL);
                    have = true;
                } else {
                    continue;
                }
            } else {
                int done = 0;
                if (W.lane == 0) {
                    const unsigned long long idl = ctl_load(P, kCtlIdle);
                    const unsigned long long reg = ctl_load(P, kCtlRegistered);
                    const unsigned long long h2 = ctl_load(P, kCtlPoolHead);
                    const unsigned long long t2 = ctl_load(P, kCtlPoolTail);
                    const unsigned long long takes1 = ctl_load(P, kCtlTakes);
                    done = (idl == reg && h2 == t2 && takes1 == takes0) ? 1 : 0;
                }
                if (__shfl(done, 0)) break;
                __builtin_amdgcn_s_sleep(8);
                continue;
            }
        } else {
            break;
        }
        if (!have) continue;

        // ---- process the node in L.node
        const int cfg = uniform((int)L.node[0]);
        switch_config(P, W, cfg);
        if (config_found(P, cfg)) continue;
        const ConfigDesc &C = P.cfgs[cfg];
        const int d = uniform((int)(L.node[1] & 0xFF));
        const int e = uniform((int)((L.node[1] >> 8) & 0xFF));
        if ((C.accept[d] >> e) & 1ULL) {
            report_found(P, W, L, cfg, d, -1);
            continue;
        }
        if (d >= (int)C.max_depth || !((C.reach[d] >> e) & 1ULL)) continue;
        W.cnt.v[kCntExpansions]++;
        W.cnt.v[kCntGenerated] += C.triple_count;
        const int goal = expand(P, W, L, C, false);
        if (goal >= 0) report_found(P, W, L, cfg, d + 1, goal);

        // donation: give the oldest half of the stack to idle waves
        if (P.donate && ((++iter & 3) == 0) && W.hi - W.lo >= 2) {
            unsigned long long idl = 0;
            if (W.lane == 0) idl = ctl_load(P, kCtlIdle);
            idl = (unsigned long long)__shfl((int)idl, 0);
            if (idl > 0) {
                uint32_t give = (W.hi - W.lo) / 2;
                if (give > idl) give = (uint32_t)idl;
                if (give > 16) give = 16;
                for (uint32_t g = 0; g < give; ++g)
                    if (!donate_one(P, W)) break;
                if (W.lane == 0) fetch_add(&P.ctrl[kCtlDonations].v, (unsigned long long)give);
            }
        }
    }
    flush_counters(P, W);
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
    }
    if (tid < (size_t)P.pool_cap) P.pool_seq[tid] = tid;
}

// Debug/validation kernel: odd-cycle count of every child of one node,
// computed with the same LDS cycle analysis and parity delta as the search.
__global__ __launch_bounds__(64) void eval_children_kernel(const uint8_t *pi, const uint8_t *spi, int n,
                                                           const uint32_t *tri, int T, int e,
                                                           int32_t *out) {
    __shared__ WaveLds L;
    const int lane = threadIdx.x;
    const int sw = spi_words_for(kMaxN);
    uint32_t w = 0;
    if (lane < sw) {
        for (int b = 0; b < 4; ++b) {
            const int s = 4 * lane + b;
            if (s < n) w |= (uint32_t)spi[s] << (8 * b);
        }
        L.node[kNodeHeaderWords + lane] = w;
        w = 0;
        for (int b = 0; b < 4; ++b) {
            const int s = 4 * lane + b;
            if (s < n) w |= (uint32_t)pi[s] << (8 * b);
        }
        L.node[kNodeHeaderWords + sw + lane] = w;
    }
    wave_sync();
    build_posinfo(L, n, sw, lane);
    for (int t = lane; t < T; t += kWave) {
        const uint32_t tp = tri[t];
        out[t] = e + parity_delta(L.posinfo[tp & 0xFF], L.posinfo[(tp >> 8) & 0xFF],
                                  L.posinfo[(tp >> 16) & 0xFF]);
    }
}

}  // namespace tdp
