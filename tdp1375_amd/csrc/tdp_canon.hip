// Configuration canonicalisation on the GPU: the signature of every rotation
// of pi and of its mirror (2n equivalent signatures), the least by CRC32
// hash then content, and the configuration that signature stands for.
//
// Follows Configuration.java (reference): signature :53-101, ofSignature
// :103-157, getCanonical :159-176, getEquivalentSignatures / mirror
// :182-198, Signature.hashCode :283-297; areSymbolsInCyclicOrder is
// CommonOperations.java:167-182.
//
// One wave per configuration (n <= 64): lanes are symbols for the cycle
// analysis and positions for the signature; the 2n signatures are kept in
// LDS (dynamic, sized by the batch's largest n: 10 KiB at n = 36 where a
// fixed n = 64 array took 32 KiB and held a CU to four waves) and every lane
// hashes its own candidates.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tdp_canon.h"

namespace tdp {
namespace {

constexpr int kW = 64;
constexpr int kMaxCand = 2 * kMaxCanonN;

struct CanonLds {
    uint32_t crc_table[256];
    uint32_t root[kW], cpos[kW], clen[kW];   // cycle of each symbol: min symbol, steps from it, length
    uint32_t cnt[kW];                  // per cycle root: ascents along the cycle
    uint32_t first[kW];                // per cycle root: first position in pi
    uint32_t minp[kW];                 // per cycle root: steps of its first-positioned member
    uint32_t label[kW];                // per cycle root: label
    uint32_t pos[kW];                  // position of each symbol in the rotated pi
    uint32_t perm[kW];                 // the permutation being signed (one-line)
    uint32_t inv[kW];                  // spi^-1
    uint32_t sym[kW];                  // symbol at each position of pi
    int32_t hash[kMaxCand];
    uint32_t lab_of[kW], rank_of[kW];  // canonical signature: label and rank per position
};

__device__ __forceinline__ void sync_wave() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int scan_add(int v) {
    // inclusive prefix sum over the 64 lanes
#pragma unroll
    for (int off = 1; off < kW; off <<= 1) {
        const int o = __shfl_up(v, off);
        if ((threadIdx.x & 63) >= off) v += o;
    }
    return v;
}

// Cycle analysis of L.perm over n symbols (lane = symbol): L.root, L.cpos,
// L.clen by pointer jumping.
__device__ __forceinline__ void analyse(CanonLds &L, int n, int lane) {
    const int nxt = lane < n ? (int)L.perm[lane] : lane;
    int m = lane, p = nxt;
    for (int r = 0; r < 6; ++r) {
        const int mo = __shfl(m, p);
        const int pp = __shfl(p, p);
        m = min(m, mo);
        p = pp;
    }
    const int root = m;
    int f = lane == root ? 0 : 1;
    int q = lane == root ? lane : nxt;
    for (int r = 0; r < 6; ++r) {
        const int fq = __shfl(f, q);
        const int qq = __shfl(q, q);
        f += fq;
        q = qq;
    }
    const int f_next = __shfl(f, nxt);
    const int len = __shfl(f_next + 1, root);
    L.root[lane] = (uint32_t)root;
    L.cpos[lane] = (uint32_t)(f == 0 ? 0 : len - f);
    L.clen[lane] = (uint32_t)len;
    sync_wave();
}

// Signature of (L.perm, pi rotated by r) into candidate c of `sig` (rows of
// `stride` floats); L.pos holds the positions in the rotated pi (lane =
// symbol) and `at` the symbol at this lane's position.
__device__ __forceinline__ void sign(CanonLds &L, float *sig, int stride, int n, int lane, int at, int c) {
    if (lane < n) {
        L.cnt[lane] = 0;
        L.first[lane] = 0xFFFFFFFFu;
    }
    sync_wave();
    if (lane < n) {
        const uint32_t rt = L.root[lane];
        const uint32_t ps = L.pos[lane];
        // an ascent of the position from this symbol to its image; a cycle
        // with two or more is not in the cyclic order of pi^-1 (oriented)
        atomicAdd(&L.cnt[rt], L.pos[L.perm[lane]] > ps ? 1u : 0u);
        atomicMin(&L.first[rt], ps);
    }
    sync_wave();
    if (lane < n && L.pos[lane] == L.first[L.root[lane]]) L.minp[L.root[lane]] = L.cpos[lane];
    // labels by first appearance along pi
    const uint32_t rt = lane < n ? L.root[at] : 0u;
    const bool first = lane < n && L.first[rt] == (uint32_t)lane;
    const int lbl = scan_add(first ? 1 : 0);
    if (first) L.label[rt] = (uint32_t)lbl;
    sync_wave();
    if (lane < n) {
        const float label = (float)L.label[rt];
        float v = label;
        if (L.cnt[rt] >= 2u) {
            const uint32_t len = L.clen[at];
            const uint32_t k = (L.cpos[at] + len - L.minp[rt]) % len + 1u;
            v = label + (float)k / 100.0f;
        }
        sig[c * stride + lane] = v;
    }
    sync_wave();
}

// stride: the batch's largest n; the launch gives 2 * stride * stride floats
// of dynamic LDS for the candidates.
__global__ __launch_bounds__(64) void canon_kernel(const uint8_t *spis, const uint8_t *pis, const int32_t *ns,
                                                    int count, int stride, float *out_sig, int32_t *out_hash,
                                                    int32_t *out_index, uint8_t *out_spi) {
    __shared__ CanonLds L;
    extern __shared__ float sig[];
    const int lane = threadIdx.x;
    // CRC-32 (reflected polynomial 0xEDB88320), the table once per block
    for (int b = lane; b < 256; b += kW) {
        uint32_t r = (uint32_t)b;
        for (int k = 0; k < 8; ++k) r = (r >> 1) ^ ((r & 1u) ? 0xEDB88320u : 0u);
        L.crc_table[b] = r;
    }
    const int cfg = blockIdx.x;
    if (cfg >= count) return;
    const int n = ns[cfg];
    const uint8_t *spi = spis + (size_t)cfg * kMaxCanonN;
    const uint8_t *pi = pis + (size_t)cfg * kMaxCanonN;
    if (lane < n) {
        L.sym[lane] = pi[lane];
        L.inv[spi[lane]] = (uint32_t)lane;
    }
    sync_wave();
    const int nc = 2 * n;
    for (int r = 0; r < n; ++r) {
        // pi rotated to start at its r-th symbol
        const int at = lane < n ? (int)L.sym[(lane + r) % n] : 0;
        if (lane < n) L.pos[at] = (uint32_t)lane;
        // the configuration itself
        if (lane < n) L.perm[lane] = spi[lane];
        sync_wave();
        if (r == 0) analyse(L, n, lane);   // the same cycles for every rotation
        else sync_wave();
        sign(L, sig, stride, n, lane, at, 2 * r);
        // its mirror: spi conjugated by tau = prod (pi_r[i] pi_r[n-1-i]),
        // then inverted: x -> tau(spi^-1(tau(x)))
        if (lane < n) {
            const int tx = (int)L.sym[(n - 1 - (int)L.pos[lane] + r) % n];   // tau(lane)
            const int y = (int)L.inv[tx];
            L.perm[lane] = L.sym[(n - 1 - (int)L.pos[y] + r) % n];          // tau(y)
        }
        sync_wave();
        analyse(L, n, lane);
        sign(L, sig, stride, n, lane, at, 2 * r + 1);
        // the cycles of spi again for the next rotation
        if (lane < n) L.perm[lane] = spi[lane];
        sync_wave();
        analyse(L, n, lane);
    }
    // every lane hashes its candidates: CRC32 of the big-endian float bytes
    for (int c = lane; c < nc; c += kW) {
        uint32_t crc = 0xFFFFFFFFu;
        for (int i = 0; i < n; ++i) {
            const uint32_t b = __float_as_uint(sig[c * stride + i]);
#pragma unroll
            for (int s = 24; s >= 0; s -= 8) crc = L.crc_table[(crc ^ (b >> s)) & 0xFFu] ^ (crc >> 8);
        }
        L.hash[c] = (int32_t)(crc ^ 0xFFFFFFFFu);
    }
    sync_wave();
    // the least hash, ties by the least content (Float.compare order, which
    // for these positive values is the order of the bit patterns)
    int best = 0;
    if (lane == 0) {
        for (int c = 1; c < nc; ++c) {
            const int32_t h = L.hash[c], hb = L.hash[best];
            bool less = h < hb;
            if (h == hb) {
                for (int i = 0; i < n; ++i) {
                    const float a = sig[c * stride + i], bb = sig[best * stride + i];
                    if (a != bb) {
                        less = a < bb;
                        break;
                    }
                }
            }
            if (less) best = c;
        }
    }
    best = __shfl(best, 0);
    if (lane < n) out_sig[(size_t)cfg * kMaxCanonN + lane] = sig[best * stride + lane];
    if (lane == 0) {
        out_hash[cfg] = L.hash[best];
        out_index[cfg] = best;
    }
    // the configuration of that signature (pi the identity): an oriented
    // cycle in rank order, any other by decreasing position
    if (lane < n) {
        const float v = sig[best * stride + lane];
        const uint32_t lab = (uint32_t)v;
        const bool oriented = v != (float)lab;
        L.lab_of[lane] = lab;
        L.rank_of[lane] = oriented ? (uint32_t)__float2int_rd((v - (float)lab) * 100.0f + 0.5f) : 0u;
    }
    sync_wave();
    if (lane < n) {
        const uint32_t lab = L.lab_of[lane], rk = L.rank_of[lane];
        int next = -1, wrap = -1;
        for (int j = 0; j < n; ++j) {
            if (L.lab_of[j] != lab) continue;
            if (rk) {
                if (L.rank_of[j] == rk + 1) next = j;
                if (L.rank_of[j] == 1) wrap = j;
            } else {
                if (j < lane) next = j;          // the largest smaller position
                wrap = j;                        // the largest position
            }
        }
        out_spi[(size_t)cfg * kMaxCanonN + lane] = (uint8_t)(next >= 0 ? next : wrap);
    }
}

}  // namespace

void launch_canon(const uint8_t *spis, const uint8_t *pis, const int32_t *ns, int count, int max_n, float *out_sig,
                  int32_t *out_hash, int32_t *out_index, uint8_t *out_spi, hipStream_t stream) {
    const int stride = max_n < 1 ? 1 : max_n;
    const size_t lds = (size_t)2 * stride * stride * sizeof(float);
    hipLaunchKernelGGL(canon_kernel, dim3((unsigned)count), dim3(kW), lds, stream, spis, pis, ns, count, stride,
                       out_sig, out_hash, out_index, out_spi);
}

}  // namespace tdp
