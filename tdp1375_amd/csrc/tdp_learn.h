// Learned move ordering: features of a candidate move and the scoring MLP
// (128 binary inputs -> 256 ReLU -> 1 logit, bf16 weights, fp32 accumulate)
// on gfx950 MFMA.  No counterpart in the reference, which orders children by
// triple index only (sorting_search.cu:533-535); the ordering changes which
// child a wave takes first, never which children exist, so verdicts are those
// of the plain search.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tdp_common.h"

namespace tdp {

constexpr int kFeatWords = 4;      // 128 feature bits
constexpr int kHidden = 256;
constexpr int kIn = 128;

// Model in device memory: W1 transposed (hidden-major, 128 bf16 per hidden
// unit) so that one MFMA B fragment is one 16-byte load.
struct MlpWeights {
    const __bf16 *w1t;   // [kHidden][kIn]
    const float *b1;     // [kHidden]
    const __bf16 *w2;    // [kHidden]
    const float *b2;     // [1]
};

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Feature bits of the move (i, j, k) of a node (layout in learned.py
// `FEATURE_LAYOUT`):
//   0-2    delta of the odd-cycle count (-2, 0, +2) one-hot
//   3-18   child depth d+1, one-hot (0..15)
//   19-34  child gain e' - e0, clamped to 0..15, one-hot
//   35-50  moves left after the child, D - (d+1), one-hot (0..15)
//   51-98  lengths of the cycles of a, b, c (min(len,16)-1), one-hot each
//   99-103 which symbols share a cycle: none, a=b only, b=c only, a=c only, all
//   104-106 parity of the spi-steps a->b, b->c, c->a inside a shared cycle
//   107    a -> b -> c follows the cycle (all three in one cycle)
//   108-123 n / 4 clamped to 15, one-hot
//   124-127 j = i+1, k = j+1, i = 0, k = n-1
// ia/ib/ic: cycle_info of the symbols at positions i, j, k.
__device__ __forceinline__ void move_features(uint32_t ia, uint32_t ib, uint32_t ic, int i, int j, int k, int n,
                                              int d, int D, int e0, int echild, int delta, uint32_t (&f)[kFeatWords]) {
    f[0] = f[1] = f[2] = f[3] = 0u;
    auto set = [&](int bit) { f[bit >> 5] |= 1u << (bit & 31); };
    set((delta + 2) >> 1);
    set(3 + min(max(d + 1, 0), 15));
    set(19 + min(max(echild - e0, 0), 15));
    set(35 + min(max(D - (d + 1), 0), 15));
    const int ra = ia & 63, rb = ib & 63, rc = ic & 63;
    const int la = (ia >> 16) & 127, lb = (ib >> 16) & 127, lc = (ic >> 16) & 127;
    set(51 + min(la, 16) - 1);
    set(67 + min(lb, 16) - 1);
    set(83 + min(lc, 16) - 1);
    const bool ab = ra == rb, bc = rb == rc, ac = ra == rc;
    const int pattern = (ab && bc) ? 4 : ab ? 1 : bc ? 2 : ac ? 3 : 0;
    set(99 + pattern);
    const int pa = (ia >> 8) & 63, pb = (ib >> 8) & 63, pc = (ic >> 8) & 63;
    if (ab) { int s = pb - pa; s += s < 0 ? la : 0; if (s & 1) set(104); }
    if (bc) { int s = pc - pb; s += s < 0 ? lb : 0; if (s & 1) set(105); }
    if (ac) { int s = pa - pc; s += s < 0 ? lc : 0; if (s & 1) set(106); }
    if (ab && bc) {
        int kab = pb - pa; kab += kab < 0 ? la : 0;
        int kbc = pc - pb; kbc += kbc < 0 ? la : 0;
        int kca = pa - pc; kca += kca < 0 ? la : 0;
        if (kab + kbc + kca == la) set(107);
    }
    set(108 + min(n >> 2, 15));
    if (j == i + 1) set(124);
    if (k == j + 1) set(125);
    if (i == 0) set(126);
    if (k == n - 1) set(127);
}

// A fragment of k-step ks (features 32 ks .. 32 ks + 31) for the row whose
// feature words are `f`: lane l supplies row l % 16, features 8 (l / 16) + t.
__device__ __forceinline__ bf16x8_t feature_fragment(const uint32_t (&f)[kFeatWords], int ks, int lane) {
    const uint32_t byte = (f[ks] >> (8 * (lane >> 4))) & 0xFFu;
    bf16x8_t a;
#pragma unroll
    for (int t = 0; t < 8; ++t) a[t] = (byte >> t) & 1u ? (__bf16)1.0f : (__bf16)0.0f;
    return a;
}

// Logits of 16 rows held by lanes (l % 16) of the wave: row r's feature words
// are read from lane `src_lane(r)` by the caller into fw (per lane, the row
// l % 16 of this tile).  Returns, in every lane, the logit of row
// 4 (l / 16) + i in out[i].
__device__ __forceinline__ void mlp_tile_logits(const MlpWeights &M, const uint32_t (&fw)[kFeatWords], int lane,
                                                float (&out)[4]) {
    bf16x8_t a[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a[ks] = feature_fragment(fw, ks, lane);
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    // A hidden block's four k-step weight fragments, bias and output weight
    // are loaded together before its products: one L2 round trip per block,
    // not one per fragment waited for before its product.
    const bf16x8_t *wrow = reinterpret_cast<const bf16x8_t *>(M.w1t + (size_t)(lane & 15) * kIn + 8 * (lane >> 4));
    constexpr int kRowStep = 16 * kIn / 8;   // bf16x8 units from one hidden block to the next
#pragma unroll 1
    for (int nt = 0; nt < kHidden / 16; ++nt) {
        bf16x8_t b[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) b[ks] = wrow[nt * kRowStep + 4 * ks];
        const float bias = M.b1[nt * 16 + (lane & 15)], w = (float)M.w2[nt * 16 + (lane & 15)];
        f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b[ks], c, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) y[i] += fmaxf(c[i] + bias, 0.f) * w;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = y[i];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        out[i] = v + *M.b2;
    }
}

constexpr int kMlpParams = kIn * kHidden + 2 * kHidden + 1;   // [W1 (k-major) | b1 | w2 | b2]

__global__ void mlp_forward_kernel(const uint32_t *X, int N, MlpWeights M, float *hpre, float *logits);
__global__ void group_softmax_kernel(const float *logits, const uint8_t *labels, const int32_t *gstart, int G,
                                     float scale, float loss_scale, float *dy, float *loss);
__global__ void mlp_dw1_kernel(const uint32_t *X, int N, const float *hpre, const float *dy, const __bf16 *w2,
                               int rows_per_chunk, float *dw1);
__global__ void mlp_column_kernel(int N, const float *hpre, const float *dy, const __bf16 *w2, float *db1,
                                  float *dw2, float *db2);
__global__ void mlp_adam_kernel(float *p, const float *g, float *m, float *v, int count, float lr, float b1c,
                                float b2c, __bf16 *w1t, float *b1, __bf16 *w2, float *b2);
__global__ void mlp_publish_kernel(const float *p, __bf16 *w1t, float *b1, __bf16 *w2, float *b2);

}  // namespace tdp
