// Training kernels of the move-ordering MLP (no reference counterpart; see
// tdp_learn.h).  Rows are candidate moves with 128 feature bits; rows of one
// search state form a group, exactly one row of a group being the move a
// found sequence took.  The loss is the softmax cross-entropy of each group's
// logits against that row, averaged over the groups.
//
//   forward   hpre = X W1 + b1 (MFMA 16x16x32 bf16, fp32 accumulate),
//             logit = relu(hpre) . w2 + b2
//   backward  dy = (softmax_group(logit) - label) / G
//             dw2 = relu(hpre)^T dy, db2 = sum dy
//             dhpre = dy w2^T * [hpre > 0], db1 = sum dhpre
//             dW1 = X^T dhpre (MFMA, dhpre rounded to bf16, fp32 accumulate)
//   update    Adam on fp32 master weights, bf16 copies for the MFMA passes
#include <hip/hip_runtime.h>

#include "tdp_learn.h"

namespace tdp {

// one wave per 16 rows; hpre (N x 256, fp32) kept for the backward pass
__global__ __launch_bounds__(256) void mlp_forward_kernel(const uint32_t *X, int N, MlpWeights M, float *hpre,
                                                          float *logits) {
    const int lane = threadIdx.x & 63;
    const int row0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 16;
    if (row0 >= N) return;
    uint32_t fw[kFeatWords];
    const int myrow = row0 + (lane & 15);
#pragma unroll
    for (int w = 0; w < kFeatWords; ++w) fw[w] = myrow < N ? X[(size_t)myrow * kFeatWords + w] : 0u;
    bf16x8_t a[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a[ks] = feature_fragment(fw, ks, lane);
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    for (int nt = 0; nt < kHidden / 16; ++nt) {
        const int col = nt * 16 + (lane & 15);
        f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(M.w1t + (size_t)col * kIn + 32 * ks + 8 * (lane >> 4));
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b, c, 0, 0, 0);
        }
        const float bias = M.b1[col], w = (float)M.w2[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = row0 + 4 * (lane >> 4) + i;
            const float h = c[i] + bias;
            if (hpre && r < N) hpre[(size_t)r * kHidden + col] = h;
            y[i] += fmaxf(h, 0.f) * w;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float v = y[i];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        const int r = row0 + 4 * (lane >> 4) + i;
        if ((lane & 15) == 0 && r < N) logits[r] = v + *M.b2;
    }
}

// one thread per group: dy = (softmax - label) * scale;
// loss += -log p_label * loss_scale
__global__ void group_softmax_kernel(const float *logits, const uint8_t *labels, const int32_t *gstart, int G,
                                     float scale, float loss_scale, float *dy, float *loss) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const int s = gstart[g], e = gstart[g + 1];
    float m = -3.0e38f;
    for (int r = s; r < e; ++r) m = fmaxf(m, logits[r]);
    float z = 0.f;
    for (int r = s; r < e; ++r) z += expf(logits[r] - m);
    float l = 0.f;
    int pos = 0;
    for (int r = s; r < e; ++r) {
        const float p = expf(logits[r] - m) / z;
        const float lab = labels[r] ? 1.f : 0.f;
        pos += labels[r] ? 1 : 0;
        dy[r] = (p - lab) * scale;
        if (labels[r]) l = -(logits[r] - m - logf(z));
    }
    if (pos != 1)
        for (int r = s; r < e; ++r) dy[r] = 0.f;   // malformed group: no gradient
    else
        atomicAdd(loss, l * loss_scale);
}

// dW1 tile (16 features x 16 hidden units) over a chunk of rows; one wave per
// (tile, chunk), partial sums added to dw1 (k-major, 128 x 256, zeroed).
__global__ __launch_bounds__(256) void mlp_dw1_kernel(const uint32_t *X, int N, const float *hpre, const float *dy,
                                                      const __bf16 *w2, int rows_per_chunk, float *dw1) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int tile = wave % 128, chunk = wave / 128;
    const int kt = tile / 16, jt = tile % 16;
    const int r_begin = chunk * rows_per_chunk;
    if (r_begin >= N) return;
    const int r_end = min(N, r_begin + rows_per_chunk);
    const int feat = kt * 16 + (lane & 15);
    const int j = jt * 16 + (lane & 15);
    const float wj = (float)w2[j];
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
    for (int r0 = r_begin; r0 < r_end; r0 += 32) {
        bf16x8_t a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = r0 + 8 * (lane >> 4) + t;
            const bool ok = r < r_end;
            const uint32_t word = ok ? X[(size_t)r * kFeatWords + (feat >> 5)] : 0u;
            a[t] = ((word >> (feat & 31)) & 1u) ? (__bf16)1.0f : (__bf16)0.0f;
            const float h = ok ? hpre[(size_t)r * kHidden + j] : 0.f;
            b[t] = (__bf16)((ok && h > 0.f) ? dy[r] * wj : 0.f);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&dw1[(size_t)(kt * 16 + 4 * (lane >> 4) + i) * kHidden + jt * 16 + (lane & 15)], c[i]);
}

// per hidden unit j: db1[j] = sum dhpre, dw2[j] = sum relu(hpre) dy; block 0
// thread 0 also adds db2 = sum dy.  One block of 256 threads per unit.
__global__ __launch_bounds__(256) void mlp_column_kernel(int N, const float *hpre, const float *dy, const __bf16 *w2,
                                                         float *db1, float *dw2, float *db2) {
    const int j = blockIdx.x;
    const float wj = (float)w2[j];
    float s1 = 0.f, s2 = 0.f, s3 = 0.f;
    for (int r = threadIdx.x; r < N; r += blockDim.x) {
        const float h = hpre[(size_t)r * kHidden + j];
        if (h > 0.f) {
            s1 += dy[r] * wj;
            s2 += h * dy[r];
        }
        if (j == 0) s3 += dy[r];
    }
    __shared__ float red[3][256];
    red[0][threadIdx.x] = s1;
    red[1][threadIdx.x] = s2;
    red[2][threadIdx.x] = s3;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        db1[j] = red[0][0];
        dw2[j] = red[1][0];
        if (j == 0) *db2 = red[2][0];
    }
}

// Adam on the flat fp32 parameter vector [W1 (k-major) | b1 | w2 | b2], then
// the bf16 copies used by the MFMA passes (W1 transposed) and b2.
__global__ void mlp_adam_kernel(float *p, const float *g, float *m, float *v, int count, float lr, float b1c,
                                float b2c, __bf16 *w1t, float *b1, __bf16 *w2, float *b2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const float gq = g[q];
    const float mq = 0.9f * m[q] + 0.1f * gq;
    const float vq = 0.999f * v[q] + 0.001f * gq * gq;
    m[q] = mq;
    v[q] = vq;
    const float val = p[q] - lr * (mq / b1c) / (sqrtf(vq / b2c) + 1e-8f);
    p[q] = val;
    const int nw1 = kIn * kHidden;
    if (q < nw1) {
        const int k = q / kHidden, j = q % kHidden;
        w1t[(size_t)j * kIn + k] = (__bf16)val;
    } else if (q < nw1 + kHidden) {
        b1[q - nw1] = val;
    } else if (q < nw1 + 2 * kHidden) {
        w2[q - nw1 - kHidden] = (__bf16)val;
    } else {
        *b2 = val;
    }
}

// bf16 copies from the fp32 parameter vector (after the host sets weights)
__global__ void mlp_publish_kernel(const float *p, __bf16 *w1t, float *b1, __bf16 *w2, float *b2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw1 = kIn * kHidden;
    if (q < nw1) {
        const int k = q / kHidden, j = q % kHidden;
        w1t[(size_t)j * kIn + k] = (__bf16)p[q];
    } else if (q < nw1 + kHidden) {
        b1[q - nw1] = p[q];
    } else if (q < nw1 + 2 * kHidden) {
        w2[q - nw1 - kHidden] = (__bf16)p[q];
    } else if (q == nw1 + 2 * kHidden) {
        *b2 = p[q];
    }
}

}  // namespace tdp
