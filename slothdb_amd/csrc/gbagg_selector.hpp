// gbagg_selector.hpp -- the learned strategy selector.
//
// A 32 -> 64 -> 64 -> 64 -> 5 ReLU MLP (FP32) reads features of the
// heavy-hitter sample (the strided sample every AUTO plan takes first) and
// predicts the distinct-group count, as log2((distinct + 1) / (default + 1)) / 4
// over the sample's own estimate, and four path logits
// (DIRECT, LDS_HASH, GLOBAL_HASH, RADIX). The operator uses it to pick the path and
// to size the hash table; results never depend on it (every path is exact and
// retries on table overflow), and a prediction that disagrees with the
// sample's own estimate by more than 4x is ignored in favour of the default
// plan.
//
//   k_sel_plan      one block: features from the sample's statistics (gathered by
//                   k_heavy_hist), then the forward pass, into the plan block
//   k_sel_infer     one block of 64 per input row: forward pass
//   k_sel_train     one block: minibatch Adam over the whole training set,
//                   forward and backward passes with activations in LDS
#pragma once

#include "gbagg_radix.hpp"

namespace gbagg {

constexpr int SEL_IN = 32, SEL_H = 64, SEL_OUT = 5;
constexpr int SEL_CLS = SEL_OUT - 1;  // path logits
// parameter layout (floats): W1[H][IN] b1[H] W2[H][H] b2[H] W3[H][H] b3[H]
// W4[OUT][H] b4[OUT], then the input standardisation mu[IN] sigma[IN]
constexpr int SEL_OFF_W1 = 0;
constexpr int SEL_OFF_B1 = SEL_OFF_W1 + SEL_H * SEL_IN;
constexpr int SEL_OFF_W2 = SEL_OFF_B1 + SEL_H;
constexpr int SEL_OFF_B2 = SEL_OFF_W2 + SEL_H * SEL_H;
constexpr int SEL_OFF_W3 = SEL_OFF_B2 + SEL_H;
constexpr int SEL_OFF_B3 = SEL_OFF_W3 + SEL_H * SEL_H;
constexpr int SEL_OFF_W4 = SEL_OFF_B3 + SEL_H;
constexpr int SEL_OFF_B4 = SEL_OFF_W4 + SEL_OUT * SEL_H;
constexpr int SEL_NW = SEL_OFF_B4 + SEL_OUT;  // trained parameters
constexpr int SEL_OFF_MU = SEL_NW;
constexpr int SEL_OFF_SIG = SEL_OFF_MU + SEL_IN;
constexpr int SEL_NPARAM = SEL_OFF_SIG + SEL_IN;

// ---- features -----------------------------------------------------------
// feat[0] log2(rows+1)/32        feat[1] log2(n+1)/32 (sampled rows)
// feat[2] log2(d+1)/32 (distinct in the sample)   feat[3] d/n
// feat[4..6] f1/d f2/d f3/d (keys seen once, twice, three times)
// feat[7] log2(max+1)/32         feat[8] max/n
// feat[9] sample entropy / log2(n)
// feat[10..25] share of sampled keys with count in [2^j, 2^(j+1)), j = 0..15
// feat[26] log2(GEE+1)/32, GEE = d - f1 + f1 sqrt(rows/n)
// feat[27] log2(Chao1+1)/32, Chao1 = d + f1^2 / (2 max(f2, 1))
// feat[28] key bytes / 16 (4, 8 or 16)    feat[29] two key columns
// feat[30] log2(span + 1)/64, span = largest - smallest sampled key word
//          (single-column keys; 1 otherwise)
// feat[31] log2(rows/n + 1)/32
__device__ __forceinline__ void sel_layer(const float *W, const float *b, const float *in, int nin, float *out,
                                          bool relu);

// The sample's features from the statistics k_heavy_hist gathered in the
// plan block, and with parameters P (may be null) the MLP's outputs, both
// into the plan block (one block of SEL_H threads; no scan of the sample
// table, no copy of its own: the host reads them with the plan).
__global__ __launch_bounds__(SEL_H) void k_sel_plan(DevPlan *plan, uint64_t rows, uint32_t key_kind,
                                                    const float *__restrict__ P) {
    __shared__ float x[SEL_IN], h1[SEL_H], h2[SEL_H], h3[SEL_H];
    if (threadIdx.x == 0) {
        float *feat = plan->feat;
        const uint64_t d = plan->sample_distinct, f1 = plan->sample_f1, f2 = plan->sample_f2, f3 = plan->sample_f3;
        const uint64_t mx = plan->sample_max;
        const double n = (double)max<uint64_t>(plan->sample_rows, 1), D = (double)max<uint64_t>(d, 1);
        const double R = (double)rows;
        const double ent = (double)plan->sample_ent / 1048576.0;
        const double gee = (double)(d - f1) + (double)f1 * sqrt(R / n);
        const double chao = (double)d + (double)f1 * (double)f1 / (2.0 * (double)max<uint64_t>(f2, 1));
        feat[0] = (float)(log2(R + 1) / 32);
        feat[1] = (float)(log2(n + 1) / 32);
        feat[2] = (float)(log2((double)d + 1) / 32);
        feat[3] = (float)((double)d / n);
        feat[4] = (float)(f1 / D);
        feat[5] = (float)(f2 / D);
        feat[6] = (float)(f3 / D);
        feat[7] = (float)(log2((double)mx + 1) / 32);
        feat[8] = (float)((double)mx / n);
        feat[9] = (float)((log2(n) - ent / n) / max(1.0, log2(n)));
        for (int j = 0; j < 16; j++) feat[10 + j] = (float)(plan->sample_hist[j] / D);
        feat[26] = (float)(log2(gee + 1) / 32);
        feat[27] = (float)(log2(chao + 1) / 32);
        // key_kind: 0 one 4-byte column, 1 one 8-byte column, 2 two 4-byte
        // columns, 3 a 16-byte key
        feat[28] = key_kind == 0 ? 0.25f : key_kind == 3 ? 1.f : 0.5f;
        feat[29] = key_kind >= 2;
        const int64_t lo = plan->key_min, hi = plan->key_max;  // sampled key word range
        const double span = lo <= hi ? (double)((unsigned long long)hi - (unsigned long long)lo) : 0.0;
        feat[30] = key_kind <= 1 ? (float)(log2(span + 1) / 64) : 1.f;
        feat[31] = (float)(log2(R / n + 1) / 32);
    }
    __syncthreads();
    if (!P) return;
    if (threadIdx.x < SEL_IN)
        x[threadIdx.x] = (plan->feat[threadIdx.x] - P[SEL_OFF_MU + threadIdx.x]) / P[SEL_OFF_SIG + threadIdx.x];
    __syncthreads();
    sel_layer(P + SEL_OFF_W1, P + SEL_OFF_B1, x, SEL_IN, h1, true);
    __syncthreads();
    sel_layer(P + SEL_OFF_W2, P + SEL_OFF_B2, h1, SEL_H, h2, true);
    __syncthreads();
    sel_layer(P + SEL_OFF_W3, P + SEL_OFF_B3, h2, SEL_H, h3, true);
    __syncthreads();
    if (threadIdx.x < SEL_OUT) {
        float z = P[SEL_OFF_B4 + threadIdx.x];
        for (int i = 0; i < SEL_H; i++) z += P[SEL_OFF_W4 + threadIdx.x * SEL_H + i] * h3[i];
        plan->pred[threadIdx.x] = z;
    }
}

// ---- inference: one block of SEL_H threads per input row ---------------------
__device__ __forceinline__ void sel_layer(const float *W, const float *b, const float *in, int nin, float *out,
                                          bool relu) {
    const int j = threadIdx.x;
    float z = b[j];
    for (int i = 0; i < nin; i++) z += W[j * nin + i] * in[i];
    out[j] = relu ? fmaxf(z, 0.f) : z;
}

__global__ __launch_bounds__(SEL_H) void k_sel_infer(const float *__restrict__ P, const float *__restrict__ X,
                                                     float *Y) {
    __shared__ float x[SEL_IN], h1[SEL_H], h2[SEL_H], h3[SEL_H];
    const float *xr = X + (uint64_t)blockIdx.x * SEL_IN;
    if (threadIdx.x < SEL_IN)
        x[threadIdx.x] = (xr[threadIdx.x] - P[SEL_OFF_MU + threadIdx.x]) / P[SEL_OFF_SIG + threadIdx.x];
    __syncthreads();
    sel_layer(P + SEL_OFF_W1, P + SEL_OFF_B1, x, SEL_IN, h1, true);
    __syncthreads();
    sel_layer(P + SEL_OFF_W2, P + SEL_OFF_B2, h1, SEL_H, h2, true);
    __syncthreads();
    sel_layer(P + SEL_OFF_W3, P + SEL_OFF_B3, h2, SEL_H, h3, true);
    __syncthreads();
    if (threadIdx.x < SEL_OUT) {
        float z = P[SEL_OFF_B4 + threadIdx.x];
        for (int i = 0; i < SEL_H; i++) z += P[SEL_OFF_W4 + threadIdx.x * SEL_H + i] * h3[i];
        Y[(uint64_t)blockIdx.x * SEL_OUT + threadIdx.x] = z;
    }
}

// ---- training -------------------------------------------------------------
// Loss per row: (out0 - y_log)^2 + lambda * cross-entropy(out1..SEL_CLS, y_class),
// averaged over a minibatch of SEL_B rows (the last one may be short). Rows
// are visited in the given order every epoch. Adam (beta1 0.9, beta2 0.999,
// eps 1e-8). M and V hold the moments (SEL_NW floats each, zeroed by the
// caller). loss[e] receives the mean loss of epoch e.
constexpr int SEL_B = 64;
constexpr int SEL_TT = 256;

__device__ __forceinline__ void sel_adam(float *P, float *M, float *V, int i, float g, float lr, float c1, float c2) {
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

// dst[b][j] = act(sum_i W[j][i] src[b][i] + bias[j]) for nb rows
__device__ __forceinline__ void sel_fwd(const float *W, const float *bias, const float *src, int nin, float *dst,
                                        int nout, int nb, bool relu) {
    for (int e = threadIdx.x; e < nb * nout; e += blockDim.x) {
        const int b = e / nout, j = e % nout;
        float z = bias[j];
        for (int i = 0; i < nin; i++) z += W[j * nin + i] * src[b * nin + i];
        dst[b * nout + j] = relu ? fmaxf(z, 0.f) : z;
    }
}

// dsrc[b][i] = (sum_j W[j][i] dz[b][j]) * (src[b][i] > 0), the gradient at the
// previous layer's pre-activation
__device__ __forceinline__ void sel_bwd_act(const float *W, const float *dz, int nout, const float *src, int nin,
                                            float *dsrc, int nb) {
    for (int e = threadIdx.x; e < nb * nin; e += blockDim.x) {
        const int b = e / nin, i = e % nin;
        float g = 0;
        for (int j = 0; j < nout; j++) g += W[j * nin + i] * dz[b * nout + j];
        dsrc[b * nin + i] = src[b * nin + i] > 0.f ? g : 0.f;
    }
}

// gradient of W (nout x nin) and bias, then the Adam step on them
__device__ __forceinline__ void sel_bwd_w(float *P, float *M, float *V, int offW, int offB, const float *dz,
                                          const float *src, int nin, int nout, int nb, float lr, float c1, float c2) {
    for (int e = threadIdx.x; e < nout * nin + nout; e += blockDim.x) {
        float g = 0;
        if (e < nout * nin) {
            const int j = e / nin, i = e % nin;
            for (int b = 0; b < nb; b++) g += dz[b * nout + j] * src[b * nin + i];
            sel_adam(P, M, V, offW + e, g, lr, c1, c2);
        } else {
            const int j = e - nout * nin;
            for (int b = 0; b < nb; b++) g += dz[b * nout + j];
            sel_adam(P, M, V, offB + j, g, lr, c1, c2);
        }
    }
}

__global__ __launch_bounds__(SEL_TT) void k_sel_train(float *P, float *M, float *V, const float *__restrict__ X,
                                                      const float *__restrict__ Y, uint32_t n, uint32_t epochs,
                                                      float lr, float lambda, float *loss) {
    __shared__ float a0[SEL_B * SEL_IN], a1[SEL_B * SEL_H], a2[SEL_B * SEL_H], a3[SEL_B * SEL_H];
    __shared__ float o[SEL_B * SEL_OUT], d3[SEL_B * SEL_H], d2[SEL_B * SEL_H];
    __shared__ float s_loss;
    uint32_t step = 0;
    for (uint32_t ep = 0; ep < epochs; ep++) {
        if (threadIdx.x == 0) s_loss = 0;
        for (uint32_t r0 = 0; r0 < n; r0 += SEL_B) {
            const int nb = (int)min<uint32_t>(SEL_B, n - r0);
            step++;
            const float c1 = 1.f - powf(0.9f, (float)step), c2 = 1.f - powf(0.999f, (float)step);
            for (int e = threadIdx.x; e < nb * SEL_IN; e += blockDim.x) {
                const int b = e / SEL_IN, i = e % SEL_IN;
                a0[e] = (X[(uint64_t)(r0 + b) * SEL_IN + i] - P[SEL_OFF_MU + i]) / P[SEL_OFF_SIG + i];
            }
            __syncthreads();
            sel_fwd(P + SEL_OFF_W1, P + SEL_OFF_B1, a0, SEL_IN, a1, SEL_H, nb, true);
            __syncthreads();
            sel_fwd(P + SEL_OFF_W2, P + SEL_OFF_B2, a1, SEL_H, a2, SEL_H, nb, true);
            __syncthreads();
            sel_fwd(P + SEL_OFF_W3, P + SEL_OFF_B3, a2, SEL_H, a3, SEL_H, nb, true);
            __syncthreads();
            sel_fwd(P + SEL_OFF_W4, P + SEL_OFF_B4, a3, SEL_H, o, SEL_OUT, nb, false);
            __syncthreads();
            // output gradient (in place over o) and the minibatch loss
            if (threadIdx.x < (uint32_t)nb) {
                const int b = threadIdx.x;
                const float *y = Y + (uint64_t)(r0 + b) * 2;
                float *ob = o + b * SEL_OUT;
                const float diff = ob[0] - y[0];
                const int cls = (int)y[1];
                float mxl = ob[1];
                for (int c = 1; c < SEL_CLS; c++) mxl = fmaxf(mxl, ob[1 + c]);
                float e[SEL_CLS], se = 0.f;
                for (int c = 0; c < SEL_CLS; c++) {
                    e[c] = expf(ob[1 + c] - mxl);
                    se += e[c];
                }
                const float lc = ob[1 + cls] - mxl;
                atomicAdd(&s_loss, diff * diff + lambda * (logf(se) - lc));
                const float inv = 1.f / (float)nb;
                ob[0] = 2.f * diff * inv;
                for (int c = 0; c < SEL_CLS; c++) ob[1 + c] = lambda * (e[c] / se - (cls == c)) * inv;
            }
            __syncthreads();
            // layer 4, then 3, 2, 1: the activation gradient below each layer
            // uses the layer's weights before its update
            sel_bwd_act(P + SEL_OFF_W4, o, SEL_OUT, a3, SEL_H, d3, nb);
            __syncthreads();
            sel_bwd_w(P, M, V, SEL_OFF_W4, SEL_OFF_B4, o, a3, SEL_H, SEL_OUT, nb, lr, c1, c2);
            __syncthreads();
            sel_bwd_act(P + SEL_OFF_W3, d3, SEL_H, a2, SEL_H, d2, nb);
            __syncthreads();
            sel_bwd_w(P, M, V, SEL_OFF_W3, SEL_OFF_B3, d3, a2, SEL_H, SEL_H, nb, lr, c1, c2);
            __syncthreads();
            sel_bwd_act(P + SEL_OFF_W2, d2, SEL_H, a1, SEL_H, d3, nb);  // d3 now holds layer-1 gradients
            __syncthreads();
            sel_bwd_w(P, M, V, SEL_OFF_W2, SEL_OFF_B2, d2, a1, SEL_H, SEL_H, nb, lr, c1, c2);
            __syncthreads();
            sel_bwd_w(P, M, V, SEL_OFF_W1, SEL_OFF_B1, d3, a0, SEL_IN, SEL_H, nb, lr, c1, c2);
            __syncthreads();
        }
        if (threadIdx.x == 0) loss[ep] = s_loss / (float)n;
        __syncthreads();
    }
}

}  // namespace gbagg
