// Match-scoring MLP for guided RANSAC sampling: 8 -> 64 (ReLU) -> 64 (ReLU) -> 1 logit.
//
// Parameters (4801 f32 master copies): W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3.
// The network computes with the BF16 (round-to-nearest-even) values of the weights and FP32
// accumulation; training updates the FP32 masters with Adam from the gradient taken at the
// BF16 weights. Kernels:
//   * features: the 8 inputs of every putative match of a batch of image pairs;
//   * forward: one thread per match, the hidden layer-2 weights broadcast from LDS;
//   * grad: forward + backward over chunks of 64 matches per workgroup (one thread per
//     match for the per-match passes, then one thread per hidden unit for the weight
//     gradient sums over the chunk), per-workgroup partial gradients;
//   * reduce: partial gradients summed over workgroups in a fixed order (deterministic);
//   * adam: the optimizer step on the masters.
#include <algorithm>

#include "ps_internal.h"

namespace ps {

namespace {

constexpr int kP = 4801;
constexpr int oW1 = 0, oB1 = 512, oW2 = 576, oB2 = 4672, oW3 = 4736, oB3 = 4800;
constexpr int kChunk = 64;   // matches per workgroup pass of the gradient kernel

__device__ __forceinline__ float bf16_value(float x) {
    uint32_t b = __float_as_uint(x);
    b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    return __uint_as_float(b);
}

struct MlpLds {
    float W2[64][64];
    float W1[64][8];
    float b1[64], b2[64], w3[64];
    float b3;
};

__device__ void load_weights(const float *params, MlpLds &L) {
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) (&L.W2[0][0])[i] = bf16_value(params[oW2 + i]);
    for (int i = threadIdx.x; i < 512; i += blockDim.x) (&L.W1[0][0])[i] = bf16_value(params[oW1 + i]);
    for (int i = threadIdx.x; i < 64; i += blockDim.x) {
        L.b1[i] = bf16_value(params[oB1 + i]);
        L.b2[i] = bf16_value(params[oB2 + i]);
        L.w3[i] = bf16_value(params[oW3 + i]);
    }
    if (threadIdx.x == 0) L.b3 = bf16_value(params[oB3]);
}

__device__ __forceinline__ void layer1(const MlpLds &L, const float (&x)[8], float (&h1)[64]) {
#pragma unroll
    for (int i = 0; i < 64; i++) {
        float a = L.b1[i];
#pragma unroll
        for (int k = 0; k < 8; k++) a = __builtin_fmaf(L.W1[i][k], x[k], a);
        h1[i] = fmaxf(a, 0.0f);
    }
}

__device__ __forceinline__ float layer2_unit(const MlpLds &L, int i, const float (&h1)[64]) {
    float a = L.b2[i];
#pragma unroll
    for (int j = 0; j < 64; j += 4) {
        const float4 w = *reinterpret_cast<const float4 *>(&L.W2[i][j]);
        a = __builtin_fmaf(w.x, h1[j], a);
        a = __builtin_fmaf(w.y, h1[j + 1], a);
        a = __builtin_fmaf(w.z, h1[j + 2], a);
        a = __builtin_fmaf(w.w, h1[j + 3], a);
    }
    return a;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *params, const float *x, uint32_t n,
                                                          float *logit, float *prob) {
    __shared__ __attribute__((aligned(16))) MlpLds L;
    load_weights(params, L);
    __syncthreads();
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        float xv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) xv[k] = x[(size_t)s * 8 + k];
        float h1[64];
        layer1(L, xv, h1);
        float z = L.b3;
#pragma unroll 2
        for (int i = 0; i < 64; i++) z = __builtin_fmaf(L.w3[i], fmaxf(layer2_unit(L, i, h1), 0.0f), z);
        logit[s] = z;
        if (prob) prob[s] = 1.0f / (1.0f + expf(-z));
    }
}

// One wave per workgroup; persistent per-lane gradient accumulators: lane t owns W2 row t,
// W1 row t, b1[t], b2[t], w3[t] (lane 0 also b3).
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float *params, const float *x, const float *y,
                                                      uint32_t n, float inv_n, float *partial, float *loss_partial) {
    __shared__ __attribute__((aligned(16))) MlpLds L;
    __shared__ float A[kChunk][65];     // z2, then dz1
    __shared__ __attribute__((aligned(16))) float Bh[kChunk][64];   // h1
    __shared__ float X[kChunk][8];
    __shared__ float D[kChunk];         // dL/dz (already scaled by 1/n); 0 for padding
    load_weights(params, L);
    __syncthreads();
    const int t = threadIdx.x;
    float gW2[64], gW1[8];
#pragma unroll
    for (int j = 0; j < 64; j++) gW2[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) gW1[k] = 0.0f;
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
    for (uint32_t c0 = blockIdx.x * kChunk; c0 < n; c0 += gridDim.x * kChunk) {
        // per match (lane = match): forward, dL/dz, dh1 -> dz1 (kept in registers)
        const uint32_t s = c0 + t;
        const bool valid = s < n;
        float xv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            xv[k] = valid ? x[(size_t)s * 8 + k] : 0.0f;
            X[t][k] = xv[k];
        }
        float h1[64];
        layer1(L, xv, h1);
#pragma unroll
        for (int j = 0; j < 64; j += 4)
            *reinterpret_cast<float4 *>(&Bh[t][j]) = make_float4(h1[j], h1[j + 1], h1[j + 2], h1[j + 3]);
        float z = L.b3;
#pragma unroll 2
        for (int i = 0; i < 64; i++) {
            float a = layer2_unit(L, i, h1);
            A[t][i] = a;
            z = __builtin_fmaf(L.w3[i], fmaxf(a, 0.0f), z);
        }
        const float yt = valid ? y[s] : 0.0f;
        const float dz = valid ? (1.0f / (1.0f + expf(-z)) - yt) * inv_n : 0.0f;
        if (valid) loss += fmaxf(z, 0.0f) - z * yt + log1pf(expf(-fabsf(z)));
        D[t] = dz;
        // dh1[j] = sum_i W2[i][j] dz2[i], dz2[i] = dz * w3[i] * [z2_i > 0]
        float dh1[64];
#pragma unroll
        for (int j = 0; j < 64; j++) dh1[j] = 0.0f;
#pragma unroll 2
        for (int i = 0; i < 64; i++) {
            const float d2 = A[t][i] > 0.0f ? dz * L.w3[i] : 0.0f;
#pragma unroll
            for (int j = 0; j < 64; j += 4) {
                const float4 w = *reinterpret_cast<const float4 *>(&L.W2[i][j]);
                dh1[j] = __builtin_fmaf(w.x, d2, dh1[j]);
                dh1[j + 1] = __builtin_fmaf(w.y, d2, dh1[j + 1]);
                dh1[j + 2] = __builtin_fmaf(w.z, d2, dh1[j + 2]);
                dh1[j + 3] = __builtin_fmaf(w.w, d2, dh1[j + 3]);
            }
        }
#pragma unroll
        for (int j = 0; j < 64; j++) dh1[j] = h1[j] > 0.0f ? dh1[j] : 0.0f;   // dz1
        __syncthreads();
        // per hidden unit (lane = unit t): sums over the chunk's matches
#pragma unroll 2
        for (int m = 0; m < kChunk; m++) {
            const float dzm = D[m];
            const float z2 = A[m][t];
            const float d2 = z2 > 0.0f ? dzm * L.w3[t] : 0.0f;
            gb2 += d2;
            gw3 = __builtin_fmaf(dzm, fmaxf(z2, 0.0f), gw3);
#pragma unroll
            for (int j = 0; j < 64; j += 4) {
                const float4 h = *reinterpret_cast<const float4 *>(&Bh[m][j]);
                gW2[j] = __builtin_fmaf(d2, h.x, gW2[j]);
                gW2[j + 1] = __builtin_fmaf(d2, h.y, gW2[j + 1]);
                gW2[j + 2] = __builtin_fmaf(d2, h.z, gW2[j + 2]);
                gW2[j + 3] = __builtin_fmaf(d2, h.w, gW2[j + 3]);
            }
            if (t == 0) gb3 += dzm;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 64; j++) A[t][j] = dh1[j];
        __syncthreads();
        for (int m = 0; m < kChunk; m++) {
            const float d1 = A[m][t];
            gb1 += d1;
#pragma unroll
            for (int k = 0; k < 8; k++) gW1[k] = __builtin_fmaf(d1, X[m][k], gW1[k]);
        }
        __syncthreads();
    }
    float *out = partial + (size_t)blockIdx.x * kP;
#pragma unroll
    for (int j = 0; j < 64; j++) out[oW2 + t * 64 + j] = gW2[j];
#pragma unroll
    for (int k = 0; k < 8; k++) out[oW1 + t * 8 + k] = gW1[k];
    out[oB1 + t] = gb1;
    out[oB2 + t] = gb2;
    out[oW3 + t] = gw3;
    if (t == 0) out[oB3] = gb3;
    // loss: wave sum
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (t == 0) loss_partial[blockIdx.x] = loss;
}

__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float *partial, const float *loss_partial,
                                                         uint32_t nblocks, float inv_n, float *grad, float *loss) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)kP) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nblocks; b++) s += partial[(size_t)b * kP + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.0f;
        for (uint32_t b = 0; b < nblocks; b++) s += loss_partial[b];
        *loss = s * inv_n;
    }
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float *params, float *m1, float *m2, const float *grad,
                                                       float lr, float beta1, float beta2, float eps, float c1,
                                                       float c2) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kP) return;
    const float g = grad[p];
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    params[p] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// Inputs of every putative match: ratio, log(scale_t / scale_q), sin and cos of the
// orientation change, descriptor distance, query x / width, query y / height, and the
// match's rank by (ratio, index) within its pair divided by the pair's match count.
__global__ __launch_bounds__(256) void mlp_features_kernel(const ps_feature_match *m, const ps_feature *qf,
                                                           const ps_feature *tf, const ps_mlp_pair *pairs,
                                                           float *x) {
    const ps_mlp_pair P = pairs[blockIdx.x];
    const ps_feature_match *pm = m + P.match_offset;
    for (uint32_t i = threadIdx.x; i < P.n; i += blockDim.x) {
        const ps_feature_match &a = pm[i];
        float ri = a.ratio;
        if (!(ri == ri)) ri = 1.0f;
        uint32_t rank = 0;
        for (uint32_t j = 0; j < P.n; j++) {
            float rj = pm[j].ratio;
            if (!(rj == rj)) rj = 1.0f;
            rank += (rj < ri) || (rj == ri && j < i);
        }
        const ps_feature &q = qf[P.qfeat_offset + (uint32_t)a.query_index];
        const ps_feature &t = tf[P.tfeat_offset + (uint32_t)a.target_index];
        const float dth = t.orientation - q.orientation;
        float *o = x + ((size_t)P.match_offset + i) * 8;
        o[0] = fminf(ri, 2.0f);
        o[1] = logf(fmaxf(t.scale, 1e-6f) / fmaxf(q.scale, 1e-6f));
        o[2] = sinf(dth);
        o[3] = cosf(dth);
        o[4] = a.distance;
        o[5] = a.qx / P.width;
        o[6] = a.qy / P.height;
        o[7] = (float)rank / (float)P.n;
    }
}

}  // namespace

constexpr uint32_t kGradBlocks = 512;

size_t mlp_grad_scratch_bytes() { return ((size_t)kGradBlocks * kP + kGradBlocks + 64) * sizeof(float); }

hipError_t launch_mlp_forward(const float *params, const float *x, uint32_t n, float *logit, float *prob,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = std::min<uint32_t>((n + 255) / 256, 2048);
    mlp_forward_kernel<<<blocks, 256, 0, st>>>(params, x, n, logit, prob);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *params, const float *x, const float *y, uint32_t n, void *scratch,
                           float *grad, float *loss, hipStream_t st) {
    float *partial = static_cast<float *>(scratch);
    float *lpart = partial + (size_t)kGradBlocks * kP;
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(kGradBlocks, (n + kChunk - 1) / kChunk));
    const float inv_n = n ? 1.0f / (float)n : 0.0f;
    mlp_grad_kernel<<<blocks, 64, 0, st>>>(params, x, y, n, inv_n, partial, lpart);
    mlp_reduce_kernel<<<(kP + 255) / 256, 256, 0, st>>>(partial, lpart, blocks, inv_n, grad, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *params, float *m1, float *m2, const float *grad, uint32_t step, float lr,
                           hipStream_t st) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)step)));
    mlp_adam_kernel<<<(kP + 255) / 256, 256, 0, st>>>(params, m1, m2, grad, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_features(const ps_feature_match *m, const ps_feature *qf, const ps_feature *tf,
                               const ps_mlp_pair *pairs, uint32_t n_pairs, float *x, hipStream_t st) {
    if (n_pairs == 0) return hipSuccess;
    mlp_features_kernel<<<n_pairs, 256, 0, st>>>(m, qf, tf, pairs, x);
    return hipGetLastError();
}

}  // namespace ps
