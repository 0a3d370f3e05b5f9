// vx_mlp.hip -- the node-growth predictor: an FP32 MLP 16 -> 64 -> 64 -> 1 (ReLU),
// forward, backward and Adam kernels.
//
// Parameters, packed float32 (kMlpParams):
//   W1 [64][16] | b1 [64] | W2 [64][64] | b2 [64] | W3 [64] | b3
// Loss: mean squared error over the batch, L = (1/n) sum_s (y_s - t_s)^2.
//
// One wavefront per sample: lane j owns hidden unit j of both layers.  Weights are
// staged in LDS once per block, W1 and W2 transposed for the forward products (lane
// j reads column j: consecutive lanes, consecutive banks) and W2 as stored for the
// backward product dh1 = W2^T dh2 (lane k reads column k of W2).  Per-parameter
// gradients are reduced over the batch by one lane per parameter in sample order, so
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace vx {

constexpr int kIn = 16, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kHid * kIn, kOffW2 = kOffB1 + kHid,
              kOffB2 = kOffW2 + kHid * kHid, kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid;
constexpr int kMlpParams = kOffB3 + 1;  // 5313
constexpr int kMlpWaves = 4;            // samples in flight per block

__device__ inline float wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct MlpLds {
    float w1t[kIn][kHid];   // w1t[i][j] = W1[j][i]
    float w2t[kHid][kHid];  // w2t[k][j] = W2[j][k]
    float w2[kHid][kHid];   // as stored (backward)
    float h1[kMlpWaves][kHid];
    float dh2[kMlpWaves][kHid];
};

__device__ inline void mlp_stage(const float* P, MlpLds& S, bool backward) {
    for (int e = threadIdx.x; e < kHid * kIn; e += blockDim.x) {
        int j = e / kIn, i = e % kIn;
        S.w1t[i][j] = P[kOffW1 + e];
    }
    for (int e = threadIdx.x; e < kHid * kHid; e += blockDim.x) {
        int j = e / kHid, k = e % kHid;
        float w = P[kOffW2 + e];
        S.w2t[k][j] = w;
        if (backward) S.w2[j][k] = w;
    }
    __syncthreads();
}

// Forward (and, with target, the per-sample backward terms) of one sample per wave.
// ws (backward only): per sample h1[64], h2[64], dh1[64], dh2[64], then dy[n].
__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_fwd(const float* P, const float* x, uint32_t n,
                                                            float* y, const float* target, float* ws) {
    __shared__ MlpLds S;
    const bool backward = target != nullptr;
    mlp_stage(P, S, backward);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float b1 = P[kOffB1 + lane], b2 = P[kOffB2 + lane], w3 = P[kOffW3 + lane], b3 = P[kOffB3];
    const uint32_t stride = gridDim.x * kMlpWaves;
    for (uint32_t s0 = blockIdx.x * kMlpWaves; s0 < n; s0 += stride) {
        // the block's waves take samples s0 .. s0+3; a wave past n idles (block-uniform loop)
        uint32_t s = s0 + w;
        bool ok = s < n;
        float a1 = b1;
        if (ok) {
#pragma unroll
            for (int i = 0; i < kIn; i++) a1 = fmaf(S.w1t[i][lane], x[(size_t)s * kIn + i], a1);
        }
        float h1 = a1 > 0.f ? a1 : 0.f;
        S.h1[w][lane] = h1;
        __syncthreads();
        float a2 = b2;
#pragma unroll 8
        for (int k = 0; k < kHid; k++) a2 = fmaf(S.w2t[k][lane], S.h1[w][k], a2);
        float h2 = a2 > 0.f ? a2 : 0.f;
        float out = wave_sum_f(w3 * h2) + b3;
        if (ok && lane == 0 && y) y[s] = out;
        if (backward) {
            float dy = ok ? 2.f * (out - target[s]) / (float)n : 0.f;
            float dh2 = a2 > 0.f ? dy * w3 : 0.f;
            S.dh2[w][lane] = dh2;
            __syncthreads();
            float g = 0.f;
#pragma unroll 8
            for (int j = 0; j < kHid; j++) g = fmaf(S.w2[j][lane], S.dh2[w][j], g);
            float dh1 = a1 > 0.f ? g : 0.f;
            if (ok) {
                float* r = ws + (size_t)s * 4 * kHid;
                r[lane] = h1;
                r[kHid + lane] = h2;
                r[2 * kHid + lane] = dh1;
                r[3 * kHid + lane] = dh2;
                if (lane == 0) ws[(size_t)n * 4 * kHid + s] = dy;
            }
        }
        __syncthreads();  // S.h1 / S.dh2 are rewritten by the next sample
    }
}

// dL/dparam for every parameter, reduced over the batch in sample order; the lane
// of b3 also writes the loss.
__global__ void k_mlp_grad(const float* x, const float* target, const float* y, uint32_t n, const float* ws,
                           float* grad, float* loss) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kMlpParams) return;
    const float* dy = ws + (size_t)n * 4 * kHid;
    float g = 0.f;
    if (p < (uint32_t)kOffB1) {
        int j = p / kIn, i = p % kIn;
        for (uint32_t s = 0; s < n; s++) g = fmaf(ws[(size_t)s * 4 * kHid + 2 * kHid + j], x[(size_t)s * kIn + i], g);
    } else if (p < (uint32_t)kOffW2) {
        int j = p - kOffB1;
        for (uint32_t s = 0; s < n; s++) g += ws[(size_t)s * 4 * kHid + 2 * kHid + j];
    } else if (p < (uint32_t)kOffB2) {
        int e = p - kOffW2, j = e / kHid, k = e % kHid;
        for (uint32_t s = 0; s < n; s++) {
            const float* r = ws + (size_t)s * 4 * kHid;
            g = fmaf(r[3 * kHid + j], r[k], g);
        }
    } else if (p < (uint32_t)kOffW3) {
        int j = p - kOffB2;
        for (uint32_t s = 0; s < n; s++) g += ws[(size_t)s * 4 * kHid + 3 * kHid + j];
    } else if (p < (uint32_t)kOffB3) {
        int j = p - kOffW3;
        for (uint32_t s = 0; s < n; s++) g = fmaf(dy[s], ws[(size_t)s * 4 * kHid + kHid + j], g);
    } else {
        float l = 0.f;
        for (uint32_t s = 0; s < n; s++) {
            g += dy[s];
            float e = y[s] - target[s];
            l = fmaf(e, e, l);
        }
        if (loss) *loss = l / (float)n;
    }
    grad[p] = g;
}

// Adam (Kingma & Ba, 2015) with bias correction; step counts from 1.
__global__ void k_mlp_adam(float* P, const float* grad, float* m, float* v, uint32_t step, float lr, float b1,
                           float b2, float eps) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kMlpParams) return;
    float g = grad[p];
    float mm = b1 * m[p] + (1.f - b1) * g;
    float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace vx
