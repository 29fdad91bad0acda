// mlp.hpp -- kernels of the learned Nu predictor: a ReLU multilayer perceptron
// whose products run on the bf16 matrix cores (v_mfma_f32_16x16x16_bf16) with
// FP32 accumulation.  Weights are kept in FP32 for the optimiser and rounded to
// bf16 (round to nearest even) at every use, activations likewise, so the
// arithmetic is that of bf16 weights and activations with FP32 sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axg {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[M][N] = A[M][K] B[K][N] on strided FP32 operands (any transpose is a
// choice of strides), then the epilogue: + bias[n], ReLU, or times the mask
// (mask[m][n] > 0, the derivative of a ReLU whose output is the mask).
struct GemmArgs {
    int32_t M, N, K;
    const float *A;
    int64_t a_rs, a_cs;
    const float *B;
    int64_t b_rs, b_cs;
    float *C;
    int64_t c_rs;
    const float *bias;   // [N] or null
    const float *mask;   // [M][mask_rs] or null
    int64_t mask_rs;
    int32_t relu;
};

// One wave per 16 x 16 tile of C, four waves per block (32 x 32).  The MFMA
// operand layout (16x16x16 bf16): lane l holds A[l % 16][4 (l / 16) + r] and
// B[4 (l / 16) + r][l % 16] for r < 4, and receives D[4 (l / 16) + r][l % 16].
__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = (blockIdx.y * 2 + (wave >> 1)) * 16, n0 = (blockIdx.x * 2 + (wave & 1)) * 16;
    if (m0 >= g.M || n0 >= g.N) return;  // whole wave leaves together
    const int li = lane & 15, lk = (lane >> 4) * 4;
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const int am = m0 + li, bn = n0 + li;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4_t a, b;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int k = k0 + lk + r;
            const float av = (am < g.M && k < g.K) ? g.A[am * g.a_rs + k * g.a_cs] : 0.f;
            const float bv = (bn < g.N && k < g.K) ? g.B[k * g.b_rs + bn * g.b_cs] : 0.f;
            a[r] = f2bf(av);
            b[r] = f2bf(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = n0 + li;
    if (n >= g.N) return;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = m0 + lk + r;
        if (m >= g.M) continue;
        float v = acc[r];
        if (g.bias) v += g.bias[n];
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.mask && !(g.mask[m * g.mask_rs + n] > 0.f)) v = 0.f;
        g.C[m * g.c_rs + n] = v;
    }
}

// rows idx[r] of x [n][nin] and y [n] into a contiguous batch
__global__ void mlp_gather_kernel(int32_t rows, int32_t nin, const int32_t *idx, const float *x, const float *y,
                                  float *xb, float *yb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)rows * nin) return;
    const int r = (int)(t / nin), c = (int)(t - (int64_t)r * nin);
    const int64_t src = idx[r];
    xb[t] = x[src * nin + c];
    if (c == 0) yb[r] = y[src];
}

// gradient of L = (1/B) sum_i w_i (z_i - y_i)^2 with respect to z, where
// w_i = under_weight when the prediction is below the target, else 1
__global__ void mlp_loss_grad_kernel(int32_t B, const float *z, const float *y, float under_weight, float *dz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const float d = z[i] - y[i];
    const float w = d < 0.f ? under_weight : 1.f;
    dz[i] = 2.f * w * d / (float)B;
}

// db[n] = sum over the batch of dz[i][n]
__global__ void mlp_bias_grad_kernel(int32_t B, int32_t N, const float *dz, float *db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int i = 0; i < B; i++) s += dz[(int64_t)i * N + n];
    db[n] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 parameters; c1 = 1 - beta1^t, c2 = 1 - beta2^t
__global__ void mlp_adam_kernel(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace axg
