// Small multilayer perceptron for the learned-cutoff form, gfx950: forward, backward and the
// optimizer step, so the network is trained on the GPU by this library's own kernels.
//
// The network is in -> width -> width -> out with ReLU after the two hidden layers.  Weights
// are stored in BF16 (row-major [fan_out][fan_in]); the master copy the optimizer updates is
// FP32.  Every product runs on v_mfma_f32_16x16x32_bf16 with FP32 accumulation: activations
// and gradients live in FP32 and are rounded to BF16 (nearest-even) as they are loaded.
//
// One generic kernel computes every product of the forward and backward passes:
//   C[m][n] = epilogue( sum_k A(m, k) * B(n, k) ),   A(m, k) = A[m * sam + k * sak],
//                                                    B(n, k) = B[n * sbn + k * sbk],
// with the epilogue adding a bias, applying ReLU, or masking by a ReLU output (H > 0).
// One wave computes one 16 x 16 tile of C, stepping K by 32.
#include <string.h>

#include "common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;

enum Epilogue { kEpiNone = 0, kEpiBias = 1, kEpiBiasRelu = 2, kEpiReluMask = 3 };

template <typename T>
__device__ __forceinline__ __bf16 to_bf16(T x);
template <>
__device__ __forceinline__ __bf16 to_bf16<float>(float x) {
    return (__bf16)x;
}
template <>
__device__ __forceinline__ __bf16 to_bf16<__bf16>(__bf16 x) {
    return x;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(kWave) void mlp_gemm_kernel(int M, int N, int K, const TA* __restrict__ A,
                                                         int64_t sam, int64_t sak, const TB* __restrict__ B,
                                                         int64_t sbn, int64_t sbk, float* __restrict__ C, int64_t ldc,
                                                         int epi, const float* __restrict__ bias,
                                                         const float* __restrict__ mask, int64_t ldmask) {
    const int lane = threadIdx.x;
    const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
    const int am = m0 + lane % 16, bn = n0 + lane % 16;
    const int kq = 8 * (lane / 16);
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = k0 + kq + i;
            a[i] = (am < M && k < K) ? to_bf16<TA>(A[am * sam + k * sak]) : (__bf16)0.f;
            b[i] = (bn < N && k < K) ? to_bf16<TB>(B[bn * sbn + k * sbk]) : (__bf16)0.f;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    // output register r of lane l holds C[m0 + 4 (l / 16) + r][n0 + l % 16]
    const int cn = n0 + lane % 16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int cm = m0 + 4 * (lane / 16) + r;
        if (cm >= M || cn >= N) continue;
        float v = acc[r];
        if (epi == kEpiBias || epi == kEpiBiasRelu) v += bias[cn];
        if (epi == kEpiBiasRelu) v = v > 0.f ? v : 0.f;
        if (epi == kEpiReluMask) v = mask[cm * ldmask + cn] > 0.f ? v : 0.f;
        C[cm * ldc + cn] = v;
    }
}

template <typename TA, typename TB>
int gemm(bsmps_ctx* ctx, int M, int N, int K, const TA* A, int64_t sam, int64_t sak, const TB* B, int64_t sbn,
         int64_t sbk, float* C, int64_t ldc, int epi, const float* bias, const float* mask, int64_t ldmask) {
    if (M <= 0 || N <= 0) return BSMPS_OK;
    dim3 grid((N + 15) / 16, (M + 15) / 16);
    if (grid.y > 65535u) {
        bsmps_set_error("mlp: %d rows exceed one launch", M);
        return BSMPS_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB>), grid, dim3(kWave), 0, ctx->stream, M, N, K, A, sam, sak, B, sbn, sbk,
                       C, ldc, epi, bias, mask, ldmask);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

// activations (FP32) times weights (BF16), and activations times activations
int gemm_fb(bsmps_ctx* ctx, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const __bf16* B,
            int64_t sbn, int64_t sbk, float* C, int64_t ldc, int epi, const float* bias, const float* mask,
            int64_t ldmask) {
    return gemm<float, __bf16>(ctx, M, N, K, A, sam, sak, B, sbn, sbk, C, ldc, epi, bias, mask, ldmask);
}
int gemm_ff(bsmps_ctx* ctx, int M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
            int64_t sbn, int64_t sbk, float* C, int64_t ldc, int epi, const float* bias, const float* mask,
            int64_t ldmask) {
    return gemm<float, float>(ctx, M, N, K, A, sam, sak, B, sbn, sbk, C, ldc, epi, bias, mask, ldmask);
}

// db[c] = sum_r G[r][c], rows added in order
__global__ void colsum_kernel(int rows, int cols, const float* __restrict__ G, float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += G[(int64_t)r * cols + c];
    out[c] = s;
}

int colsum(bsmps_ctx* ctx, int rows, int cols, const float* G, float* out) {
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 63) / 64), dim3(64), 0, ctx->stream, rows, cols, G, out);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

__global__ void f32_to_bf16_kernel(int64_t count, const float* __restrict__ src, __bf16* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (__bf16)src[i];
}

// Adam (Kingma & Ba 2015) with bias correction, one thread per parameter.
__global__ void adam_kernel(int64_t count, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float beta1, float beta2, float eps, float c1, float c2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = beta1 * m[i] + (1.f - beta1) * gi;
        const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Mean squared error over n rows x d outputs: dY = 2 (Y - T) / (n d), err = (Y - T)^2.
__global__ void mse_grad_kernel(int64_t count, float scale, const float* __restrict__ Y, const float* __restrict__ T,
                                float* __restrict__ dY, float* __restrict__ err) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const float e = Y[i] - T[i];
        dY[i] = scale * e;
        if (err) err[i] = e * e;
    }
}

unsigned blocks_for(int64_t count) {
    const int64_t b = (count + 255) / 256;
    return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

bool bad_dims(int n, int d_in, int width, int d_out) {
    return n < 0 || d_in <= 0 || width <= 0 || d_out <= 0 || d_in > 4096 || width > 4096 || d_out > 4096;
}

}  // namespace

extern "C" int bsmps_mlp_forward(bsmps_ctx* ctx, int n, int d_in, int width, int d_out, const float* X,
                                 const void* W1, const float* b1, const void* W2, const float* b2, const void* W3,
                                 const float* b3, float* H1, float* H2, float* Y) {
    if (!ctx || bad_dims(n, d_in, width, d_out)) {
        bsmps_set_error("bsmps_mlp_forward: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    if (!X || !W1 || !b1 || !W2 || !b2 || !W3 || !b3 || !H1 || !H2 || !Y) {
        bsmps_set_error("bsmps_mlp_forward: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const __bf16* w1 = static_cast<const __bf16*>(W1);
    const __bf16* w2 = static_cast<const __bf16*>(W2);
    const __bf16* w3 = static_cast<const __bf16*>(W3);
    // H1 = relu(X W1^T + b1); H2 = relu(H1 W2^T + b2); Y = H2 W3^T + b3
    BSMPS_TRY(gemm_fb(ctx, n, width, d_in, X, d_in, 1, w1, d_in, 1, H1, width, kEpiBiasRelu, b1, nullptr, 0));
    BSMPS_TRY(gemm_fb(ctx, n, width, width, H1, width, 1, w2, width, 1, H2, width, kEpiBiasRelu, b2, nullptr,
                                  0));
    BSMPS_TRY(gemm_fb(ctx, n, d_out, width, H2, width, 1, w3, width, 1, Y, d_out, kEpiBias, b3, nullptr, 0));
    return BSMPS_OK;
}

extern "C" int bsmps_mlp_backward(bsmps_ctx* ctx, int n, int d_in, int width, int d_out, const float* X,
                                  const void* W2, const void* W3, const float* H1, const float* H2, const float* dY,
                                  float* scratch, float* dW1, float* db1, float* dW2, float* db2, float* dW3,
                                  float* db3) {
    if (!ctx || bad_dims(n, d_in, width, d_out)) {
        bsmps_set_error("bsmps_mlp_backward: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (!X || !W2 || !W3 || !H1 || !H2 || !dY || !scratch || !dW1 || !db1 || !dW2 || !db2 || !dW3 || !db3) {
        bsmps_set_error("bsmps_mlp_backward: null pointer");
        return BSMPS_ERR_INVALID;
    }
    const __bf16* w2 = static_cast<const __bf16*>(W2);
    const __bf16* w3 = static_cast<const __bf16*>(W3);
    float* dZ2 = scratch;                      // [n][width]
    float* dZ1 = scratch + (int64_t)n * width;  // [n][width]
    // output layer: dW3[o][w] = sum_r dY[r][o] H2[r][w]; db3 = colsum(dY)
    BSMPS_TRY(gemm_ff(ctx, d_out, width, n, dY, 1, d_out, H2, 1, width, dW3, width, kEpiNone, nullptr, nullptr,
                                 0));
    BSMPS_TRY(colsum(ctx, n, d_out, dY, db3));
    // dZ2 = (dY W3) masked by H2 > 0
    BSMPS_TRY(gemm_fb(ctx, n, width, d_out, dY, d_out, 1, w3, 1, width, dZ2, width, kEpiReluMask, nullptr,
                                  H2, width));
    BSMPS_TRY(gemm_ff(ctx, width, width, n, dZ2, 1, width, H1, 1, width, dW2, width, kEpiNone, nullptr,
                                 nullptr, 0));
    BSMPS_TRY(colsum(ctx, n, width, dZ2, db2));
    // dZ1 = (dZ2 W2) masked by H1 > 0
    BSMPS_TRY(gemm_fb(ctx, n, width, width, dZ2, width, 1, w2, 1, width, dZ1, width, kEpiReluMask, nullptr,
                                  H1, width));
    BSMPS_TRY(gemm_ff(ctx, width, d_in, n, dZ1, 1, width, X, 1, d_in, dW1, d_in, kEpiNone, nullptr, nullptr,
                                 0));
    BSMPS_TRY(colsum(ctx, n, width, dZ1, db1));
    return BSMPS_OK;
}

extern "C" int bsmps_f32_to_bf16(bsmps_ctx* ctx, int64_t count, const float* src, void* dst) {
    if (!ctx || count < 0 || (count > 0 && (!src || !dst))) {
        bsmps_set_error("bsmps_f32_to_bf16: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (count == 0) return BSMPS_OK;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(blocks_for(count)), dim3(256), 0, ctx->stream, count, src,
                       static_cast<__bf16*>(dst));
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_adam_step(bsmps_ctx* ctx, int64_t count, float* param, const float* grad, float* m, float* v,
                               float lr, float beta1, float beta2, float eps, int step) {
    if (!ctx || count < 0 || step < 1 || (count > 0 && (!param || !grad || !m || !v))) {
        bsmps_set_error("bsmps_adam_step: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (count == 0) return BSMPS_OK;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count)), dim3(256), 0, ctx->stream, count, param, grad, m, v, lr,
                       beta1, beta2, eps, c1, c2);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}

extern "C" int bsmps_mse_grad(bsmps_ctx* ctx, int n, int d_out, const float* Y, const float* T, float* dY, float* err) {
    if (!ctx || n < 0 || d_out <= 0 || (n > 0 && (!Y || !T || !dY))) {
        bsmps_set_error("bsmps_mse_grad: invalid arguments");
        return BSMPS_ERR_INVALID;
    }
    if (n == 0) return BSMPS_OK;
    const int64_t count = (int64_t)n * d_out;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(count)), dim3(256), 0, ctx->stream, count,
                       2.f / (float)count, Y, T, dY, err);
    BSMPS_HIP(hipGetLastError());
    return BSMPS_OK;
}
