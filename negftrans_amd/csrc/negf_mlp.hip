// negf_mlp.hip -- the kernels of the learned adaptive energy grid (extension; the
// reference has no counterpart): a multilayer perceptron that maps low-order Fourier
// samples of a device potential to T(E) on a fixed grid plus a per-bin log-variance.
//
// Arithmetic: weights and activations are kept in FP32 in HBM; every product runs on the
// matrix cores with both operands rounded to bfloat16 (round to nearest even) as they
// are staged in LDS, and FP32 accumulation (v_mfma_f32_16x16x16_bf16).  Element-wise
// work (bias, GELU, its derivative, the Gaussian negative log-likelihood, Adam) is FP32.
//
// All entry points take device pointers and a stream (hipStream_t passed as void *);
// matrices are row-major with explicit leading dimensions.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/negf_gpu.h"

int negf_fail_internal(int code, const std::string &msg);

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even; NaN stays NaN.
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

#define MLP_BM 64
#define MLP_BN 64
#define MLP_BK 16

// C (M x N) = op(A) (M x K) * op(B) (K x N) with bf16 operands and FP32 accumulation.
// op(A)(i, k) = transA ? A[k lda + i] : A[i lda + k]; op(B)(k, j) = transB ? B[j ldb + k]
// : B[k ldb + j].  mode 0: C = acc (+ C when accumulate); mode 1: Z = acc + bias,
// C = gelu(Z); mode 2: C = acc + bias.
// Workgroup: 256 threads = 4 waves, one 64 x 64 tile of C; wave w owns rows
// 32 (w / 2) .. +32 and columns 32 (w % 2) .. +32 as 2 x 2 MFMA 16 x 16 tiles.
// MFMA 16x16x16 bf16 lane layout: A lane l holds A[l % 16][4 (l / 16) + 0..3], B lane l
// holds B[4 (l / 16) + 0..3][l % 16], D lane l holds D[4 (l / 16) + r][l % 16], r = 0..3.
__global__ __launch_bounds__(256) void k_mlp_gemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                                  int transA, const float *__restrict__ B, int ldb, int transB,
                                                  float *C, int ldc, const float *__restrict__ bias, float *Z,
                                                  int mode, int accumulate) {
    __shared__ unsigned short As[MLP_BM][MLP_BK + 2];
    __shared__ unsigned short Bs[MLP_BK][MLP_BN + 2];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * MLP_BM, n0 = blockIdx.x * MLP_BN;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int ai = t >> 2, ak = (t & 3) * 4;        // A stage: row ai, columns ak .. ak+3
    const int bk = t >> 4, bn = (t & 15) * 4;       // B stage: row bk, columns bn .. bn+3
    for (int k0 = 0; k0 < K; k0 += MLP_BK) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = m0 + ai, k = k0 + ak + q;
            float v = 0.f;
            if (i < M && k < K) v = transA ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k];
            As[ai][ak + q] = f2bf(v);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = k0 + bk, j = n0 + bn + q;
            float v = 0.f;
            if (k < K && j < N) v = transB ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j];
            Bs[bk][bn + q] = f2bf(v);
        }
        __syncthreads();
        const int fi = lane & 15, fk = (lane >> 4) * 4;
        bf16x4 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) af[a][q] = (short)As[wr + 16 * a + fi][fk + q];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) bfr[b][q] = (short)Bs[fk + q][wc + 16 * b + fi];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    const int cj = lane & 15, ci = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + wr + 16 * a + ci + r, j = n0 + wc + 16 * b + cj;
                if (i >= M || j >= N) continue;
                float v = acc[a][b][r];
                const size_t o = (size_t)i * ldc + j;
                if (mode == 0) {
                    C[o] = accumulate ? C[o] + v : v;
                } else {
                    v += bias[j];
                    if (mode == 1) {
                        Z[o] = v;
                        C[o] = gelu(v);
                    } else {
                        C[o] = v;
                    }
                }
            }
}

// dZ = dY * gelu'(Z), element-wise over n entries.
__global__ __launch_bounds__(256) void k_gelu_backward(int n, const float *Z, const float *dY, float *dZ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dY[i] * gelu_grad(Z[i]);
}

// out[j] = sum_i X[i ldx + j] (bias gradients), one thread per column.
__global__ __launch_bounds__(256) void k_colsum(int M, int N, const float *X, int ldx, float *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int i = 0; i < M; ++i) s += X[(size_t)i * ldx + j];
    out[j] = s;
}

// Gaussian negative log-likelihood of targets T (B x P) under the outputs O (B x 2P):
// mean mu = O[:, :P], log-variance s = clamp(O[:, P:], -SMAX, SMAX):
//   loss = 1/(B P) sum 0.5 ((T - mu)^2 e^{-s} + s)
// grad (B x 2P) = d loss / d O (0 for s where the clamp is active).  loss[0] is added to
// atomically (the caller zeroes it).
#define MLP_SMAX 20.0f
__global__ __launch_bounds__(256) void k_nll(int B, int P, const float *O, const float *T, float *G, float *loss) {
    __shared__ float red[4];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    float part = 0.f;
    if (e < B * P) {
        const int b = e / P, p = e % P;
        const float mu = O[(size_t)b * 2 * P + p], sr = O[(size_t)b * 2 * P + P + p];
        const float s = fminf(fmaxf(sr, -MLP_SMAX), MLP_SMAX);
        const float r = T[e] - mu, ex = expf(-s), sc = 1.0f / ((float)B * (float)P);
        part = 0.5f * (r * r * ex + s) * sc;
        G[(size_t)b * 2 * P + p] = -r * ex * sc;
        G[(size_t)b * 2 * P + P + p] = (sr > -MLP_SMAX && sr < MLP_SMAX) ? 0.5f * (1.0f - r * r * ex) * sc : 0.0f;
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

// Adam with bias correction folded into lr_t = lr sqrt(1 - b2^t) / (1 - b1^t).
__global__ __launch_bounds__(256) void k_adam(int n, float *w, const float *g, float *m, float *v, float lr_t,
                                              float b1, float b2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

int launch_check(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return negf_fail_internal(-100 - (int)e, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

}  // namespace

extern "C" {

int negf_mlp_gemm(int M, int N, int K, const float *A, int lda, int transA, const float *B, int ldb, int transB,
                  float *C, int ldc, const float *bias, float *Z, int mode, int accumulate, void *stream) {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C) return negf_fail_internal(-2, "bad arguments to negf_mlp_gemm");
    if (mode < 0 || mode > 2 || (mode >= 1 && !bias) || (mode == 1 && !Z))
        return negf_fail_internal(-2, "negf_mlp_gemm: mode 1 needs bias and Z, mode 2 needs bias");
    if ((transA ? lda < M : lda < K) || (transB ? ldb < K : ldb < N) || ldc < N)
        return negf_fail_internal(-2, "negf_mlp_gemm: leading dimension too small");
    if (M == 0 || N == 0) return 0;
    const dim3 grid((N + MLP_BN - 1) / MLP_BN, (M + MLP_BM - 1) / MLP_BM);
    hipLaunchKernelGGL(k_mlp_gemm, grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, transA, B, ldb,
                       transB, C, ldc, bias, Z, mode, accumulate);
    return launch_check("k_mlp_gemm");
}

int negf_mlp_gelu_backward(int n, const float *Z, const float *dY, float *dZ, void *stream) {
    if (n < 0 || !Z || !dY || !dZ) return negf_fail_internal(-2, "bad arguments to negf_mlp_gelu_backward");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gelu_backward, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, Z, dY, dZ);
    return launch_check("k_gelu_backward");
}

int negf_mlp_colsum(int M, int N, const float *X, int ldx, float *out, void *stream) {
    if (M < 0 || N < 0 || ldx < N || !X || !out) return negf_fail_internal(-2, "bad arguments to negf_mlp_colsum");
    if (N == 0) return 0;
    hipLaunchKernelGGL(k_colsum, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, M, N, X, ldx, out);
    return launch_check("k_colsum");
}

int negf_mlp_nll(int B, int P, const float *O, const float *T, float *G, float *loss, void *stream) {
    if (B < 0 || P < 0 || !O || !T || !G || !loss) return negf_fail_internal(-2, "bad arguments to negf_mlp_nll");
    if (B * P == 0) return 0;
    hipLaunchKernelGGL(k_nll, dim3((B * P + 255) / 256), dim3(256), 0, (hipStream_t)stream, B, P, O, T, G, loss);
    return launch_check("k_nll");
}

int negf_dev_alloc(void **ptr, size_t bytes) {
    if (!ptr) return negf_fail_internal(-2, "ptr is null");
    *ptr = nullptr;
    const hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) return negf_fail_internal(-100 - (int)e, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

int negf_dev_free(void *ptr) {
    if (!ptr) return 0;
    const hipError_t e = hipFree(ptr);
    if (e != hipSuccess) return negf_fail_internal(-100 - (int)e, std::string("hipFree: ") + hipGetErrorString(e));
    return 0;
}

int negf_dev_copy(void *dst, const void *src, size_t bytes, int kind) {
    if ((!dst || !src) && bytes) return negf_fail_internal(-2, "null pointer to negf_dev_copy");
    if (kind < 0 || kind > 2) return negf_fail_internal(-2, "negf_dev_copy: kind must be 0, 1 or 2");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    const hipError_t e = hipMemcpy(dst, src, bytes, k);
    if (e != hipSuccess) return negf_fail_internal(-100 - (int)e, std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

int negf_mlp_adam(int n, float *w, const float *g, float *m, float *v, float lr_t, float b1, float b2, float eps,
                  void *stream) {
    if (n < 0 || !w || !g || !m || !v) return negf_fail_internal(-2, "bad arguments to negf_mlp_adam");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, lr_t, b1,
                       b2, eps);
    return launch_check("k_adam");
}

}  // extern "C"
