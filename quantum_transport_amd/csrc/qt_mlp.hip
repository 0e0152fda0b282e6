// qt_mlp.hip -- the small multilayer perceptron of the learned lead warm
// start: (E, kz) features -> diag + rank-R factors of the lead's surface
// Green's function g00 = diag(d) + U V^H.  Weights are BF16 (with FP32
// master copies for Adam), products accumulate in FP32 on the BF16 matrix
// cores, activations are kept in FP32 and rounded to BF16 where they enter a
// product.  Row-major storage throughout (a batch of samples is a matrix
// with one sample per row).
//
// Kernels: a tiled BF16 GEMM on v_mfma_f32_16x16x16_bf16 (64 x 64 output
// tiles, K staged through LDS in chunks of 32, either operand optionally
// transposed), GELU forward / backward, FP32 -> BF16 rounding, bias
// gradients (column sums), Adam, and the conversions between the network's
// output row and the complex128 factors d, U, V.
#include <cmath>

#include "qt_internal.h"

namespace qt {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs), the rounding of the CPU oracle
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

constexpr int TM = 64, TN = 64, KC = 32, KP = KC + 4;  // LDS rows padded to 36 bf16

// C[m][n] = sum_k opA(m, k) opB(k, n) (+ bias[n]).
//   transA = 0: A is M x K, (m, k) at A[m*lda + k];  1: A is K x M, at A[k*lda + m]
//   transB = 0: B is K x N, (k, n) at B[k*ldb + n];  1: B is N x K, at B[n*ldb + k]
// K % 32 == 0, the leading dimensions are multiples of 8; rows beyond M / N
// read as zero and are not stored.
template <int TA, int TB>
__global__ __launch_bounds__(256) void bf16_gemm_kernel(int M, int N, int K, const uint16_t* __restrict__ A,
                                                        int lda, const uint16_t* __restrict__ B, int ldb,
                                                        float* __restrict__ C, int ldc,
                                                        const float* __restrict__ bias) {
    __shared__ uint16_t As[TM][KP];  // As[m][k]
    __shared__ uint16_t Bt[TN][KP];  // Bt[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += KC) {
        // 64 x 32 bf16 per operand tile = 256 chunks of 8 values, one per thread
        {
            uint16_t v[8];
            if (TA == 0) {
                const int r = tid >> 2, kk = (tid & 3) * 8, m = m0 + r;
                if (m < M) {
                    const uint4 q = *(const uint4*)(A + (int64_t)m * lda + k0 + kk);
                    *(uint4*)v = q;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = 0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) As[r][kk + e] = v[e];
            } else {
                const int kk = tid >> 3, r = (tid & 7) * 8;  // 8 consecutive m at fixed k
                const uint16_t* src = A + (int64_t)(k0 + kk) * lda + m0 + r;
                if (m0 + r + 8 <= M) {
                    *(uint4*)v = *(const uint4*)src;  // 16-byte aligned: lda % 8 == 0, m0 + r % 8 == 0
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = m0 + r + e < M ? src[e] : (uint16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) As[r + e][kk] = v[e];
            }
        }
        {
            uint16_t v[8];
            if (TB == 1) {
                const int c = tid >> 2, kk = (tid & 3) * 8, n = n0 + c;
                if (n < N) {
                    const uint4 q = *(const uint4*)(B + (int64_t)n * ldb + k0 + kk);
                    *(uint4*)v = q;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = 0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) Bt[c][kk + e] = v[e];
            } else {
                const int kk = tid >> 3, c = (tid & 7) * 8;  // 8 consecutive n at fixed k
                const uint16_t* src = B + (int64_t)(k0 + kk) * ldb + n0 + c;
                if (n0 + c + 8 <= N) {
                    *(uint4*)v = *(const uint4*)src;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = n0 + c + e < N ? src[e] : (uint16_t)0;
                }
#pragma unroll
                for (int e = 0; e < 8; ++e) Bt[c + e][kk] = v[e];
            }
        }
        __syncthreads();
        // lane l: row / column l % 16 of a 16 x 16 block, k = 4 (l / 16) + 0..3
        const int fr = lane & 15, fk = (lane >> 4) * 4;
#pragma unroll
        for (int ks = 0; ks < KC; ks += 16) {
            short4_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *(const short4_t*)&As[wm + 16 * i + fr][ks + fk];  // 8-byte aligned
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *(const short4_t*)&Bt[wn + 16 * j + fr][ks + fk];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D lane layout: column l % 16, rows 4 (l / 16) + 0..3
    const int col_l = lane & 15, row_l = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + wn + 16 * j + col_l;
            if (n >= N) continue;
            const float bb = bias ? bias[n] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + 16 * i + row_l + q;
                if (m < M) C[(int64_t)m * ldc + n] = acc[i][j][q] + bb;
            }
        }
}

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void gelu_fwd_kernel(int64_t n, const float* Z, uint16_t* H) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        H[i] = f2bf(gelu_f(Z[i]));
}

__global__ void gelu_bwd_kernel(int64_t n, const float* dA, const float* Z, float* dZ) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dZ[i] = dA[i] * gelu_d(Z[i]);
}

__global__ void to_bf16_kernel(int64_t n, const float* X, uint16_t* Y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Y[i] = f2bf(X[i]);
}

// db[n] = sum over the rows of dZ (rows x cols)
__global__ void colsum_kernel(int rows, int cols, const float* dZ, float* db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += dZ[(int64_t)r * cols + c];
    db[c] = s;
}

__global__ void adam_kernel(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* wbf, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wbf) wbf[i] = f2bf(wi);
    }
}

// network output row (2n + 4nR values, row stride W >= that: d re | d im |
// U re | U im | V re | V im, U(i, k) at i R + k) -> complex128 d [b][n] and column-major U, V [b][n x R]
__global__ void out_to_lowrank_kernel(int batch, int n, int R, const float* Y, int W, double2* d, double2* U,
                                      double2* V) {
    const int b = blockIdx.y;
    const float* y = Y + (int64_t)b * W;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * R; t += gridDim.x * blockDim.x) {
        const int i = t / R, k = t % R;
        const int64_t o = (int64_t)b * n * R + i + (int64_t)k * n;
        U[o] = make_double2(y[2 * n + t], y[2 * n + n * R + t]);
        V[o] = make_double2(y[2 * n + 2 * n * R + t], y[2 * n + 3 * n * R + t]);
        if (k == 0) d[(int64_t)b * n + i] = make_double2(y[i], y[n + i]);
    }
}

// gradients with respect to the real and imaginary parts of d, U, V
// (complex128, same layouts) -> network output gradient row
// gradient row of member b from the residual E (its diagonal is the
// gradient of d) and dU = E V, dV = E^H U, all scaled by scale[b]
__global__ void lowrank_to_out_kernel(int batch, int n, int R, const double2* E, int64_t sE, const double2* dU,
                                      const double2* dV, const double* scale, float* dY, int W) {
    const int b = blockIdx.y;
    const double sc = scale[b];
    float* y = dY + (int64_t)b * W;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n * R; t += gridDim.x * blockDim.x) {
        const int i = t / R, k = t % R;
        const int64_t o = (int64_t)b * n * R + i + (int64_t)k * n;
        y[2 * n + t] = (float)(sc * dU[o].x);
        y[2 * n + n * R + t] = (float)(sc * dU[o].y);
        y[2 * n + 2 * n * R + t] = (float)(sc * dV[o].x);
        y[2 * n + 3 * n * R + t] = (float)(sc * dV[o].y);
        if (k == 0) {
            const double2 e = E[b * sE + (int64_t)i * (n + 1)];  // diagonal of the residual
            y[i] = (float)(sc * e.x);
            y[n + i] = (float)(sc * e.y);
        }
    }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

}  // namespace

void mlp_gemm(hipStream_t s, int transA, int transB, int M, int N, int K, const uint16_t* A, int lda,
              const uint16_t* B, int ldb, float* C, int ldc, const float* bias) {
    if (M <= 0 || N <= 0) return;
    if (K % KC != 0 || lda % 8 != 0 || ldb % 8 != 0) throw hip_error("mlp_gemm: K % 32 or ld % 8", 1);
    const dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    if (transA == 0 && transB == 0)
        hipLaunchKernelGGL((bf16_gemm_kernel<0, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias);
    else if (transA == 0 && transB == 1)
        hipLaunchKernelGGL((bf16_gemm_kernel<0, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias);
    else if (transA == 1 && transB == 0)
        hipLaunchKernelGGL((bf16_gemm_kernel<1, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias);
    else
        hipLaunchKernelGGL((bf16_gemm_kernel<1, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias);
    QT_HIP(hipGetLastError());
}

void mlp_gelu_fwd(hipStream_t s, int64_t n, const float* Z, uint16_t* H) {
    hipLaunchKernelGGL(gelu_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Z, H);
    QT_HIP(hipGetLastError());
}
void mlp_gelu_bwd(hipStream_t s, int64_t n, const float* dA, const float* Z, float* dZ) {
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, dA, Z, dZ);
    QT_HIP(hipGetLastError());
}
void mlp_to_bf16(hipStream_t s, int64_t n, const float* X, uint16_t* Y) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, X, Y);
    QT_HIP(hipGetLastError());
}
void mlp_colsum(hipStream_t s, int rows, int cols, const float* dZ, float* db) {
    hipLaunchKernelGGL(colsum_kernel, dim3((cols + 255) / 256), dim3(256), 0, s, rows, cols, dZ, db);
    QT_HIP(hipGetLastError());
}
void mlp_adam(hipStream_t s, int64_t n, float* w, const float* g, float* m, float* v, uint16_t* wbf, float lr,
              float b1, float b2, float eps, int step) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, w, g, m, v, wbf, lr, b1, b2, eps, c1, c2);
    QT_HIP(hipGetLastError());
}
void mlp_out_to_lowrank(hipStream_t s, int batch, int n, int R, const float* Y, int ldy, double2* d, double2* U,
                        double2* V) {
    hipLaunchKernelGGL(out_to_lowrank_kernel, dim3((n * R + 255) / 256, batch), dim3(256), 0, s, batch, n, R, Y,
                       ldy, d, U, V);
    QT_HIP(hipGetLastError());
}
void mlp_lowrank_to_out(hipStream_t s, int batch, int n, int R, const double2* E, int64_t sE, const double2* dU,
                        const double2* dV, const double* scale, float* dY, int ldy) {
    hipLaunchKernelGGL(lowrank_to_out_kernel, dim3((n * R + 255) / 256, batch), dim3(256), 0, s, batch, n, R, E,
                       sE, dU, dV, scale, dY, ldy);
    QT_HIP(hipGetLastError());
}

}  // namespace qt
