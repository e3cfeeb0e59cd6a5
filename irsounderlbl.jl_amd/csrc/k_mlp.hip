// BF16 MFMA kernels of the emulator MLP: GEMM with fused bias / GELU /
// GELU-derivative epilogues (FP32 accumulation), MSE gradient, bias
// gradients, Adam, conversions.
//
// GEMM: C[M x N] = A[M x K] . B[K x N], bf16 operands from LDS into
// v_mfma_f32_16x16x32_bf16.  A workgroup (4 waves) owns a 64 x 64 tile of C,
// each wave a 32 x 32 quarter (2 x 2 MFMA blocks); K advances 32 per step
// through LDS tiles stored k-contiguous ([row][k] for A, [col][k] for B) so a
// lane reads its 8 consecutive k with one 16-byte LDS load.  Operand lane
// layout of the 16x16x32 bf16 MFMA: lane l supplies A[l % 16][8 (l / 16) + 0..7]
// and B[8 (l / 16) + 0..7][l % 16]; it receives C[4 (l / 16) + 0..3][l % 16].
// Transposed operands (A = X^T for weight gradients, B = W^T for input
// gradients) are transposed while the tile is staged into LDS.
// M, N must be multiples of 64 and K of 32 (the caller pads).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "irs_mlp.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_d(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

#define GT 64       // C tile
#define GK 32       // K step
#define GPAD 8      // LDS row padding (bf16 elements)

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[GT][GK + GPAD];
    __shared__ __attribute__((aligned(16))) uint16_t sB[GT][GK + GPAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        // stage A(m, k) for m in [m0, m0+64), k in [k0, k0+32): 2048 values, 8 per thread
        if (!TA) {   // A[m * lda + k], k contiguous
            const int m = t >> 2, k = (t & 3) * 8;
            const uint4 v = *(const uint4 *)(g.A + (int64_t)(m0 + m) * g.lda + k0 + k);
            *(uint4 *)&sA[m][k] = v;
        } else {     // A[k * lda + m], m contiguous
            const int k = t >> 3, m = (t & 7) * 8;
            const uint4 v = *(const uint4 *)(g.A + (int64_t)(k0 + k) * g.lda + m0 + m);
            const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
            for (int j = 0; j < 8; ++j) sA[m + j][k] = e[j];
        }
        if (TB) {    // B[n * ldb + k], k contiguous
            const int n = t >> 2, k = (t & 3) * 8;
            const uint4 v = *(const uint4 *)(g.B + (int64_t)(n0 + n) * g.ldb + k0 + k);
            *(uint4 *)&sB[n][k] = v;
        } else {     // B[k * ldb + n], n contiguous
            const int k = t >> 3, n = (t & 7) * 8;
            const uint4 v = *(const uint4 *)(g.B + (int64_t)(k0 + k) * g.ldb + n0 + n);
            const uint16_t *e = (const uint16_t *)&v;
#pragma unroll
            for (int j = 0; j < 8; ++j) sB[n + j][k] = e[j];
        }
        __syncthreads();
        const int r = lane & 15, kq = (lane >> 4) * 8;
        bf16x8_t a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const uint4 va = *(const uint4 *)&sA[wm + 16 * i + r][kq];
            a[i] = __builtin_bit_cast(bf16x8_t, va);
            const uint4 vb = *(const uint4 *)&sB[wn + 16 * i + r][kq];
            b[i] = __builtin_bit_cast(bf16x8_t, vb);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // epilogue: lane holds C[wm + 16 i + 4 (lane / 16) + r][wn + 16 j + lane % 16]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + rr;
                const int n = n0 + wn + 16 * j + (lane & 15);
                float v = acc[i][j][rr] * g.alpha;
                if (g.bias) v += g.bias[n];
                const int64_t o = (int64_t)m * g.ldc + n;
                if (g.flags & GEMM_DGELU) v *= gelu_d(g.Zin[o]);
                if (g.flags & GEMM_GELU) {
                    if (g.C32) g.C32[o] = v;
                    if (g.C16) g.C16[o] = f2bf(gelu(v));
                } else {
                    if (g.C32) g.C32[o] = v;
                    if (g.C16) g.C16[o] = f2bf(v);
                }
            }
}

hipError_t launch_gemm_bf16(const GemmArgs &g, bool ta, bool tb, hipStream_t s) {
    if (g.M % GT || g.N % GT || g.K % GK || g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    dim3 grid(g.N / GT, g.M / GT);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

// ---- elementwise / reductions ----------------------------------------------
__global__ void f32_to_bf16_kernel(const float *x, uint16_t *y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

hipError_t launch_f32_to_bf16(const float *x, uint16_t *y, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

// Gathers rows idx[r] of X [*, n_in] (f32) into a padded bf16 batch [B, ld]
// (zeros beyond n_in), and the matching target rows into T [B, n_out] f32.
__global__ void gather_batch_kernel(const float *X, const float *Tsrc, const int32_t *idx, int B, int n_in,
                                    int ld, int n_out, uint16_t *xb, float *tb) {
    const int r = blockIdx.x;
    const int64_t src = idx ? idx[r] : r;
    for (int c = threadIdx.x; c < ld; c += 256) xb[(int64_t)r * ld + c] = c < n_in ? f2bf(X[src * n_in + c]) : 0;
    if (Tsrc && tb)
        for (int c = threadIdx.x; c < n_out; c += 256) tb[(int64_t)r * n_out + c] = Tsrc[src * n_out + c];
}

hipError_t launch_gather_batch(const float *X, const float *T, const int32_t *idx, int B, int n_in, int ld,
                               int n_out, uint16_t *xb, float *tb, hipStream_t s) {
    hipLaunchKernelGGL(gather_batch_kernel, dim3(B), dim3(256), 0, s, X, T, idx, B, n_in, ld, n_out, xb, tb);
    return hipGetLastError();
}

// dY = scale * (Y - T) (bf16 for the next GEMM) and per-block partial sums of (Y - T)^2.
__global__ void mse_grad_kernel(const float *Y, const float *T, int64_t n, float scale, uint16_t *dY,
                                float *partial) {
    __shared__ float red[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float e2 = 0.0f;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = f2bf(scale * d);
        e2 = d * d;
    }
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_xor(e2, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_mse_grad(const float *Y, const float *T, int64_t n, float scale, uint16_t *dY, float *partial,
                           hipStream_t s) {
    hipLaunchKernelGGL(mse_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Y, T, n, scale, dY,
                       partial);
    return hipGetLastError();
}

// db[n] = sum_m dZ[m][n] (dZ given in bf16 as fed to the GEMMs); one thread per column.
__global__ void colsum_bf16_kernel(const uint16_t *dZ, int M, int N, float *db) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f, c = 0.0f;
    for (int m = 0; m < M; ++m) {   // compensated: M up to a few thousand
        const float y = bf2f(dZ[(int64_t)m * N + n]) - c;
        const float tt = s + y;
        c = (tt - s) - y;
        s = tt;
    }
    db[n] = s;
}

hipError_t launch_colsum_bf16(const uint16_t *dZ, int M, int N, float *db, hipStream_t s) {
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((N + 255) / 256), dim3(256), 0, s, dZ, M, N, db);
    return hipGetLastError();
}

// Adam (Kingma & Ba 2015) on FP32 master parameters; refreshes the bf16 copy
// of weight entries (w16 may be null for biases).
__global__ void adam_kernel(float *p, const float *gr, float *m, float *v, uint16_t *w16, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    if (w16) w16[i] = f2bf(pi);
}

hipError_t launch_adam(float *p, const float *g, float *m, float *v, uint16_t *w16, int64_t n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, g, m, v, w16, n, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}

// Tangent of GELU for forward-mode input derivatives: dH = gelu'(Z) * dZ (bf16 out).
__global__ void gelu_tangent_kernel(const float *Z, const float *dZ, int64_t n, uint16_t *dH) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dH[i] = f2bf(gelu_d(Z[i]) * dZ[i]);
}

hipError_t launch_gelu_tangent(const float *Z, const float *dZ, int64_t n, uint16_t *dH, hipStream_t s) {
    hipLaunchKernelGGL(gelu_tangent_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Z, dZ, n, dH);
    return hipGetLastError();
}

// out[0] = scale * sum(partial[0..n))  (one workgroup, FP64 accumulate)
__global__ void sum_partials_kernel(const float *partial, int n, float scale, float *out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += (double)partial[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (float)(scale * (red[0] + red[1] + red[2] + red[3]));
}

hipError_t launch_sum_partials(const float *partial, int n, float scale, float *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, n, scale, out);
    return hipGetLastError();
}

// out[b][n] = row[n]
__global__ void broadcast_row_kernel(const float *row, int B, int N, float *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (int64_t)B * N) out[i] = row[i % N];
}

hipError_t launch_broadcast_row(const float *row, int B, int N, float *out, hipStream_t s) {
    const int64_t n = (int64_t)B * N;
    hipLaunchKernelGGL(broadcast_row_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, row, B, N, out);
    return hipGetLastError();
}
