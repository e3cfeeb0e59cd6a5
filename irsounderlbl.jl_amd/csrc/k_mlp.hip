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
    const int ks = g.splitk > 1 ? g.K / g.splitk : g.K;
    const int kbeg = (int)blockIdx.z * ks, kend = kbeg + ks;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
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
    if (g.splitk > 1) {
        float *cp = g.Cpart + (int64_t)blockIdx.z * g.M * g.ldc;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + rr;
                    const int n = n0 + wn + 16 * j + (lane & 15);
                    cp[(int64_t)m * g.ldc + n] = acc[i][j][rr] * g.alpha;
                }
        return;
    }
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

// Split-K reduction: C[m][n] = sum over slices in order (deterministic).
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const float *part, int S, int M, int N, int ldc,
                                                            float *C) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)M * N) return;
    const int64_t o = (i / N) * ldc + i % N;
    const int64_t stride = (int64_t)M * ldc;
    float s = 0.0f;
    for (int k = 0; k < S; ++k) s += part[k * stride + o];
    C[o] = s;
}

hipError_t launch_gemm_bf16(const GemmArgs &g, bool ta, bool tb, hipStream_t s) {
    if (g.M % GT || g.N % GT || g.K % GK || g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    if (g.splitk > 1 && (!g.Cpart || !g.C32 || g.C16 || g.bias || g.flags || (g.K / g.splitk) % GK ||
                         g.K % g.splitk))
        return hipErrorInvalidValue;
    dim3 grid(g.N / GT, g.M / GT, g.splitk > 1 ? g.splitk : 1);
    if (!ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, g);
    else if (!ta && tb) hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, g);
    else if (ta && !tb) hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, g);
    if (g.splitk > 1) {
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int64_t n = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g.Cpart,
                           g.splitk, g.M, g.N, g.ldc, g.C32);
    }
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

// db[n] = sum_m dZ[m][n] (dZ given in bf16 as fed to the GEMMs).  Rows are cut into
// S slices; a workgroup covers 64 consecutive columns (one per lane, coalesced
// rows) of four slices, each thread a compensated sum over its slice; the slice
// partials are then summed in slice order, so the result does not depend on
// scheduling.
__global__ void __launch_bounds__(256) colsum_part_kernel(const uint16_t *dZ, int M, int N, int S, float *part) {
    const int t = threadIdx.x;
    const int n = blockIdx.x * 64 + (t & 63);
    const int sl = blockIdx.y * 4 + (t >> 6);
    if (n >= N || sl >= S) return;
    const int rows = (M + S - 1) / S, m0 = sl * rows, m1 = min(M, m0 + rows);
    float s = 0.0f, c = 0.0f;
    for (int m = m0; m < m1; ++m) {
        const float y = bf2f(dZ[(int64_t)m * N + n]) - c;
        const float tt = s + y;
        c = (tt - s) - y;
        s = tt;
    }
    part[(int64_t)sl * N + n] = s;
}

__global__ void __launch_bounds__(256) colsum_reduce_kernel(const float *part, int N, int S, float *db) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f, c = 0.0f;
    for (int k = 0; k < S; ++k) {
        const float y = part[(int64_t)k * N + n] - c;
        const float tt = s + y;
        c = (tt - s) - y;
        s = tt;
    }
    db[n] = s;
}

hipError_t launch_colsum_bf16(const uint16_t *dZ, int M, int N, float *db, float *scratch, int64_t scratch_floats,
                              hipStream_t s) {
    int S = 1;
    while (S < 128 && M / (2 * S) >= 16 && (int64_t)2 * S * N <= scratch_floats) S *= 2;
    S = (S + 3) / 4 * 4;
    if ((int64_t)S * N > scratch_floats) S = 4;
    if ((int64_t)S * N > scratch_floats) return hipErrorInvalidValue;
    hipLaunchKernelGGL(colsum_part_kernel, dim3((N + 63) / 64, S / 4), dim3(256), 0, s, dZ, M, N, S, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(colsum_reduce_kernel, dim3((N + 255) / 256), dim3(256), 0, s, scratch, N, S, db);
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
