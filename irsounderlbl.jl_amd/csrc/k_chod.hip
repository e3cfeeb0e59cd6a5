// Per-layer channel optical depths along the viewing path (the emulator's
// training target); tau_j are vertical layer optical depths, mu the cosine of
// the profile's zenith angle.
//
//   T^ch_k(c)   = sum_taps ILS(c) * exp(-sum_{j>=k} tau_j / mu)  level-to-space, k = 0..L
//   t^ch_k(c)   = sum_taps ILS(c) * exp(-tau_k / mu)             layer alone
//   tau^ch_k(c) = ln(T^ch_{k+1} / T^ch_k)          where T^ch_{k+1} >= floor
//               = -ln t^ch_k                       otherwise (the layer sits under an opaque
//                                                  column and does not reach the top)
//
// With these, prod_{j>=k} exp(-tau^ch_j) = T^ch_k wherever the column above
// transmits, so a channel-space RTE reproduces the ILS-convolved level
// transmittances of the scene it was computed on.
//
// A second pass gives the channel weight of the CIM source correction
// (src/Solver/schwarzschild.jl:130-209: the layer adds
// (B(T_top) - B(T_ave)) (T_{k+1} - T_k) f_cim(s_k), s_k = tau_k / mu):
//   G_k = sum_taps ILS * (T_{k+1} - T_k) f_cim(s_k),   g_k = G_k / T^ch_{k+1}
// (or, under an opaque column, the layer-alone sum_taps ILS (1 - e^{-s_k}) f_cim(s_k)),
// so that the channel RTE  sum_k B(T_ave)(T^ch_{k+1} - T^ch_k)
// + (B(T_top) - B(T_ave)) g_k T^ch_{k+1}  matches the ILS-convolved line-by-line
// radiance of the scene up to the variation of B across the ILS.  With raw_G
// the pass stores G_k itself (the emulator's level-transmittance target).
//
// One workgroup per (channel, profile); thread t walks tap points t, t+256, ...
// of the channel's window, forms the top-down optical depth in FP32 and keeps
// one accumulator per level and per layer in registers (static indices up to
// IRS_CHOD_MAXL layers), then a wave-shuffle + LDS reduction.
#include "irs_devmath.h"
#include "irs_internal.h"

#define IRS_CHOD_MAXL 64

template <int MAXL>
__global__ void __launch_bounds__(256) chod_kernel(const ChodArgs a) {
    __shared__ float red[4][2 * MAXL + 1];
    const int c = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
    const int L = a.n_layers;
    const float *tau = a.tau + (int64_t)q * L * a.n;
    const int64_t ic = a.ch_idx[c];
    const float imu = (float)(1.0 / a.scal[3 * q + 1]);
    float lev[MAXL + 1], lay[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) lev[k] = lay[k] = 0.0f;
    lev[MAXL] = 0.0f;
    const int64_t nt = 2 * a.half + 1;
    for (int64_t k0 = t; k0 < nt; k0 += 256) {
        const int64_t i = ic + k0 - a.half;
        if (i < 0 || i >= a.n) continue;
        const float w = (float)a.taps[k0];
        float cum = 0.0f;
        lev[MAXL] += w;                         // TOA: transmittance 1
#pragma unroll
        for (int k = MAXL - 1; k >= 0; --k) {
            if (k < L) {
                const float tk = tau[(int64_t)k * a.n + i] * imu;
                cum += tk;
                lev[k] += w * __expf(-cum);
                lay[k] += w * __expf(-tk);
            }
        }
    }
    // reduce 2*MAXL+1 values: lev[0..L-1], lev[MAXL] (top), lay[0..L-1]
    const int wv = t >> 6, ln = t & 63;
#pragma unroll
    for (int k = 0; k <= MAXL; ++k) {
        if (k < L || k == MAXL) {
            float v = lev[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (ln == 0) red[wv][k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            float v = lay[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (ln == 0) red[wv][MAXL + 1 + k] = v;
        }
    }
    __syncthreads();
    if (t < L) {
        const int k = t;
        const double sc = a.dnu;
        const double up = sc * ((double)red[0][k + 1 == L ? MAXL : k + 1] + red[1][k + 1 == L ? MAXL : k + 1] +
                                red[2][k + 1 == L ? MAXL : k + 1] + red[3][k + 1 == L ? MAXL : k + 1]);
        const double here = sc * ((double)red[0][k] + red[1][k] + red[2][k] + red[3][k]);
        const double alone = sc * ((double)red[0][MAXL + 1 + k] + red[1][MAXL + 1 + k] + red[2][MAXL + 1 + k] +
                                   red[3][MAXL + 1 + k]);
        double od;
        if (up >= a.floor && here > 0.0) od = log(up / here);
        else od = alone > 0.0 ? -log(alone) : (double)a.od_max;
        od = fmin(fmax(od, 0.0), (double)a.od_max);
        a.out[((int64_t)q * L + k) * a.n_ch + c] = (float)od;
        if (a.lev_out) {
            a.lev_out[((int64_t)q * (L + 1) + k) * a.n_ch + c] = (float)here;
            if (k == L - 1) a.lev_out[((int64_t)q * (L + 1) + L) * a.n_ch + c] = (float)up;
        }
    }
}

__device__ __forceinline__ float cim_f32(float s, float e) {   // f_cim, schwarzschild.jl:59-64
    return s < 0.06f ? s * (1.0f / 6.0f) : 1.0f - 2.0f * (e / expm1f(-s) + 1.0f / s);
}

template <int MAXL>
__global__ void __launch_bounds__(256) chod_cim_kernel(const ChodArgs a) {
    __shared__ float red[4][2 * MAXL];
    const int c = blockIdx.x, q = blockIdx.y, t = threadIdx.x;
    const int L = a.n_layers;
    const float *tau = a.tau + (int64_t)q * L * a.n;
    const int64_t ic = a.ch_idx[c];
    const float imu = (float)(1.0 / a.scal[3 * q + 1]);
    float G[MAXL], Ga[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) G[k] = Ga[k] = 0.0f;
    const int64_t nt = 2 * a.half + 1;
    for (int64_t k0 = t; k0 < nt; k0 += 256) {
        const int64_t i = ic + k0 - a.half;
        if (i < 0 || i >= a.n) continue;
        const float w = (float)a.taps[k0];
        float cum = 0.0f;
#pragma unroll
        for (int k = MAXL - 1; k >= 0; --k) {
            if (k < L) {
                const float sk = tau[(int64_t)k * a.n + i] * imu;
                const float e = __expf(-sk);
                const float top = __expf(-cum);
                cum += sk;
                const float f = cim_f32(sk, e);
                const float u = 1.0f - e;
                G[k] += w * (top * u) * f;
                Ga[k] += w * u * f;
            }
        }
    }
    const int wv = t >> 6, ln = t & 63;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            float v = G[k], va = Ga[k];
            for (int o = 32; o > 0; o >>= 1) {
                v += __shfl_xor(v, o, 64);
                va += __shfl_xor(va, o, 64);
            }
            if (ln == 0) {
                red[wv][k] = v;
                red[wv][MAXL + k] = va;
            }
        }
    }
    __syncthreads();
    if (t < L) {
        const int k = t;
        const double sc = a.dnu;
        const double g = sc * ((double)red[0][k] + red[1][k] + red[2][k] + red[3][k]);
        const double ga = sc * ((double)red[0][MAXL + k] + red[1][MAXL + k] + red[2][MAXL + k] + red[3][MAXL + k]);
        const double up = (double)a.lev_in[((int64_t)q * (L + 1) + k + 1) * a.n_ch + c];
        const double v = a.raw_G ? g : (up >= a.floor ? g / up : ga);
        a.g_out[((int64_t)q * L + k) * a.n_ch + c] = (float)fmin(fmax(v, 0.0), 1.0);
    }
}

hipError_t launch_chod(const ChodArgs &a, int n_prof, hipStream_t s) {
    if (a.n_layers > IRS_CHOD_MAXL || a.n_layers < 1) return hipErrorInvalidValue;
    if (a.g_out && !(a.lev_out && a.lev_in == a.lev_out)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(chod_kernel<IRS_CHOD_MAXL>, dim3((unsigned)a.n_ch, (unsigned)n_prof), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.g_out) return e;
    hipLaunchKernelGGL(chod_cim_kernel<IRS_CHOD_MAXL>, dim3((unsigned)a.n_ch, (unsigned)n_prof), dim3(256), 0, s, a);
    return hipGetLastError();
}
