// Emulated forward model: per (profile, channel) two targets per layer are
// decoded from the MLP's PCA coefficients,
//     x_k(c) = mean_k(c) + sum_m a_{k,m} V_m(c)    (first target on columns
//                                                  0..n_ch-1, second on n_ch..2 n_ch-1),
// either (kind 0) the layer channel transmittance t_k = T_k / T_{k+1} and the
// CIM weight g_k = G_k / T_{k+1} (irs_channel_od), or (kind 1) the level-to-space
// transmittance T_k and the CIM sum G_k (irs_channel_transmittance), and fed to
// the channel-space RTE with the CIM source of the line-by-line RTE
// (src/Solver/schwarzschild.jl:130-209) summed over the ILS:
//     R = eps B(T_sfc) T_0 + sum_k B(T_ave,k)(T_{k+1} - T_k) + (B(T_{k+1}) - B(T_ave,k)) G_k,
// Planck at the channel centre, T_L = 1.  The transmittances are along the
// viewing path (the emulator is trained with 1/mu as an input).  For eps < 1
// the reflected downwelling term uses the layer transmittances T_k / T_{k+1}
// with layer-mean Planck emission (an approximation; eps = 1 is exact).
//
// One thread per channel, one workgroup row per profile; the profile's
// coefficients [L][n_pc] sit in LDS and are read as broadcasts, the basis is
// read coalesced once per thread (2 n_pc loads), the decode keeps two FP32
// accumulators per layer, the RTE runs in FP64.
#include "irs_devmath.h"
#include "irs_internal.h"
#include "irs_mlp.h"

#define EMU_MAXL 64

template <int MAXL>
__global__ void __launch_bounds__(256) emu_rte_kernel(const EmuRteArgs a) {
    extern __shared__ float s_coef[];          // [L][n_pc]
    const int q = blockIdx.y, L = a.L, npc = a.n_pc;
    const float *cq = a.coef + (int64_t)q * L * npc;
    for (int i = threadIdx.x; i < L * npc; i += 256) s_coef[i] = cq[i] * a.coef_scale[i % npc];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t W = 2 * (int64_t)a.n_ch;
    float T[MAXL], G[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        T[k] = k < L ? a.t_mean[(int64_t)k * W + c] : 1.0f;
        G[k] = k < L ? a.t_mean[(int64_t)k * W + a.n_ch + c] : 0.0f;
    }
    for (int m = 0; m < npc; ++m) {
        const float v = a.basis[(int64_t)m * W + c];
        const float vg = a.basis[(int64_t)m * W + a.n_ch + c];
#pragma unroll
        for (int k = 0; k < MAXL; ++k)
            if (k < L) {
                T[k] = fmaf(s_coef[k * npc + m], v, T[k]);
                G[k] = fmaf(s_coef[k * npc + m], vg, G[k]);
            }
    }
    const double nu = a.ch_nu[c];
    const double *T_lev = a.T_lev + (int64_t)q * (L + 1), *T_ave = a.T_ave + (int64_t)q * L;
    const double T_sfc = a.scal[3 * q], eps = a.scal[3 * q + 2];
    double tr_top = 1.0, acc = 0.0;
    const bool ratio = a.target_kind == 0;
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
        if (k < L) {
            const double x = fmin(fmax((double)T[k], 0.0), 1.0);
            const double tr_bot = ratio ? tr_top * x : x;
            const double Gk = ratio ? (double)fminf(fmaxf(G[k], 0.0f), 1.0f) * tr_top : (double)G[k];
            if (a.t_out) a.t_out[((int64_t)q * L + k) * a.n_ch + c] = (float)tr_bot;
            const double ba = planck64(nu, T_ave[k]), bt = planck64(nu, T_lev[k + 1]);
            acc += ba * (tr_top - tr_bot) + (bt - ba) * Gk;
            T[k] = (float)tr_bot;
            tr_top = tr_bot;
        }
    }
    double I = eps * planck64(nu, T_sfc) * tr_top + acc;
    if (eps < 1.0) {
        double down = 0.0, g = 1.0, up = 1.0;
#pragma unroll
        for (int k = 0; k < MAXL; ++k) {
            if (k < L) {
                const double tb = (double)T[k];                         // level transmittances now
                const double tt = k + 1 < L ? (double)T[k + 1] : 1.0;
                const double tl = tt > 1e-12 ? fmin(tb / tt, 1.0) : 0.0;
                down += g * (1.0 - tl) * planck64(nu, T_ave[k]);
                g *= tl;
                (void)up;
            }
        }
        I += (1.0 - eps) * down * tr_top;
    }
    const int64_t o = (int64_t)q * a.n_ch + c;
    a.R[o] = I;
    a.BT[o] = I <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / I);
}

hipError_t launch_emu_rte(const EmuRteArgs &a, hipStream_t s) {
    if (a.L < 1 || a.L > EMU_MAXL) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)a.L * a.n_pc;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(emu_rte_kernel<EMU_MAXL>, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)a.n_prof),
                       dim3(256), lds, s, a);
    return hipGetLastError();
}
