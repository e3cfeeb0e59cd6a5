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
    for (int k = 0; k < MAXL; ++k) T[k] = G[k] = 0.0f;
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
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            const int64_t o = (int64_t)k * W + c;
            const float wt = a.inv_w ? a.inv_w[o] : 1.0f, wg = a.inv_w ? a.inv_w[o + a.n_ch] : 1.0f;
            T[k] = fmaf(T[k], wt, a.t_mean[o]);
            G[k] = fmaf(G[k], wg, a.t_mean[o + a.n_ch]);
        } else {
            T[k] = 1.0f;
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

__device__ __forceinline__ double dbdt_c(double nu, double T) {
    const double x = IRS_C2_PL * nu / T, ex = exp(x);
    return IRS_C1_PL * IRS_C2_PL * nu * nu * nu * nu * ex / (T * T * (ex - 1.0) * (ex - 1.0));
}

// Adjoint of the emulated channel RTE (target kind 0: t_k, g_k).  With
// T_k = x_k T_{k+1} (x_k = clamp(t_k)), R = eps B_s T_0 + sum_k [ba_k (T_{k+1} - T_k)
// + (bt_k - ba_k) g_k T_{k+1}], the total derivative lambda_k = dR/dT_k obeys
//   lambda_0 = eps B_s - ba_0,
//   lambda_k = -ba_k + ba_{k-1} + (bt_{k-1} - ba_{k-1}) g_{k-1} + lambda_{k-1} x_{k-1},
// and dR/dt_k = lambda_k T_{k+1}, dR/dg_k = (bt_k - ba_k) T_{k+1}.  Every output is
// multiplied by dBT/dR.  The reflected term (eps < 1) is held fixed.
template <int MAXL>
__global__ void __launch_bounds__(256) emu_adjoint_kernel(const EmuRteArgs a, EmuAdjArgs j) {
    extern __shared__ float s_coef[];
    const int q = blockIdx.y, L = a.L, npc = a.n_pc;
    const float *cq = a.coef + (int64_t)q * L * npc;
    for (int i = threadIdx.x; i < L * npc; i += 256) s_coef[i] = cq[i] * a.coef_scale[i % npc];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t W = 2 * (int64_t)a.n_ch;
    float T[MAXL], G[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) T[k] = G[k] = 0.0f;
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
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            const int64_t o = (int64_t)k * W + c;
            const float wt = a.inv_w ? a.inv_w[o] : 1.0f, wg = a.inv_w ? a.inv_w[o + a.n_ch] : 1.0f;
            T[k] = fmaf(T[k], wt, a.t_mean[o]);
            G[k] = fmaf(G[k], wg, a.t_mean[o + a.n_ch]);
        } else {
            T[k] = 1.0f;
        }
    }
    const double nu = a.ch_nu[c];
    const double *T_lev = a.T_lev + (int64_t)q * (L + 1), *T_ave = a.T_ave + (int64_t)q * L;
    const double T_sfc = a.scal[3 * q], eps = a.scal[3 * q + 2];
    // forward: keep T_{k+1} (in T[k] after the pass: transmittance above layer k)
    double tr_top = 1.0, acc = 0.0;
    float above[MAXL];
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
        if (k < L) {
            const float xr = T[k];
            const double x = fmin(fmax((double)xr, 0.0), 1.0);
            const double g = fmin(fmax((double)G[k], 0.0), 1.0);
            above[k] = (float)tr_top;
            const double tr_bot = tr_top * x;
            const double ba = planck64(nu, T_ave[k]), bt = planck64(nu, T_lev[k + 1]);
            acc += ba * (tr_top - tr_bot) + (bt - ba) * g * tr_top;
            T[k] = (xr > 0.0f && xr < 1.0f) ? xr : (xr <= 0.0f ? -0.0f : 2.0f);   // keep the clamp state
            G[k] = (float)g;
            tr_top = tr_bot;
        }
    }
    const double bs = planck64(nu, T_sfc);
    const double I = eps * bs * tr_top + acc;
    const double dbt = I > 0.0 ? ({
        const double A = IRS_C1_PL * nu * nu * nu, Lg = log1p(A / I);
        IRS_C2_PL * nu * A / (Lg * Lg * I * (I + A));
    }) : 0.0;
    const int64_t o = (int64_t)q * a.n_ch + c;
    j.BT[o] = I <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / I);
    j.dTsfc[o] = dbt * eps * tr_top * dbdt_c(nu, T_sfc);
    double lam = 0.0, prev_x = 0.0, prev_term = 0.0;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < L) {
            const double ba = planck64(nu, T_ave[k]), bt = planck64(nu, T_lev[k + 1]);
            const double x = fmin(fmax((double)T[k], 0.0), 1.0);
            const bool live = T[k] > 0.0f && T[k] < 1.0f;
            const double g = (double)G[k], up = (double)above[k];
            lam = k == 0 ? eps * bs - ba : -ba + prev_term + lam * prev_x;
            const int64_t ok = ((int64_t)q * L + k) * a.n_ch + c;
            j.dt[ok] = (float)(live ? dbt * lam * up : 0.0);
            const bool glive = g > 0.0 && g < 1.0;
            j.dg[ok] = (float)(glive ? dbt * (bt - ba) * up : 0.0);
            j.dTave[ok] = (float)(dbt * (up - up * x - g * up) * dbdt_c(nu, T_ave[k]));
            j.dTlev[((int64_t)q * (L + 1) + k + 1) * a.n_ch + c] = (float)(dbt * g * up * dbdt_c(nu, T_lev[k + 1]));
            prev_term = ba + (bt - ba) * g;
            prev_x = x;
        }
    }
    j.dTlev[((int64_t)q * (L + 1)) * a.n_ch + c] = 0.0f;
}

// dBT/dfeature: F[q][k][i][c] = sum_m Jm[q][k][m][i] (dt[q][k][c] V_m(c) iw_t + dg[q][k][c] Vg_m(c) iw_g) * scale_m
// (iw: the decode weights of the layer and channel).
// One workgroup per (channel block, layer, profile); J of the layer in LDS.
__global__ void __launch_bounds__(256) emu_feature_jac_kernel(const EmuRteArgs a, EmuAdjArgs j, int n_in) {
    extern __shared__ float s_j[];             // [n_pc][n_in]
    const int k = blockIdx.y, q = blockIdx.z, L = a.L, npc = a.n_pc;
    const float *Jk = j.Jmlp + (((int64_t)q * L + k) * npc) * n_in;
    for (int i = threadIdx.x; i < npc * n_in; i += 256) s_j[i] = Jk[i] * a.coef_scale[i / n_in];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t W = 2 * (int64_t)a.n_ch;
    const int64_t ok = ((int64_t)q * L + k) * a.n_ch + c;
    const float at = j.dt[ok] * (a.inv_w ? a.inv_w[(int64_t)k * W + c] : 1.0f);
    const float ag = j.dg[ok] * (a.inv_w ? a.inv_w[(int64_t)k * W + a.n_ch + c] : 1.0f);
    float acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    for (int m = 0; m < npc; ++m) {
        const float w = at * a.basis[(int64_t)m * W + c] + ag * a.basis[(int64_t)m * W + a.n_ch + c];
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (i < n_in) acc[i] = fmaf(s_j[m * n_in + i], w, acc[i]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
        if (i < n_in) j.dfeat[(((int64_t)q * L + k) * n_in + i) * a.n_ch + c] = acc[i];
}

hipError_t launch_emu_adjoint(const EmuRteArgs &a, const EmuAdjArgs &j, int n_in, hipStream_t s) {
    if (a.L < 1 || a.L > EMU_MAXL || n_in < 1 || n_in > 16) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)a.L * a.n_pc;
    hipLaunchKernelGGL(emu_adjoint_kernel<EMU_MAXL>, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)a.n_prof),
                       dim3(256), lds, s, a, j);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(emu_feature_jac_kernel, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)a.L, (unsigned)a.n_prof),
                       dim3(256), sizeof(float) * (size_t)a.n_pc * n_in, s, a, j, n_in);
    return hipGetLastError();
}
