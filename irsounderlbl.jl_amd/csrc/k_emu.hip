// Emulated forward model.  Per profile, layer k and channel c three targets
//     y_k(c) = (t_k, g_k, m_k): the layer channel transmittance t_k = T_k / T_{k+1},
//     the CIM weight g_k = G_k / T_{k+1} and the normalised ILS first moment
//     m_k = ILS (x) [(nu - nu_c)(T_{k+1} - T_k)] / T_{k+1} (irs_channel_od_moments)
// are decoded from n_pc coefficients in the layer's own basis,
//     y_k(c) = mean_k(c) + inv_w_k(c) sum_m a_{k,m} V_{k,m}(c),
// and fed to the channel-space RTE with the CIM source of the line-by-line RTE
// (src/Solver/schwarzschild.jl:130-209) summed over the ILS, with Planck linear in
// nu across the ILS (B(nu) = B(nu_c) + B'(nu_c)(nu - nu_c)):
//     R = eps [B(T_sfc) T_0 + B'(T_sfc) M_s] + sum_k T_{k+1} [B(T_ave,k)(1 - t_k)
//         + (B(T_{k+1}) - B(T_ave,k)) g_k + B'(T_ave,k) m_k],   M_s = -sum_k m_k T_{k+1},
// T_L = 1, T_k = t_k T_{k+1}.  The transmittances are along the viewing path
// (the emulator is trained with 1/mu as an input).  For eps < 1 the reflected
// downwelling term (1 - eps) T_0 D_0 uses the layer transmittances,
// D_k = (1 - t_k) B(T_ave,k) + t_k D_{k+1} (layer-mean Planck emission, no moment
// term: an approximation; eps = 1 is exact up to the emulation).
#include "irs_devmath.h"
#include "irs_internal.h"
#include "irs_mlp.h"

#define EMU_MAXL 64
#define EMU_P 4            // profiles per workgroup of the forward kernel (basis loads shared)
#define EMU_MU_MAX 4.0     // |m_k| bound of the decode (cm^-1; half the ILS window is 16)

__device__ __forceinline__ double planck_dnu(double nu, double T) {   // dB/dnu at fixed T
    const double x = IRS_C2_PL * nu / T, em1 = expm1(x);
    const double B = IRS_C1_PL * nu * nu * nu / em1;
    return B * (3.0 / nu - (IRS_C2_PL / T) * (em1 + 1.0) / em1);
}

__device__ __forceinline__ double dbdt_c(double nu, double T) {
    const double x = IRS_C2_PL * nu / T, ex = exp(x);
    return IRS_C1_PL * IRS_C2_PL * nu * nu * nu * nu * ex / (T * T * (ex - 1.0) * (ex - 1.0));
}

__device__ __forceinline__ double d2b_dnu_dT(double nu, double T) {
    const double h = 1e-2;
    return (planck_dnu(nu, T + h) - planck_dnu(nu, T - h)) / (2.0 * h);
}

__device__ __forceinline__ double dbt_dr(double nu, double I) {
    if (!(I > 0.0)) return 0.0;
    const double A = IRS_C1_PL * nu * nu * nu, Lg = log1p(A / I);
    return IRS_C2_PL * nu * A / (Lg * Lg * I * (I + A));
}

// a[r][m] = scale y[r][m] + sum_j A[l][j][m] phi[r][j] for rows r0 .. r0 + rows - 1 (l = r mod L)
__global__ void __launch_bounds__(256) emu_coef_kernel(const float *y, const float *phi, const float *A, float scale,
                                                       int64_t r0, int64_t rows, int L, int n_phi, int n_pc,
                                                       float *coef) {
    const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= rows * n_pc) return;
    const int64_t r = id / n_pc;
    const int m = (int)(id % n_pc);
    const int l = (int)((r0 + r) % L);
    const float *ph = phi + (r0 + r) * n_phi;
    const float *Al = A + (int64_t)l * n_phi * n_pc + m;
    float acc = scale * y[id];
    for (int j = 0; j < n_phi; ++j) acc = fmaf(Al[(int64_t)j * n_pc], ph[j], acc);
    coef[(r0 + r) * n_pc + m] = acc;
}

hipError_t launch_emu_coef(const float *y, const float *phi, const float *A, float scale, int64_t r0, int64_t rows,
                           int L, int n_phi, int n_pc, float *coef, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(emu_coef_kernel, dim3((unsigned)((rows * n_pc + 255) / 256)), dim3(256), 0, s, y, phi, A, scale,
                       r0, rows, L, n_phi, n_pc, coef);
    return hipGetLastError();
}

// One thread per channel and EMU_P profiles per workgroup: each basis value is
// loaded once (coalesced) and used by the EMU_P profiles, whose coefficients
// [P][L][n_pc] sit in LDS (broadcast reads); the layers are decoded top-down
// and the RTE (FP64) advances with them.
template <int P>
__global__ void __launch_bounds__(256) emu_rte_kernel(const EmuRteArgs a) {
    extern __shared__ float s_coef[];          // [P][L][n_pc]
    const int L = a.L, npc = a.n_pc;
    const int q0 = blockIdx.y * P;
    const int np = min(P, a.n_prof - q0);
    for (int i = threadIdx.x; i < np * L * npc; i += 256) s_coef[i] = a.coef[(int64_t)q0 * L * npc + i];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t nch = a.n_ch, W = 3 * nch;
    const double nu = a.ch_nu[c];
    double top[P], acc[P], ms[P], dn[P];
#pragma unroll
    for (int p = 0; p < P; ++p) top[p] = 1.0, acc[p] = ms[p] = dn[p] = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        float yt[P], yg[P], ym[P];
#pragma unroll
        for (int p = 0; p < P; ++p) yt[p] = yg[p] = ym[p] = 0.0f;
        const float *bk = a.basis + (int64_t)k * npc * W + c;
        for (int m = 0; m < npc; ++m) {
            const float vt = bk[(int64_t)m * W], vg = bk[(int64_t)m * W + nch], vm = bk[(int64_t)m * W + 2 * nch];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const float cf = s_coef[(p * L + k) * npc + m];
                yt[p] = fmaf(cf, vt, yt[p]);
                yg[p] = fmaf(cf, vg, yg[p]);
                ym[p] = fmaf(cf, vm, ym[p]);
            }
        }
        const int64_t o = (int64_t)k * W + c;
        const float mt = a.mean[o], mg = a.mean[o + nch], mm = a.mean[o + 2 * nch];
        const float wt = a.inv_w[o], wg = a.inv_w[o + nch], wm = a.inv_w[o + 2 * nch];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (p < np) {
                const int q = q0 + p;
                const double x = fmin(fmax((double)fmaf(yt[p], wt, mt), 0.0), 1.0);
                const double g = fmin(fmax((double)fmaf(yg[p], wg, mg), 0.0), 1.0);
                const double mu = fmin(fmax((double)fmaf(ym[p], wm, mm), -EMU_MU_MAX), EMU_MU_MAX);
                const double Ta = a.T_ave[(int64_t)q * L + k], Tb = a.T_lev[(int64_t)q * (L + 1) + k + 1];
                const double ba = planck64(nu, Ta), bt = planck64(nu, Tb);
                const double bot = top[p] * x;
                acc[p] += top[p] * (ba * (1.0 - x) + (bt - ba) * g + planck_dnu(nu, Ta) * mu);
                ms[p] -= mu * top[p];
                dn[p] = (1.0 - x) * ba + x * dn[p];
                if (a.t_out) a.t_out[((int64_t)q * L + k) * nch + c] = (float)bot;
                top[p] = bot;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (p < np) {
            const int q = q0 + p;
            const double T_sfc = a.scal[3 * q], eps = a.scal[3 * q + 2];
            double I = eps * (planck64(nu, T_sfc) * top[p] + planck_dnu(nu, T_sfc) * ms[p]) + acc[p];
            if (eps < 1.0) I += (1.0 - eps) * dn[p] * top[p];
            const int64_t oo = (int64_t)q * nch + c;
            a.R[oo] = I;
            a.BT[oo] = I <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / I);
        }
    }
}

hipError_t launch_emu_rte(const EmuRteArgs &a, hipStream_t s) {
    if (a.L < 1 || a.L > EMU_MAXL || a.n_prof < 1) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)EMU_P * a.L * a.n_pc;
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(emu_rte_kernel<EMU_P>, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)((a.n_prof + EMU_P - 1) / EMU_P)),
                       dim3(256), lds, s, a);
    return hipGetLastError();
}

// Spot-check correction of emulated spectra: with the scene's errors at the n_spot
// spot channels e_q = BT_emulated[S] - BT_LBL[S], every channel moves by the linear
// map fitted on calibration scenes, BT[q][c] -= sum_s e_q[s] W[s][c], and R is the
// Planck radiance of the corrected BT.  One thread per (channel, scene); the
// scene's errors sit in LDS (broadcast reads), W is read along c (coalesced).
__global__ void __launch_bounds__(256) emu_correct_kernel(const float *e, const float *W, int n_spot, int64_t n_ch,
                                                          const double *ch_nu, double *R, double *BT) {
    __shared__ float s_e[EMU_MAX_SPOT];
    const int q = blockIdx.y;
    for (int i = threadIdx.x; i < n_spot; i += 256) s_e[i] = e[(int64_t)q * n_spot + i];
    __syncthreads();
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n_ch) return;
    float acc = 0.0f;
    for (int k = 0; k < n_spot; ++k) acc = fmaf(s_e[k], W[(int64_t)k * n_ch + c], acc);
    const int64_t o = (int64_t)q * n_ch + c;
    const double bt = BT[o] - (double)acc;
    BT[o] = bt;
    R[o] = bt > 0.0 ? planck64(ch_nu[c], bt) : 0.0;
}

hipError_t launch_emu_correct(const float *e, const float *W, int n_spot, int64_t n_ch, int n_prof,
                              const double *ch_nu, double *R, double *BT, hipStream_t s) {
    if (n_spot < 1 || n_spot > EMU_MAX_SPOT || n_prof < 1 || n_ch < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(emu_correct_kernel, dim3((unsigned)((n_ch + 255) / 256), (unsigned)n_prof), dim3(256), 0, s, e,
                       W, n_spot, n_ch, ch_nu, R, BT);
    return hipGetLastError();
}

// Adjoint of the emulated channel RTE (one profile per workgroup row).  With
// f_k = B_k (1 - t_k) + (bt_k - B_k) g_k + (B'_k - eps B'_s) m_k the radiance is
// R = eps B_s T_0 + sum_k T_{k+1} f_k, T_k = t_k T_{k+1}, and Lambda_k = dR/dT_k
// (T_k free, the levels below following it) obeys
//     Lambda_0 = eps B_s,   Lambda_k = f_{k-1} + t_{k-1} Lambda_{k-1},
// so dR/dt_k = T_{k+1} (Lambda_k - B_k), dR/dg_k = T_{k+1} (bt_k - B_k),
// dR/dm_k = T_{k+1} (B'_k - eps B'_s).  Every output is multiplied by dBT/dR;
// clamped targets get 0; the reflected term (eps < 1) is held fixed.  The
// forward sweep parks (t_k, g_k, m_k, T_{k+1}) in the output arrays, the
// reverse sweep reads them back and overwrites them with the derivatives.
__global__ void __launch_bounds__(256) emu_adjoint_kernel(const EmuRteArgs a, EmuAdjArgs j) {
    extern __shared__ float s_coef[];          // [L][n_pc]
    const int q = blockIdx.y, L = a.L, npc = a.n_pc;
    const float *cq = a.coef + (int64_t)q * L * npc;
    for (int i = threadIdx.x; i < L * npc; i += 256) s_coef[i] = cq[i];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t nch = a.n_ch, W = 3 * nch;
    const double nu = a.ch_nu[c];
    const double *T_lev = a.T_lev + (int64_t)q * (L + 1), *T_ave = a.T_ave + (int64_t)q * L;
    const double T_sfc = a.scal[3 * q], eps = a.scal[3 * q + 2];
    double top = 1.0, acc = 0.0, ms = 0.0, dn = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        float yt = 0.0f, yg = 0.0f, ym = 0.0f;
        const float *bk = a.basis + (int64_t)k * npc * W + c;
        for (int m = 0; m < npc; ++m) {
            const float cf = s_coef[k * npc + m];
            yt = fmaf(cf, bk[(int64_t)m * W], yt);
            yg = fmaf(cf, bk[(int64_t)m * W + nch], yg);
            ym = fmaf(cf, bk[(int64_t)m * W + 2 * nch], ym);
        }
        const int64_t o = (int64_t)k * W + c;
        const float xr = fmaf(yt, a.inv_w[o], a.mean[o]);
        const float gr = fmaf(yg, a.inv_w[o + nch], a.mean[o + nch]);
        const float mr = fmaf(ym, a.inv_w[o + 2 * nch], a.mean[o + 2 * nch]);
        const double x = fmin(fmax((double)xr, 0.0), 1.0), g = fmin(fmax((double)gr, 0.0), 1.0);
        const double mu = fmin(fmax((double)mr, -EMU_MU_MAX), EMU_MU_MAX);
        const double ba = planck64(nu, T_ave[k]), bt = planck64(nu, T_lev[k + 1]);
        acc += top * (ba * (1.0 - x) + (bt - ba) * g + planck_dnu(nu, T_ave[k]) * mu);
        ms -= mu * top;
        dn = (1.0 - x) * ba + x * dn;
        const int64_t ok = ((int64_t)q * L + k) * nch + c;
        j.dt[ok] = (xr > 0.0f && xr < 1.0f) ? xr : (xr <= 0.0f ? -0.0f : 2.0f);   // value and clamp state
        j.dg[ok] = (gr > 0.0f && gr < 1.0f) ? gr : (gr <= 0.0f ? -0.0f : 2.0f);
        j.dm[ok] = mr;
        j.dTave[ok] = (float)top;                                                  // T_{k+1}
        top *= x;
    }
    const double bs = planck64(nu, T_sfc), bps = planck_dnu(nu, T_sfc);
    double I = eps * (bs * top + bps * ms) + acc;
    if (eps < 1.0) I += (1.0 - eps) * dn * top;
    const double dbt = dbt_dr(nu, I);
    const int64_t oq = (int64_t)q * nch + c;
    j.BT[oq] = I <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / I);
    j.dTsfc[oq] = (float)(dbt * eps * (top * dbdt_c(nu, T_sfc) + ms * d2b_dnu_dT(nu, T_sfc)));
    double lam = eps * bs;
    for (int k = 0; k < L; ++k) {
        const int64_t ok = ((int64_t)q * L + k) * nch + c;
        const float xs = j.dt[ok], gs = j.dg[ok], mr = j.dm[ok];
        const double up = (double)j.dTave[ok];
        const double x = fmin(fmax((double)xs, 0.0), 1.0), g = fmin(fmax((double)gs, 0.0), 1.0);
        const double mu = fmin(fmax((double)mr, -EMU_MU_MAX), EMU_MU_MAX);
        const double ba = planck64(nu, T_ave[k]), bt = planck64(nu, T_lev[k + 1]);
        const double bpa = planck_dnu(nu, T_ave[k]);
        j.dt[ok] = (float)((xs > 0.0f && xs < 1.0f) ? dbt * up * (lam - ba) : 0.0);
        j.dg[ok] = (float)((gs > 0.0f && gs < 1.0f) ? dbt * up * (bt - ba) : 0.0);
        j.dm[ok] = (float)((fabs((double)mr) < EMU_MU_MAX) ? dbt * up * (bpa - eps * bps) : 0.0);
        j.dTave[ok] = (float)(dbt * up * ((1.0 - x - g) * dbdt_c(nu, T_ave[k]) + mu * d2b_dnu_dT(nu, T_ave[k])));
        j.dTlev[((int64_t)q * (L + 1) + k + 1) * nch + c] = (float)(dbt * up * g * dbdt_c(nu, T_lev[k + 1]));
        lam = ba * (1.0 - x) + (bt - ba) * g + (bpa - eps * bps) * mu + x * lam;
    }
    j.dTlev[((int64_t)q * (L + 1)) * nch + c] = 0.0f;
}

// dBT/dfeature: F[q][k][i][c] = sum_m J[q][k][m][i] (dt V^t_{k,m}(c) w^t + dg V^g_{k,m}(c) w^g
// + dm V^m_{k,m}(c) w^m) with J = d a_{k,m} / d feature_i (the host folds the MLP
// input Jacobian, the regression and the standardisation into it) and w the decode
// weights.  One workgroup per (channel block, layer, profile); J of the layer in LDS.
#define EMU_MAXF 24
__global__ void __launch_bounds__(256) emu_feature_jac_kernel(const EmuRteArgs a, EmuAdjArgs j, int n_feat) {
    extern __shared__ float s_j[];             // [n_pc][n_feat]
    const int k = blockIdx.y, q = blockIdx.z, L = a.L, npc = a.n_pc;
    const float *Jk = j.Jcoef + (((int64_t)q * L + k) * npc) * n_feat;
    for (int i = threadIdx.x; i < npc * n_feat; i += 256) s_j[i] = Jk[i];
    __syncthreads();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_ch) return;
    const int64_t nch = a.n_ch, W = 3 * nch;
    const int64_t ok = ((int64_t)q * L + k) * nch + c, o = (int64_t)k * W + c;
    const float at = j.dt[ok] * a.inv_w[o], ag = j.dg[ok] * a.inv_w[o + nch], am = j.dm[ok] * a.inv_w[o + 2 * nch];
    float acc[EMU_MAXF];
#pragma unroll
    for (int i = 0; i < EMU_MAXF; ++i) acc[i] = 0.0f;
    const float *bk = a.basis + (int64_t)k * npc * W + c;
    for (int m = 0; m < npc; ++m) {
        const float w = at * bk[(int64_t)m * W] + ag * bk[(int64_t)m * W + nch] + am * bk[(int64_t)m * W + 2 * nch];
#pragma unroll
        for (int i = 0; i < EMU_MAXF; ++i)
            if (i < n_feat) acc[i] = fmaf(s_j[m * n_feat + i], w, acc[i]);
    }
#pragma unroll
    for (int i = 0; i < EMU_MAXF; ++i)
        if (i < n_feat) j.dfeat[(((int64_t)q * L + k) * n_feat + i) * nch + c] = acc[i];
}

hipError_t launch_emu_adjoint(const EmuRteArgs &a, const EmuAdjArgs &j, int n_feat, hipStream_t s) {
    if (a.L < 1 || a.L > EMU_MAXL || n_feat < 1 || n_feat > EMU_MAXF) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)a.L * a.n_pc;
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(emu_adjoint_kernel, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)a.n_prof), dim3(256), lds,
                       s, a, j);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(emu_feature_jac_kernel, dim3((unsigned)((a.n_ch + 255) / 256), (unsigned)a.L, (unsigned)a.n_prof),
                       dim3(256), sizeof(float) * (size_t)a.n_pc * n_feat, s, a, j, n_feat);
    return hipGetLastError();
}
