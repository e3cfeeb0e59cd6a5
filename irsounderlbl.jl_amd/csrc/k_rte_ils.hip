// Schwarzschild RTE scan and ILS channel reduction (FP64).
//
// rte_kernel: one thread per internal grid point.  Layers are swept from the
// top of the atmosphere down, forming the level-to-space transmittance as the
// running product T_k = T_{k+1} exp(-tau_k/mu) (src/Solver/transmittance.jl:
// 26-43) and adding each layer's emission with the CIM or Toon source
// (src/Solver/schwarzschild.jl:161-193), then the surface term and, for
// eps < 1, the reflected downwelling term (schwarzschild.jl:203-206,228-274).
// tau is read once (FP32 from the line sum, or FP64 for irs_rte).
//
// ils_channels_kernel: one workgroup per sounder channel: the direct ILS
// correlation of src/Sensor/ils.jl:131-153 evaluated only at the channel's
// internal node(s), linear interpolation between two nodes when the channel
// is not a node (src/Sensor/sounder.jl:372-377), then the brightness
// temperature (src/Solver/planck.jl:43-46).
#include "irs_devmath.h"
#include "irs_internal.h"

__device__ __forceinline__ double cim_corr(double s, double e) {   // schwarzschild.jl:59-64
    return s < 0.06 ? s / 6.0 : 1.0 - 2.0 * (e / (e - 1.0) + 1.0 / s);
}

__device__ __forceinline__ double lit_corr(double s, double e) {   // schwarzschild.jl:43-47
    return s < 1e-4 ? s * (0.5 - s * (1.0 / 3.0 - s * 0.125)) : (1.0 - e) / s - e;
}

template <typename TT>
__global__ void __launch_bounds__(256) rte_kernel(const RteArgs a, const TT *__restrict__ tau_all) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int q = blockIdx.y;
    const int L = a.n_layers;
    const TT *__restrict__ tau = tau_all + (int64_t)q * L * a.n;
    const double *T_lev = a.T_lev + (int64_t)q * (L + 1);
    const double *T_ave = a.T_ave ? a.T_ave + (int64_t)q * L : nullptr;
    const double T_sfc = a.scal ? a.scal[3 * q] : a.T_sfc;
    const double mu = a.scal ? a.scal[3 * q + 1] : a.mu;
    const double eps = a.scal ? a.scal[3 * q + 2] : a.eps;
    const double nu = (a.an + (double)i * a.sn) / a.den;
    double tr_top = 1.0;
    double acc = 0.0;
    for (int k = L - 1; k >= 0; --k) {
        const double s = (double)tau[(int64_t)k * a.n + i] / mu;
        const double e = exp(-s);
        const double tr_bot = tr_top * e;
        const double dT = tr_top - tr_bot;
        const double bt = planck64(nu, T_lev[k + 1]);
        if (a.source_function == 1) {
            const double ba = planck64(nu, T_ave[k]);
            acc += ba * dT + (bt - ba) * (dT * cim_corr(s, e));
        } else {
            const double bb = planck64(nu, T_lev[k]);
            acc += bt * dT + (bb - bt) * (tr_top * lit_corr(s, e));
        }
        tr_top = tr_bot;
    }
    double I = eps * planck64(nu, T_sfc) * tr_top + acc;
    if (eps < 1.0) {
        double down = 0.0, g = 1.0;
        for (int k = 0; k < L; ++k) {
            const double s = (double)tau[(int64_t)k * a.n + i] / mu;
            const double e = exp(-s);
            double em;
            if (a.source_function == 1) {
                const double ba = planck64(nu, T_ave[k]), bb = planck64(nu, T_lev[k]);
                em = (1.0 - e) * (ba + (bb - ba) * cim_corr(s, e));
            } else {
                const double bb = planck64(nu, T_lev[k]), bt = planck64(nu, T_lev[k + 1]);
                em = bb * (1.0 - e) + (bt - bb) * lit_corr(s, e);
            }
            down += g * em;
            g *= e;
        }
        I += (1.0 - eps) * down * tr_top;
    }
    a.out[(int64_t)q * a.n + i] = I;
}

hipError_t launch_rte(const RteArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    dim3 grid((unsigned)((a.n + 255) / 256), (unsigned)(a.n_prof > 0 ? a.n_prof : 1));
    if (a.tau32) hipLaunchKernelGGL(rte_kernel<float>, grid, dim3(256), 0, s, a, a.tau32);
    else hipLaunchKernelGGL(rte_kernel<double>, grid, dim3(256), 0, s, a, a.tau64);
    return hipGetLastError();
}

// ILS taps: 2 * sum_m A(L_m) cos(2 pi L_m dnu_k) dL   (ils.jl:104-108), unnormalised.
__global__ void __launch_bounds__(256) ils_taps_kernel(const double *__restrict__ A,
                                                       const double *__restrict__ Lm, int n_pts,
                                                       double dL, double dnu, int64_t half,
                                                       double *__restrict__ taps) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k > 2 * half) return;
    const double dv = (double)(k - half) * dnu;
    double s = 0.0;
    for (int m = 0; m < n_pts; ++m) s += A[m] * cos(2.0 * IRS_PI * Lm[m] * dv);
    taps[k] = 2.0 * s * dL;
}

hipError_t launch_ils_taps(const double *A, const double *L, int n_pts, double dL, double dnu,
                           int64_t half, double *taps, hipStream_t s) {
    const int64_t n = 2 * half + 1;
    hipLaunchKernelGGL(ils_taps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, A, L, n_pts,
                       dL, dnu, half, taps);
    return hipGetLastError();
}

__device__ __forceinline__ double block_sum_d(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double ils_at(const IlsArgs &a, const double *__restrict__ spec, int64_t i,
                                         double *red) {
    double acc = 0.0;
    const int64_t nt = 2 * a.half + 1;
    for (int64_t k = threadIdx.x; k < nt; k += 256) {
        const int64_t j = i + k - a.half;
        if (j >= 0 && j < a.n) acc += spec[j] * a.taps[k];
    }
    return block_sum_d(acc, red) * a.dnu;
}

__global__ void __launch_bounds__(256) ils_channels_kernel(const IlsArgs a) {
    __shared__ double red[4];
    const int64_t c = blockIdx.x;
    const int q = blockIdx.y;
    const double *spec = a.spec + (int64_t)q * a.n;
    const int64_t i = a.ch_idx[c];
    const double w = a.ch_w[c];
    double r = ils_at(a, spec, i, red);
    if (w != 0.0) {
        const double r2 = ils_at(a, spec, i + 1, red);
        r = r * (1.0 - w) + r2 * w;
    }
    if (threadIdx.x == 0) {
        const int64_t o = (int64_t)q * a.n_ch + c;
        a.r_out[o] = r;
        if (a.bt_out) {
            const double nu = a.ch_nu[c];
            a.bt_out[o] = r <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / r);
        }
    }
}

hipError_t launch_ils_channels(const IlsArgs &a, hipStream_t s) {
    if (a.n_ch <= 0) return hipSuccess;
    hipLaunchKernelGGL(ils_channels_kernel, dim3((unsigned)a.n_ch, (unsigned)(a.n_prof > 0 ? a.n_prof : 1)),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}
