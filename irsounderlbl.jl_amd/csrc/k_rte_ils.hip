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
#include <cstdlib>

#include "irs_devmath.h"
#include "irs_internal.h"

__device__ __forceinline__ double cim_corr(double s, double e) {   // schwarzschild.jl:59-64
    return s < 0.06 ? s / 6.0 : 1.0 - 2.0 * (e / (e - 1.0) + 1.0 / s);
}

__device__ __forceinline__ double lit_corr(double s, double e) {   // schwarzschild.jl:43-47
    return s < 1e-4 ? s * (0.5 - s * (1.0 / 3.0 - s * 0.125)) : (1.0 - e) / s - e;
}

// rte_kernel: each thread owns P grid points i0 + t + 256 p of a 256 P-point
// block (coalesced loads of tau per layer).  Planck at the P points comes from
// one FP64 exponential per temperature and thread: exp(C2 nu_p / T) =
// exp(C2 nu_0 / T) rho^p with rho = exp(256 C2 dnu / T) from LDS (relative
// rounding below 1e-15 after P - 1 products), and 1 / (exp - 1) by the
// hardware reciprocal refined by two Newton steps.  The CIM correction times
// dT = T_{k+1} (1 - e) is dT + 2 T_{k+1} e - 2 dT / s for s >= 0.06 (the same
// f_cim, with e/(e - 1) dT = -T_{k+1} e), so one reciprocal per point remains.
// What is left per (point, layer) is the FP64 exp(-tau/mu) of the
// transmittance, which bounds the kernel above the HBM time of reading tau.
#define RTE_P 4
#define RTE_MAXL 255

__device__ __forceinline__ double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// exp(-s) for s >= 0 in FP64 (the layer transmittance): exp(-s) = 2^-(k/64) exp(-r),
// k = rint(64 s / ln 2), r = s - k ln2/64 in [-ln2/128, ln2/128] (Cody-Waite split of
// ln2/64), exp(-r) by its Taylor polynomial of degree 7 (first omitted term
// r^8/8! < 2e-23), 2^-(k/64) = 2^-(k >> 6) tab[k & 63] from a 64-entry table.
// Relative error a few ulp; zero below exp(-708).
__device__ __forceinline__ double exp_neg(double s, const double *tab) {
    if (!(s < 708.0)) return 0.0;
    const double kd = rint(s * 92.332482616893656);              // 64 / ln 2
    const int k = (int)kd;
    double r = fma(-kd, 1.0830424695086549e-02, s);               // ln2/64: 32-bit high part (k hi exact)
    r = fma(-kd, 1.162596423439437e-12, r);                       //         low part
    const double m = -r;
    double p = fma(m, 1.0 / 5040.0, 1.0 / 720.0);
    p = fma(p, m, 1.0 / 120.0);
    p = fma(p, m, 1.0 / 24.0);
    p = fma(p, m, 1.0 / 6.0);
    p = fma(p, m, 0.5);
    p = fma(p, m, 1.0);
    p = fma(p, m, 1.0);
    return ldexp(tab[k & 63] * p, -(k >> 6));
}

template <typename TT>
__global__ void __launch_bounds__(256) rte_kernel(const RteArgs a, const TT *__restrict__ tau_all) {
    __shared__ double s_xl[RTE_MAXL + 1], s_rl[RTE_MAXL + 1], s_xa[RTE_MAXL], s_ra[RTE_MAXL];
    __shared__ double s_e2[64];
    if (threadIdx.x < 64) s_e2[threadIdx.x] = exp2(-(double)threadIdx.x / 64.0);
    const int q = blockIdx.y;
    const int L = a.n_layers;
    const double *T_lev = a.T_lev + (int64_t)q * (L + 1);
    const double *T_ave = a.T_ave ? a.T_ave + (int64_t)q * L : nullptr;
    const double step = 256.0 * a.sn / a.den;
    for (int j = threadIdx.x; j <= L; j += 256) {
        const double x = IRS_C2_PL / T_lev[j];
        s_xl[j] = x;
        s_rl[j] = exp(x * step);
        if (j < L && T_ave) {
            const double xa = IRS_C2_PL / T_ave[j];
            s_xa[j] = xa;
            s_ra[j] = exp(xa * step);
        }
    }
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * (256 * RTE_P) + threadIdx.x;
    if (i0 >= a.n) return;
    const TT *__restrict__ tau = tau_all + (int64_t)q * L * a.n;
    const double T_sfc = a.scal ? a.scal[3 * q] : a.T_sfc;
    const double mu = a.scal ? a.scal[3 * q + 1] : a.mu;
    const double eps = a.scal ? a.scal[3 * q + 2] : a.eps;
    const double imu = 1.0 / mu;
    const int np = (int)min((int64_t)RTE_P, (a.n - i0 + 255) / 256);   // points of this thread in the grid
    double c1nu3[RTE_P], tt[RTE_P], acc[RTE_P];
    const double nu0 = (a.an + (double)i0 * a.sn) / a.den;
#pragma unroll
    for (int p = 0; p < RTE_P; ++p) {
        const double nu = (a.an + (double)(i0 + 256 * p) * a.sn) / a.den;
        c1nu3[p] = IRS_C1_PL * nu * nu * nu;
        tt[p] = 1.0;
        acc[p] = 0.0;
    }
    const bool cim = a.source_function == 1;
    for (int k = L - 1; k >= 0; --k) {
        const TT *tk = tau + (int64_t)k * a.n + i0;
        double tv[RTE_P];
#pragma unroll
        for (int p = 0; p < RTE_P; ++p) tv[p] = p < np ? (double)tk[256 * p] : 0.0;
        // Planck at T_top = T_lev[k+1] and T_other = T_ave[k] (CIM) or T_lev[k] (Toon)
        const double xo = cim ? s_xa[k] : s_xl[k], ro = cim ? s_ra[k] : s_rl[k];
        double Et = exp(nu0 * s_xl[k + 1]), Eo = exp(nu0 * xo);
        const double rt = s_rl[k + 1];
#pragma unroll
        for (int p = 0; p < RTE_P; ++p) {
            const double sv = tv[p] * imu;
            // the three reciprocals 1/(Et - 1), 1/(Eo - 1), 1/s from one: 1/((Et - 1)(Eo - 1) s')
            const double At = Et - 1.0, Ao = Eo - 1.0, s1 = sv < 0.06 ? 1.0 : sv;
            const double AtAo = At * Ao;
            const double inv = rcp64(AtAo * s1);
            const double Bt = c1nu3[p] * (Ao * s1 * inv), Bo = c1nu3[p] * (At * s1 * inv);
            Et *= rt;
            Eo *= ro;
            const double e = exp_neg(sv, s_e2);
            const double tb = tt[p] * e;
            const double dT = tt[p] - tb;
            if (cim) {
                // C_k = dT f_cim(s): s/6 below 0.06 (schwarzschild.jl:59-64)
                const double Ck = sv < 0.06 ? dT * (sv * (1.0 / 6.0)) : dT + 2.0 * tb - 2.0 * dT * (AtAo * inv);
                acc[p] += Bo * dT + (Bt - Bo) * Ck;
            } else {
                acc[p] += Bt * dT + (Bo - Bt) * (tt[p] * lit_corr(sv, e));
            }
            tt[p] = tb;
        }
    }
    double I[RTE_P];
    {
        const double xs = IRS_C2_PL / T_sfc, rs = exp(xs * step);
        double Es = exp(nu0 * xs);
#pragma unroll
        for (int p = 0; p < RTE_P; ++p) {
            I[p] = eps * (c1nu3[p] * rcp64(Es - 1.0)) * tt[p] + acc[p];
            Es *= rs;
        }
    }
    if (eps < 1.0) {
        // reflected downwelling (schwarzschild.jl:203-206,228-274), from the top down to the surface
        double down[RTE_P], g[RTE_P];
#pragma unroll
        for (int p = 0; p < RTE_P; ++p) { down[p] = 0.0; g[p] = 1.0; }
        for (int k = 0; k < L; ++k) {
            const TT *tk = tau + (int64_t)k * a.n + i0;
            double tv[RTE_P];
#pragma unroll
            for (int p = 0; p < RTE_P; ++p) tv[p] = p < np ? (double)tk[256 * p] : 0.0;
            // B_bot = B(T_lev[k]); B_x = B(T_ave[k]) (CIM) or B(T_lev[k+1]) (Toon)
            const double xx = cim ? s_xa[k] : s_xl[k + 1], rx = cim ? s_ra[k] : s_rl[k + 1];
            double Eb = exp(nu0 * s_xl[k]), Ex = exp(nu0 * xx);
            const double rb = s_rl[k];
#pragma unroll
            for (int p = 0; p < RTE_P; ++p) {
                const double Bb = c1nu3[p] * rcp64(Eb - 1.0), Bx = c1nu3[p] * rcp64(Ex - 1.0);
                Eb *= rb;
                Ex *= rx;
                const double sv = tv[p] * imu;
                const double e = exp_neg(sv, s_e2);
                double em;
                if (cim) em = (1.0 - e) * (Bx + (Bb - Bx) * cim_corr(sv, e));
                else em = Bb * (1.0 - e) + (Bx - Bb) * lit_corr(sv, e);
                down[p] += g[p] * em;
                g[p] *= e;
            }
        }
#pragma unroll
        for (int p = 0; p < RTE_P; ++p) I[p] += (1.0 - eps) * down[p] * tt[p];
    }
    double *out = a.out + (int64_t)q * a.n + i0;
#pragma unroll
    for (int p = 0; p < RTE_P; ++p)
        if (p < np) out[256 * p] = I[p];
}

hipError_t launch_rte(const RteArgs &a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    if (a.n_layers > RTE_MAXL) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.n + 256 * RTE_P - 1) / (256 * RTE_P)), (unsigned)(a.n_prof > 0 ? a.n_prof : 1));
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

// Channel ILS on equally spaced node channels (idx_c = idx_0 + c S, exact
// nodes): a workgroup evaluates ILS_G consecutive channels at once.  With
// k = r + S m (r the residue of the tap index, m its row), channel c0 + g reads
//   spec[idx_0 - half + r + S (c0 + g + m)],
// so for one residue the G channels slide along one strided sequence of the
// spectrum: out_g += sum_m spec(r, m + g) taps(r, m).  A work item is one
// residue and ILS_MC consecutive rows; it keeps a circular window of G spectrum
// values in registers (one new load and G FMAs per row, the tap shared by all G),
// loads are coalesced across the wave (consecutive residues), and the partial
// sums of the block are reduced per channel.  Workgroups map to channel groups
// so that each XCD (workgroups dispatched round-robin over 8) covers one
// contiguous eighth of the spectrum in its L2.
#define ILS_MC 32
template <int ILS_G>
__global__ void __launch_bounds__(256) ils_block_kernel(const IlsBlockArgs a) {
    __shared__ double red[4][ILS_G];
    const int n_groups = (int)((a.n_ch + ILS_G - 1) / ILS_G);
    const int per_xcd = (n_groups + 7) / 8;
    const int b = blockIdx.x;
    const int grp = (b % 8) * per_xcd + b / 8;                 // XCD-contiguous group order
    if (grp >= n_groups) return;                              // block-uniform
    const int col = blockIdx.y;
    const double *spec = a.spec + (int64_t)col * a.n;
    const int64_t c0 = (int64_t)grp * ILS_G;
    const int S = a.stride;
    const int64_t nt = 2 * a.half + 1;
    const int64_t base = a.idx0 - a.half + (int64_t)c0 * S;   // index of (r = 0, n = 0)
    const int64_t rows = (nt + S - 1) / S;                      // rows of residue 0
    const int64_t n_mc = (rows + ILS_MC - 1) / ILS_MC;
    const int64_t items = (int64_t)S * n_mc;
    double acc[ILS_G];
#pragma unroll
    for (int g = 0; g < ILS_G; ++g) acc[g] = 0.0;
    for (int64_t w = threadIdx.x; w < items; w += 256) {
        const int r = (int)(w % S);
        const int64_t m0 = (w / S) * ILS_MC;
        const int64_t mr = (nt - r + S - 1) / S;                  // rows of residue r
        if (m0 >= mr) continue;
        const int64_t m1 = min(m0 + (int64_t)ILS_MC, mr);
        const int64_t jb = base + r;                              // spec index of (m, g): jb + S (m + g)
        const double *tp = a.taps + r + (int64_t)S * m0;
        double win[ILS_G];
        if (jb + (int64_t)S * m0 >= 0 && jb + (int64_t)S * (m1 + ILS_G - 1) < a.n) {
            // interior: every window value is in the grid, pointers advance by S
            const double *sp = spec + jb + (int64_t)S * m0;
#pragma unroll
            for (int g = 0; g < ILS_G; ++g) win[g] = sp[(int64_t)S * g];
            sp += (int64_t)S * ILS_G;
            for (int64_t m = m0; m < m1; m += ILS_G) {
#pragma unroll
                for (int u = 0; u < ILS_G; ++u) {
                    if (m + u < m1) {
                        const double t = *tp;
                        tp += S;
#pragma unroll
                        for (int g = 0; g < ILS_G; ++g) acc[g] = fma(win[(u + g) % ILS_G], t, acc[g]);
                        win[u] = *sp;                                // row m + u + G (may lie past m1: unused)
                        sp += S;
                    }
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < ILS_G; ++g) {
                const int64_t j = jb + (int64_t)S * (m0 + g);
                win[g] = (j >= 0 && j < a.n) ? spec[j] : 0.0;
            }
            // rows in blocks of ILS_G with a static rotation of the window
            for (int64_t m = m0; m < m1; m += ILS_G) {
#pragma unroll
                for (int u = 0; u < ILS_G; ++u) {
                    const int64_t mm = m + u;
                    if (mm < m1) {
                        const double t = a.taps[r + (int64_t)S * mm];
#pragma unroll
                        for (int g = 0; g < ILS_G; ++g) acc[g] = fma(win[(u + g) % ILS_G], t, acc[g]);
                        const int64_t j = jb + (int64_t)S * (mm + ILS_G);
                        win[u] = (j >= 0 && j < a.n) ? spec[j] : 0.0;    // row mm + G, window slot u
                    }
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int g = 0; g < ILS_G; ++g) {
        double v = acc[g];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wv][g] = v;
    }
    __syncthreads();
    if (threadIdx.x < ILS_G) {
        const int g = threadIdx.x;
        const int64_t c = c0 + g;
        if (c < a.n_ch) {
            const double r = (red[0][g] + red[1][g] + red[2][g] + red[3][g]) * a.dnu;
            const int64_t o = (int64_t)col * a.n_ch + c;
            a.r_out[o] = a.factor ? r * a.factor[c] : r;
            if (a.bt_out) {
                const double nu = a.ch_nu[c];
                a.bt_out[o] = r <= 0.0 ? 0.0 : IRS_C2_PL * nu / log1p(IRS_C1_PL * nu * nu * nu / r);
            }
        }
    }
}

// G = 8 channels per workgroup keeps the kernel at 118 VGPRs (four waves per SIMD)
hipError_t launch_ils_block(const IlsBlockArgs &a, int n_cols, hipStream_t s) {
    if (a.n_ch <= 0) return hipSuccess;
    if (a.stride < 1) return hipErrorInvalidValue;
    static const int G = [] {
        const char *e = std::getenv("IRS_ILS_G");
        return (e && std::atoi(e) == 16) ? 16 : 8;
    }();
    const int n_groups = (int)((a.n_ch + G - 1) / G);
    const int blocks = ((n_groups + 7) / 8) * 8;
    const dim3 grid((unsigned)blocks, (unsigned)std::max(n_cols, 1));
    if (G == 16) hipLaunchKernelGGL(ils_block_kernel<16>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(ils_block_kernel<8>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ils_channels(const IlsArgs &a, hipStream_t s) {
    if (a.n_ch <= 0) return hipSuccess;
    hipLaunchKernelGGL(ils_channels_kernel, dim3((unsigned)a.n_ch, (unsigned)(a.n_prof > 0 ? a.n_prof : 1)),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}
