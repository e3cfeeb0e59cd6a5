// Jacobian kernels: per-line derivative parameters (FP64), the FP32 gradient
// line sum, the FP64 RTE Jacobian scan, column assembly and the multi-column
// channel ILS.
//
// Restates compute_voigt_cross_sections_grad / voigt_grad_kernel!
// (src/CrossSections/cross_section_jacobian.jl:43-240), the LM-aware gradient
// (:255-322, here with the analytic derivative of the dispersive term instead
// of central differences), schwarzschild_rte_jacobian and its reflected term
// (src/Solver/schwarzschild_jacobian.jl:66-266), and the column assembly of
// analytic_jacobian (src/Estimation/vmr_jacobian.jl:383-504).
#include "irs_devmath.h"
#include "irs_internal.h"
#include "irs_jac.h"

#define TIPS_T_MIN 70
#define TIPS_NT 331
#define T_REF 296.0
#define C2_LINE 1.4387752
#define C2_LM 1.4388

__device__ __forceinline__ double tips_at(const double *v, double T) {
    long lo = (long)floor(T) - TIPS_T_MIN, hi = lo + 1;
    lo = lo < 0 ? 0 : (lo > TIPS_NT - 1 ? TIPS_NT - 1 : lo);
    hi = hi < 0 ? 0 : (hi > TIPS_NT - 1 ? TIPS_NT - 1 : hi);
    const double fr = T - floor(T);
    return v[lo] * (1.0 - fr) + v[hi] * fr;
}

__device__ __forceinline__ double tips_slope(const double *v, double T) {   // partition.jl:59-66
    const long lo = (long)floor(T) - TIPS_T_MIN, hi = lo + 1;
    if (lo < 0 || hi > TIPS_NT - 1) return 0.0;
    return v[hi] - v[lo];
}

__device__ __forceinline__ double molmass(int mol) {
    const double m[11] = {18.010, 44.010, 47.998, 44.013, 28.010, 16.043, 31.999, 34.015, 64.065, 36.461, 17.031};
    return (mol < 1 || mol > 11) ? 30.0 : m[mol - 1];
}

__device__ __forceinline__ double lmmass(int mol, int iso) {
    const double co2[11] = {47.00427, 43.98983, 44.99318, 45.99398, 44.99403, 45.99706,
                            44.99706, 47.99832, 47.00134, 48.00196, 47.00427};
    const double ch4[5] = {44.0, 16.031300, 17.034655, 17.037475, 18.040830};
    if (mol == 2 && iso >= 0 && iso <= 10) return co2[iso];
    if (mol == 6 && iso >= 1 && iso <= 4) return ch4[iso];
    return 44.0;
}

// (H, dH/dx, dH/dy) with the far-wing branch (cross_section_jacobian.jl:43-56)
__device__ __forceinline__ void hgrad64(double x, double y, double xfar, double &H, double &Hx, double &Hy) {
    if (fabs(x) > xfar) {
        const double d = x * x + y * y, d2 = d * d;
        H = y * IRS_INV_SQRT_PI / d;
        Hx = -2.0 * x * y * IRS_INV_SQRT_PI / d2;
        Hy = IRS_INV_SQRT_PI * (x * x - y * y) / d2;
        return;
    }
    double K, Lw;
    w64(x, y, K, Lw);
    H = K;
    Hx = -2.0 * x * K + 2.0 * y * Lw;
    Hy = 2.0 * (x * Lw + y * K) - 2.0 * IRS_INV_SQRT_PI;
}

// One thread per (item, line): six float4 records, coef folded into amplitudes.
//   g0 = {cn, cf, fu, y}        g1 = {Sn, dSn/dT, dfu/dT, dy/dT}
//   g2 = {dy/dp, dy/dvs, dnu0/dp / dnu, dnu0/dvs / dnu}
//   g3 = {Hc, dHc/dT, dHc/dp, dHc/dvs}
//   g4 = {fu_d, y_d, YSn, dYSn/dT}     g5 = {dfu_d/dT, dy_d/dT, YSn/p, y_d/p}
__global__ void __launch_bounds__(256) prep_grad_kernel(const JacPrepArgs a) {
    const JacItemDev it = a.items[blockIdx.y];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (!it.active || j >= it.n) return;
    const LineListDev L = a.lists[it.list];
    const double T = it.T, p = it.p_atm, vs = it.vmr_self, coef = it.coef, dnu = a.dnu;
    const double nu = L.nu[j];
    const int mol = L.mol[j], iso = L.iso[j], row = L.tips_row[j];
    double qr, dlnq;
    if (row >= 0) {
        const double *v = a.tips_q + (int64_t)row * TIPS_NT;
        const double qt = tips_at(v, T);
        qr = tips_at(v, T_REF) / qt;
        dlnq = tips_slope(v, T) / qt;
    } else {
        qr = T_REF / T;
        dlnq = 1.0 / T;     // Q = T / T_REF
    }
    // temperature_scaled_intensity_deriv (broadening.jl:112-128)
    const double a2 = C2_LINE * nu;
    const double stim = (1.0 - exp(-a2 / T)) / (1.0 - exp(-a2 / T_REF));
    const double boltz = exp(-C2_LINE * L.elower[j] * (1.0 / T - 1.0 / T_REF));
    const double S = L.S[j] * qr * boltz * stim;
    const double dS = S * (-dlnq + C2_LINE * L.elower[j] / (T * T) - (a2 / (T * T)) / (exp(a2 / T) - 1.0));
    // pressure_broadened_width_deriv (broadening.jl:62-78)
    const double n_air = (double)L.nair[j];
    const double n_self = mol == 1 ? 0.95 : (mol == 2 ? 0.75 : n_air);
    const double tr = T_REF / T;
    const double ac = pow(tr, n_air) * (double)L.gair[j], bc = pow(tr, n_self) * (double)L.gself[j];
    const double A_ = ac * (1.0 - vs) * p, B_ = bc * vs * p;
    const double gl = A_ + B_;
    const double dgl = -(n_air * A_ + n_self * B_) / T;
    double gd = nu * 3.58126e-7 * sqrt(T / molmass(mol));
    const double dgd = gd / (2.0 * T);
    gd = fmax(gd, 1e-10);
    const double f = IRS_SQRT_LN2 / gd, df = -f * (dgd / gd);
    const double y = gl * f;
    const double dyT = dgl * f + gl * df, dyp = (gl / p) * f, dyv = ((bc - ac) * p) * f;
    const double sn = S * f * IRS_INV_SQRT_PI, dsnT = (dS * f + S * df) * IRS_INV_SQRT_PI;
    const double dself = mol == 1 ? 0.002 : (mol == 2 ? -0.003 : 0.0);
    const double dair = (double)L.dair[j];
    const double c = nu + (dair * (1.0 - vs) + dself * vs) * p;
    const double d0p = dair * (1.0 - vs) + dself * vs, d0v = (dself - dair) * p;
    double Hc, Hcx, Hcy;
    hgrad64(it.cutoff * f, y, it.x_far, Hc, Hcx, Hcy);
    const double u = (c - a.nu_lo) / dnu, cn = floor(u);
    const int64_t o = it.rec_off + j;
    a.g0[o] = make_float4(__int_as_float((int)cn), (float)(u - cn), (float)(dnu * f), (float)y);
    a.g1[o] = make_float4((float)(coef * sn), (float)(coef * dsnT), (float)(dnu * df), (float)dyT);
    a.g2[o] = make_float4((float)dyp, (float)dyv, (float)(d0p / dnu), (float)(d0v / dnu));
    a.g3[o] = make_float4((float)Hc, (float)(Hcx * (it.cutoff * df) + Hcy * dyT), (float)(Hcy * dyp),
                          (float)(Hcy * dyv));
    float4 g4 = make_float4(1.0f, 1.0f, 0.0f, 0.0f), g5 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (it.lm && L.ylm && L.ylm[j] != 0.0) {
        const double Y = L.ylm[j];
        const double md = lmmass(mol, iso);
        const double gdd = 3.5812e-7 * nu * sqrt(T / md);
        const double gld = (double)L.gair[j] * pow(296.0 / T, n_air) * p;
        const double dgld = -n_air * gld / T;
        const double fd = IRS_SQRT_LN2 / fmax(gdd, 1e-10), dfd = -fd / (2.0 * T);
        const double st0 = 1.0 - exp(-C2_LM * nu / 296.0);
        const double e1 = exp(-C2_LM * nu / T), stt = 1.0 - e1;
        const double st = L.S[j] * qr * exp(-C2_LM * L.elower[j] * (1.0 / T - 1.0 / 296.0)) * stt / st0;
        if (st > 0.0) {
            const double dlnst = -dlnq + C2_LM * L.elower[j] / (T * T) - (C2_LM * nu / (T * T)) * e1 / stt;
            const double ysn = coef * Y * p * st * fd * IRS_INV_SQRT_PI;
            const double yd = gld * fd;
            g4 = make_float4((float)(dnu * fd), (float)yd, (float)ysn, (float)(ysn * (dlnst + dfd / fd)));
            g5 = make_float4((float)(dnu * dfd), (float)(dgld * fd + gld * dfd), (float)(ysn / p),
                             (float)(yd / p));
        }
    }
    a.g4[o] = g4;
    a.g5[o] = g5;
}

hipError_t launch_prep_grad(const JacPrepArgs &a, int max_lines, hipStream_t s) {
    if (a.n_items <= 0 || max_lines <= 0) return hipSuccess;
    hipLaunchKernelGGL(prep_grad_kernel, dim3((max_lines + 255) / 256, a.n_items), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gradient line sum: per (tile, layer) block of 256 * JP grid points, all
// species.  Per species the seven sums (sigma, dsigma/dT, /dp, /dvs and, for
// VP_Y lines, the dispersive term with its T and p derivatives) are accumulated
// per grid point in FP32 with compensation; the clamped totals (coef folded)
// are combined with the item's chain-rule weights into tau, DTlo, DThi and,
// for retrieved species, Vlo/Vhi.  The block owns its outputs: they are
// cleared once and updated per species (read-modify-write by the owning
// thread), so no per-point state lives across species.
//
// Line ranges per species, as in the forward line sum: lines whose window
// covers the whole coarse span and that lie at least far_d grid steps from it
// (and beyond the |x| > x_far core distance from every point of the tile)
// are FAR: their seven sums are evaluated at the coarse nodes and
// interpolated; cutoff-edge lines at least JEDGE_MIN grid steps from every slice
// are evaluated at four nodes per 64-point slice (grad_edge_nodes); every other
// line in the window is evaluated pair by pair.
#define JP IRS_JAC_TILE_P
#define JFAR_MAXNC IRS_JAC_FAR_MAXNC     // coarse nodes per tile
#define JFAR_CH 160                     // FAR lines per LDS chunk
#define JNC 7                           // sums per grid point

__device__ __forceinline__ void kadd(float &s, float &c, float v) {
    const float yv = v - c;
    const float t = s + yv;
    c = (t - s) - yv;
    s = t;
}

__device__ __forceinline__ int lower_key_j(const int32_t *k, int n, int64_t v) {
    int lo = 0, hi = n;
    while (lo < hi) { const int mid = (lo + hi) >> 1; if ((int64_t)k[mid] < v) lo = mid + 1; else hi = mid; }
    return lo;
}

// Gradient records are read with a wave-uniform line index: through the constant
// address space they arrive by scalar loads into SGPRs.
typedef float jac_f4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) jac_f4 *jrec;
__device__ __forceinline__ float4 jld(jrec p, int64_t j) {
    const jac_f4 v = p[j];
    return make_float4(v.x, v.y, v.z, v.w);
}

// Far-branch (|x| > x_far) Voigt terms of one line at offset du from its
// centre, pedestal included (cross_section_jacobian.jl:43-56 and the pedestal
// of :150-200): adds the line's {sigma, dT, dp, dvs} contributions to v[0..3].
__device__ __forceinline__ void far_voigt_terms(float du, const float4 q0, const float4 q1, const float4 q2,
                                                const float4 q3, float iC2, float (&v)[4]) {
    const float ispi = 0.56418958354775628f;
    const float x = du * q0.z, y = q0.w;
    const float r = irs_rcpf(fmaf(x, x, y * y)), r2 = r * r;
    const float H = y * ispi * r;
    const float Hx = -2.0f * x * y * ispi * r2;
    const float Hy = ispi * (x * x - y * y) * r2;
    const float pedfac = fmaf(-du * du, iC2, 2.0f);
    const float hm = H - pedfac * q3.x;
    const float dHt = Hx * (du * q1.z) + Hy * q1.w;
    const float dHp = Hx * (-q2.z * q0.z) + Hy * q2.x;
    const float dHv = Hx * (-q2.w * q0.z) + Hy * q2.y;
    const float dpfp = 2.0f * du * q2.z * iC2, dpfv = 2.0f * du * q2.w * iC2;
    v[0] = fmaf(q1.x, hm, v[0]);
    v[1] += q1.y * hm + q1.x * (dHt - pedfac * q3.y);
    v[2] = fmaf(q1.x, dHp - (dpfp * q3.x + pedfac * q3.z), v[2]);
    v[3] = fmaf(q1.x, dHv - (dpfv * q3.x + pedfac * q3.w), v[3]);
}

// Far-branch dispersive (VP_Y) terms: adds {d0, dT, dp} to v[0..2].  A line
// without line mixing has q4 = {1, 1, 0, 0}, q5 = 0 and adds zero.
__device__ __forceinline__ void far_disp_terms(float du, const float4 q2, const float4 q4, const float4 q5,
                                               float (&v)[3]) {
    const float ispi = 0.56418958354775628f;
    const float xd = du * q4.x, yd = q4.y;
    const float r = irs_rcpf(fmaf(xd, xd, yd * yd)), r2 = r * r;
    const float Lv = xd * ispi * r;
    const float Lx = ispi * (yd * yd - xd * xd) * r2;
    const float Ly = -2.0f * xd * yd * ispi * r2;
    v[0] = fmaf(q4.z, Lv, v[0]);
    v[1] += q4.w * Lv + q4.z * (Lx * (du * q5.x) + Ly * q5.y);
    v[2] += q5.z * Lv + q4.z * (Lx * (-q2.z * q4.x) + Ly * q5.w);
}

// The seven terms of one line at one point (offset du from the centre, dI the
// integer part of that offset): exact branch selection, window tests and
// pedestal (cross_section_jacobian.jl:150-240).  v[0..3] (Voigt) are valid when
// vin, v[4..6] (dispersive, VP_Y lines) when din.
__device__ __forceinline__ void pair_terms(jrec g4, jrec g5, int j, const float4 q0, const float4 q1, const float4 q2,
                                           const float4 q3, bool dl, float du, int dI, float C2, float iC2, int CiD,
                                           float CfD, float xfar, float (&v)[JNC], bool &vin, bool &din) {
    const float ispi = 0.56418958354775628f;
    const float d2 = du * du;
    const float x = du * q0.z, y = q0.w;
    float H, Hx, Hy;
    if (fabsf(x) > xfar) {
        const float dd = fmaf(x, x, y * y), r = irs_rcpf(dd), r2 = r * r;
        H = y * ispi * r;
        Hx = -2.0f * x * y * ispi * r2;
        Hy = ispi * (x * x - y * y) * r2;
    } else {
        float K, Lw;
        w32(x, y, K, Lw);
        H = K;
        Hx = -2.0f * x * K + 2.0f * y * Lw;
        Hy = 2.0f * (x * Lw + y * K) - 2.0f * ispi;
    }
    vin = d2 <= C2;
    if (vin) {
        const float pedfac = fmaf(-d2, iC2, 2.0f);
        const float hm = H - pedfac * q3.x;
        const float dHt = Hx * (du * q1.z) + Hy * q1.w;
        const float dHp = Hx * (-q2.z * q0.z) + Hy * q2.x;
        const float dHv = Hx * (-q2.w * q0.z) + Hy * q2.y;
        const float dpfp = 2.0f * du * q2.z * iC2, dpfv = 2.0f * du * q2.w * iC2;
        v[0] = q1.x * hm;
        v[1] = q1.y * hm + q1.x * (dHt - pedfac * q3.y);
        v[2] = q1.x * (dHp - (dpfp * q3.x + pedfac * q3.z));
        v[3] = q1.x * (dHv - (dpfv * q3.x + pedfac * q3.w));
    }
    din = dl && in_window(dI, q0.y, CiD, CfD);
    if (din) {
        const float4 q4 = jld(g4, j), q5 = jld(g5, j);
        const float xd = du * q4.x, yd = q4.y;
        float Lv, Lx, Ly;
        if (fabsf(xd) > xfar) {
            const float dd = fmaf(xd, xd, yd * yd), r = irs_rcpf(dd), r2 = r * r;
            Lv = xd * ispi * r;
            Lx = ispi * (yd * yd - xd * xd) * r2;
            Ly = -2.0f * xd * yd * ispi * r2;
        } else {
            float K, Lw;
            w32(xd, yd, K, Lw);
            Lv = Lw;
            Lx = -2.0f * (xd * Lw + yd * K) + 2.0f * ispi;
            Ly = -2.0f * xd * K + 2.0f * yd * Lw;
        }
        v[4] = q4.z * Lv;
        v[5] = q4.w * Lv + q4.z * (Lx * (du * q5.x) + Ly * q5.y);
        v[6] = q5.z * Lv + q4.z * (Lx * (-q2.z * q4.x) + Ly * q5.w);
    }
}

// Pair-by-pair sums of the lines [jlo, jhi) at every point of the tile.
__device__ __forceinline__ void grad_pairs(jrec g0, jrec g1, jrec g2, jrec g3, jrec g4, jrec g5, int jlo, int jhi,
                                           int tl, bool lm, float C2, float iC2, int CiD, float CfD, float xfar,
                                           float (&sa)[JNC][JP], float (&ca)[JNC][JP]) {
    for (int j = jlo; j < jhi; ++j) {
        const float4 q0 = jld(g0, j), q1 = jld(g1, j), q2 = jld(g2, j), q3 = jld(g3, j);
        const int dI0 = tl - __float_as_int(q0.x);
        const float base = (float)dI0;
        const bool dl = lm && jld(g4, j).z != 0.0f;
#pragma unroll
        for (int p = 0; p < JP; ++p) {             // unrolled: static indices into the sums (330 -> 251 ms)
            const float du = (base + (float)(256 * p)) - q0.y;
            float v[JNC];
            bool vin, din;
            pair_terms(g4, g5, j, q0, q1, q2, q3, dl, du, dI0 + 256 * p, C2, iC2, CiD, CfD, xfar, v, vin, din);
            if (vin) {
#pragma unroll
                for (int c = 0; c < 4; ++c) kadd(sa[c][p], ca[c][p], v[c]);
            }
            if (din) {
#pragma unroll
                for (int c = 4; c < JNC; ++c) kadd(sa[c][p], ca[c][p], v[c]);
            }
        }
    }
}

// EDGE lines (windows reaching into the tile without covering the coarse span),
// every point in the far branch: per 64-point slice, lanes take one line each
// and evaluate its seven terms (far-branch form, pedestal included) at the four
// nodes l = 0, 21, 42, 63 of the slice; the wave sums of the node values give
// every point its value by cubic Lagrange interpolation.  The terms are smooth
// on the scale of the line's distance D (>= JEDGE_MIN grid steps): the
// interpolant's error is below 15 (21/D)^4 (1.1e-8 at D = 4000) of a line's
// contribution, and it reproduces the quadratic pedestal exactly.  A line whose
// window boundary falls inside the slice is evaluated pair by pair there.
#define JEDGE_MIN 4000
__device__ __forceinline__ void grad_edge_nodes(const JacSumArgs &a, const JacItemDev &it, jrec g0s, jrec g1s,
                                                jrec g2s, jrec g3s, jrec g4s, jrec g5s, int jl0, int jl1, int jr0,
                                                int jr1, int tl, bool lm, float C2, float iC2, int CiD, float CfD,
                                                float (&sa)[JNC][JP], float (&ca)[JNC][JP]) {
    const int nL = jl1 - jl0, ntot = nL + (jr1 - jr0);
    if (ntot <= 0) return;
    const int lane = threadIdx.x & 63;
    const int tw = __builtin_amdgcn_readfirstlane(tl & ~63);   // first point of this wave's slices
    const int tlw = tl - tw;                                      // = lane
    const float4 *g0 = a.g0 + it.rec_off, *g1 = a.g1 + it.rec_off, *g2 = a.g2 + it.rec_off;
    const float4 *g3 = a.g3 + it.rec_off, *g4 = a.g4 + it.rec_off, *g5 = a.g5 + it.rec_off;
    const float C = (float)it.C, Clm = (float)it.Clm;
    const float inV = C - 31.55f, outV = C + 31.55f, inD = Clm - 31.55f, outD = Clm + 31.55f;
    const int jsafe = nL > 0 ? jl0 : jr0;
    // cubic Lagrange weights of this lane's point on the nodes 0, 21, 42, 63
    const float u = (float)tlw * (1.0f / 21.0f);
    const float w0 = -(u - 1.0f) * (u - 2.0f) * (u - 3.0f) * (1.0f / 6.0f);
    const float w1 = u * (u - 2.0f) * (u - 3.0f) * 0.5f;
    const float w2 = -u * (u - 1.0f) * (u - 3.0f) * 0.5f;
    const float w3 = u * (u - 1.0f) * (u - 2.0f) * (1.0f / 6.0f);
#pragma unroll
    for (int p = 0; p < JP; ++p) {
        float nv[4][4], nd[3][4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            nv[0][m] = nv[1][m] = nv[2][m] = nv[3][m] = 0.0f;
            nd[0][m] = nd[1][m] = nd[2][m] = 0.0f;
        }
        for (int b0 = 0; b0 < ntot; b0 += 64) {
            const int v = b0 + lane;
            const bool ok = v < ntot;
            const int j = ok ? (v < nL ? jl0 + v : jr0 + (v - nL)) : jsafe;
            const float4 q0 = g0[j];
            const int cidx = __float_as_int(q0.x);
            const float d = (float)(tw + 256 * p - cidx) + (31.5f - q0.y);     // slice centre - line centre
            const float ad = fabsf(d);
            const bool fullV = ok && ad <= inV, bndV = ok && !fullV && ad < outV;
            bool dl = false, fullD = false, bndD = false;
            if (lm) {
                dl = ok && g4[j].z != 0.0f;
                fullD = dl && ad <= inD;
                bndD = dl && !fullD && ad < outD;
            }
            if (fullV || fullD) {
                const float4 q1 = g1[j], q2 = g2[j], q3 = g3[j];
                const float4 q4 = lm ? g4[j] : make_float4(1.0f, 1.0f, 0.0f, 0.0f);
                const float4 q5 = lm ? g5[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float du = (float)(tw + 256 * p + 21 * m - cidx) - q0.y;
                    if (fullV) {
                        float t4[4] = {nv[0][m], nv[1][m], nv[2][m], nv[3][m]};
                        far_voigt_terms(du, q0, q1, q2, q3, iC2, t4);
                        nv[0][m] = t4[0]; nv[1][m] = t4[1]; nv[2][m] = t4[2]; nv[3][m] = t4[3];
                    }
                    if (fullD) {
                        float t3[3] = {nd[0][m], nd[1][m], nd[2][m]};
                        far_disp_terms(du, q2, q4, q5, t3);
                        nd[0][m] = t3[0]; nd[1][m] = t3[1]; nd[2][m] = t3[2];
                    }
                }
            }
            // lines whose window boundary falls in this slice: pair by pair
            const unsigned long long mV = __ballot(bndV), mD = lm ? __ballot(bndD) : 0ull;
            unsigned long long mk = mV | mD;
            while (mk) {
                const int l = __builtin_ctzll(mk);
                mk &= mk - 1;
                const int jj = __builtin_amdgcn_readlane(j, l);
                const float4 e0 = jld(g0s, jj), e1 = jld(g1s, jj), e2 = jld(g2s, jj), e3 = jld(g3s, jj);
                const int dI = tl - __float_as_int(e0.x) + 256 * p;
                const float du = ((float)(tl - __float_as_int(e0.x)) + (float)(256 * p)) - e0.y;
                // far branch at every point of these lines: window tests as in pair_terms
                if (((mV >> l) & 1ull) && du * du <= C2) {
                    float t4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    far_voigt_terms(du, e0, e1, e2, e3, iC2, t4);
#pragma unroll
                    for (int c = 0; c < 4; ++c) kadd(sa[c][p], ca[c][p], t4[c]);
                }
                if (((mD >> l) & 1ull) && in_window(dI, e0.y, CiD, CfD)) {
                    float t3[3] = {0.0f, 0.0f, 0.0f};
                    far_disp_terms(du, e2, jld(g4s, jj), jld(g5s, jj), t3);
#pragma unroll
                    for (int c = 0; c < 3; ++c) kadd(sa[4 + c][p], ca[4 + c][p], t3[c]);
                }
            }
        }
        // wave sums of the node values, then the interpolant at each point
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float n0 = nv[c][0], n1 = nv[c][1], n2 = nv[c][2], n3 = nv[c][3];
            for (int o = 32; o > 0; o >>= 1) {
                n0 += __shfl_xor(n0, o, 64);
                n1 += __shfl_xor(n1, o, 64);
                n2 += __shfl_xor(n2, o, 64);
                n3 += __shfl_xor(n3, o, 64);
            }
            kadd(sa[c][p], ca[c][p], w0 * n0 + w1 * n1 + w2 * n2 + w3 * n3);
        }
        if (lm) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float n0 = nd[c][0], n1 = nd[c][1], n2 = nd[c][2], n3 = nd[c][3];
                for (int o = 32; o > 0; o >>= 1) {
                    n0 += __shfl_xor(n0, o, 64);
                    n1 += __shfl_xor(n1, o, 64);
                    n2 += __shfl_xor(n2, o, 64);
                    n3 += __shfl_xor(n3, o, 64);
                }
                kadd(sa[4 + c][p], ca[4 + c][p], w0 * n0 + w1 * n1 + w2 * n2 + w3 * n3);
            }
        }
    }
}

// FAR lines [jl0, jl1) and [jr0, jr1) at the NC coarse nodes Sl + n*HC.  The
// records stream through LDS in chunks of JFAR_CH ({centre index - Sl,
// fraction, dnu f, y} and g1..g3, plus g4, g5 with line mixing).  Thread t
// evaluates the nodes 2g and 2g+1 (g = t % NG, NG = ceil(NC/2)) for the
// chunk's lines k = s, s+S, ... (s = t / NG, S = 256 / NG slices); the node
// offsets are exact integers, so du carries one rounding (the fraction).  Per
// node the S slice sums of the seven terms are combined in FP64 through the
// (then idle) staging buffer, and every grid point of the tile adds the
// four-point (cubic) Lagrange interpolant of them.  The interpolant
// reproduces the quadratic pedestal exactly, and for a Lorentz wing at
// distance >= D its error is below 2.8 (HC/D)^4 of the line's contribution
// (15 (HC/D)^4 for the 1/du^3 derivative terms).
__device__ __forceinline__ void grad_far_coarse(const JacSumArgs &a, const JacItemDev &it, int jl0, int jl1, int jr0,
                                                int jr1, int t, int64_t i0, int64_t Sl, int HC, int NC, bool lm,
                                                float iC2, float4 *s_q, float *s_F, float (&sa)[JNC][JP],
                                                float (&ca)[JNC][JP]) {
    const int NG = (NC + 1) >> 1;
    const int S = 256 / NG;
    const int g = t % NG, sl = t / NG;
    const bool act = t < S * NG;
    const int nL = jl1 - jl0, ntot = nL + (jr1 - jr0);
    const float xrel = (float)(2 * g * HC);           // offset of node 2g from Sl (exact)
    const float fhc = (float)HC;
    float A[JNC][2], Ac[JNC][2];
#pragma unroll
    for (int c = 0; c < JNC; ++c) A[c][0] = A[c][1] = Ac[c][0] = Ac[c][1] = 0.0f;
    const float4 *g0 = a.g0 + it.rec_off, *g1 = a.g1 + it.rec_off, *g2 = a.g2 + it.rec_off;
    const float4 *g3 = a.g3 + it.rec_off, *g4 = a.g4 + it.rec_off, *g5 = a.g5 + it.rec_off;
    for (int base = 0; base < ntot; base += JFAR_CH) {
        const int cnt = min(JFAR_CH, ntot - base);
        if (t < cnt) {                                   // JFAR_CH <= 256: one record per thread
            const int v = base + t;
            const int j = v < nL ? jl0 + v : jr0 + (v - nL);
            const float4 q = g0[j];
            s_q[t] = make_float4((float)(__float_as_int(q.x) - (int)Sl), q.y, q.z, q.w);
            s_q[JFAR_CH + t] = g1[j];
            s_q[2 * JFAR_CH + t] = g2[j];
            s_q[3 * JFAR_CH + t] = g3[j];
            if (lm) {
                s_q[4 * JFAR_CH + t] = g4[j];
                s_q[5 * JFAR_CH + t] = g5[j];
            }
        }
        __syncthreads();
        if (act) {
            float part[JNC][2];
#pragma unroll
            for (int c = 0; c < JNC; ++c) part[c][0] = part[c][1] = 0.0f;
            int n = 0;
            for (int k = sl; k < cnt; k += S) {
                const float4 q0 = s_q[k], q1 = s_q[JFAR_CH + k], q2 = s_q[2 * JFAR_CH + k];
                const float4 q3 = s_q[3 * JFAR_CH + k];
                const float dI = xrel - q0.x;            // integer, exact
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    const float du = (m ? dI + fhc : dI) - q0.y;
                    float v[4] = {part[0][m], part[1][m], part[2][m], part[3][m]};
                    far_voigt_terms(du, q0, q1, q2, q3, iC2, v);
                    part[0][m] = v[0]; part[1][m] = v[1]; part[2][m] = v[2]; part[3][m] = v[3];
                }
                if (lm) {
                    const float4 q4 = s_q[4 * JFAR_CH + k], q5 = s_q[5 * JFAR_CH + k];
#pragma unroll
                    for (int m = 0; m < 2; ++m) {
                        const float du = (m ? dI + fhc : dI) - q0.y;
                        float v[3] = {part[4][m], part[5][m], part[6][m]};
                        far_disp_terms(du, q2, q4, q5, v);
                        part[4][m] = v[0]; part[5][m] = v[1]; part[6][m] = v[2];
                    }
                }
                if (++n == 32) {
#pragma unroll
                    for (int c = 0; c < JNC; ++c) {
                        kadd(A[c][0], Ac[c][0], part[c][0]);
                        kadd(A[c][1], Ac[c][1], part[c][1]);
                        part[c][0] = part[c][1] = 0.0f;
                    }
                    n = 0;
                }
            }
#pragma unroll
            for (int c = 0; c < JNC; ++c) {
                kadd(A[c][0], Ac[c][0], part[c][0]);
                kadd(A[c][1], Ac[c][1], part[c][1]);
            }
        }
        __syncthreads();
    }
    // per-(slice, node, term) partial sums in the staging buffer: S * 2NG * 7 <= 3584 floats
    float *red = (float *)s_q;
    if (act) {
#pragma unroll
        for (int c = 0; c < JNC; ++c) {
            red[((sl * 2 * NG) + 2 * g) * JNC + c] = A[c][0] - Ac[c][0];
            red[((sl * 2 * NG) + 2 * g + 1) * JNC + c] = A[c][1] - Ac[c][1];
        }
    }
    __syncthreads();
    if (t < NC) {
#pragma unroll
        for (int c = 0; c < JNC; ++c) {
            double v = 0.0;
            for (int q = 0; q < S; ++q) v += (double)red[((q * 2 * NG) + t) * JNC + c];
            s_F[c * JFAR_MAXNC + t] = (float)v;
        }
    }
    __syncthreads();
    const float ihc = 1.0f / (float)HC;
    const int nterm = lm ? JNC : 4;
#pragma unroll
    for (int p = 0; p < JP; ++p) {
        const int off = (int)(i0 - Sl) + t + 256 * p;          // grid offset from node 0
        const int n0 = off / HC;                               // node left of the point, >= 1
        const float u = (float)(off - n0 * HC) * ihc;          // in [0, 1)
        const float wm = -u * (u - 1.0f) * (u - 2.0f) * (1.0f / 6.0f);
        const float w0 = (u + 1.0f) * (u - 1.0f) * (u - 2.0f) * 0.5f;
        const float w1 = -(u + 1.0f) * u * (u - 2.0f) * 0.5f;
        const float w2 = (u + 1.0f) * u * (u - 1.0f) * (1.0f / 6.0f);
#pragma unroll
        for (int c = 0; c < JNC; ++c) {
            if (c >= nterm) break;
            const float *F = s_F + c * JFAR_MAXNC + n0;
            kadd(sa[c][p], ca[c][p], wm * F[-1] + w0 * F[0] + w1 * F[1] + w2 * F[2]);
        }
    }
    __syncthreads();
}

// Four resident workgroups per CU (128 VGPRs, some spilled): the full-IASI gradient
// sum measured 252 / 192 / 169 / 199 ms at 2 (unbounded, 193 VGPRs) / 3 / 4 / 5 waves
// per SIMD.
#ifndef IRS_JAC_MINB
#define IRS_JAC_MINB 4
#endif
__global__ void __launch_bounds__(256, IRS_JAC_MINB) grad_linesum_kernel(const JacSumArgs a) {
    __shared__ int s_rng[6];
    __shared__ float4 s_q[6 * JFAR_CH];
    __shared__ float s_F[JNC * JFAR_MAXNC];
    const int t = threadIdx.x;
    const int layer = blockIdx.y;
    const int64_t i0 = (int64_t)blockIdx.x * (256 * JP);
    const int64_t i1 = min(i0 + (int64_t)(256 * JP), a.n_hi);
    const int tl = (int)(i0 + t);
#pragma unroll
    for (int p = 0; p < JP; ++p) {
        const int64_t i = i0 + t + 256 * p;
        if (i >= a.n_hi) continue;
        const int64_t o = (int64_t)layer * a.n_hi + i;
        a.tau[o] = 0.0f;
        a.dtlo[o] = 0.0f;
        a.dthi[o] = 0.0f;
        for (int r = 0; r < a.n_retr; ++r) {
            a.vlo[((int64_t)r * a.n_layers + layer) * a.n_hi + i] = 0.0f;
            a.vhi[((int64_t)r * a.n_layers + layer) * a.n_hi + i] = 0.0f;
        }
    }
    // coarse far-wing split: nodes Sl + n*HC, n = 0..NC-1 (HC divides the tile width)
    const int HC = a.far_hc;
    const int NC = HC > 0 ? (256 * JP) / HC + 3 : 0;
    const int64_t Sl = i0 - HC, Sh = Sl + (int64_t)(NC - 1) * HC, Dd = a.far_d;
    for (int s = 0; s < a.n_species; ++s) {
        const JacItemDev it = a.items[(int64_t)layer * a.n_species + s];
        if (!it.active || it.n == 0) continue;
        const bool lm = it.lm != 0;
        const int32_t *keys = a.keys + it.key_off;
        const double Cmx = lm ? fmax((double)it.C, (double)it.Clm) : (double)it.C;
        const double Cmn = lm ? fmin((double)it.C, (double)it.Clm) : (double)it.C;
        const int64_t Cmax = (int64_t)ceil(Cmx), Cmin = (int64_t)floor(Cmn), M = it.M;
        // key bounds (centre index c in [key - M, key + M]): the window [vA, vF],
        // the core distance K (|x| <= x_far within K grid steps of the centre)
        const double numax = a.nu_lo + (double)(i1 + Cmax + M) * a.dnu;
        const int64_t K = (int64_t)ceil(it.kcore_per_nu * numax) + 2;
        const int64_t vA = i0 - Cmax - M - 1, vF = i1 - 1 + Cmax + M + 1;
        const int64_t vC = i0 - K - M - 1, vD = i1 - 1 + K + M + 1;
        const int64_t vB2 = Sh - Cmin + M + 1, vG = Sl - Dd - M - 1;
        const int64_t vH = Sh + Dd + M, vE2 = Sl + Cmin - M - 2;
        const bool far = HC > 0 && NC <= JFAR_MAXNC && vA <= vB2 && vB2 <= vG && vG < vC && vD < vH &&
                         vH <= vE2 && vE2 <= vF;
        if (t < 6) {
            int64_t v;
            switch (t) {
                case 0: v = vA; break;
                case 1: v = vF + 1; break;
                case 2: v = vB2; break;
                case 3: v = vG + 1; break;
                case 4: v = vH; break;
                default: v = vE2 + 1; break;
            }
            s_rng[t] = (t < 2 || far) ? lower_key_j(keys, it.n, v) : 0;
        }
        __syncthreads();
        const int jA = __builtin_amdgcn_readfirstlane(s_rng[0]);
        const int jF = __builtin_amdgcn_readfirstlane(s_rng[1]);
        const int jB2 = far ? __builtin_amdgcn_readfirstlane(s_rng[2]) : jF;
        const int jG = far ? __builtin_amdgcn_readfirstlane(s_rng[3]) : jF;
        const int jH = far ? __builtin_amdgcn_readfirstlane(s_rng[4]) : jF;
        const int jE2 = far ? __builtin_amdgcn_readfirstlane(s_rng[5]) : jF;
        __syncthreads();
        const jrec g0 = (jrec)(a.g0 + it.rec_off), g1 = (jrec)(a.g1 + it.rec_off), g2 = (jrec)(a.g2 + it.rec_off);
        const jrec g3 = (jrec)(a.g3 + it.rec_off), g4 = (jrec)(a.g4 + it.rec_off), g5 = (jrec)(a.g5 + it.rec_off);
        const float C2 = (float)((double)it.C * it.C);
        const double Cld = fmin(it.lm_cutoff / a.dnu, 1.0e9);
        const int CiD = (int)floor(Cld);
        const float CfD = (float)(Cld - (double)CiD);
        const float iC2 = (float)(1.0 / ((double)it.C * it.C));
        const float xfar = (float)it.x_far;
        float sa[JNC][JP], ca[JNC][JP];
#pragma unroll
        for (int c = 0; c < JNC; ++c)
#pragma unroll
            for (int p = 0; p < JP; ++p) sa[c][p] = ca[c][p] = 0.0f;
        if (far) {
            if (jG > jB2 || jE2 > jH)
                grad_far_coarse(a, it, jB2, jG, jH, jE2, t, i0, Sl, HC, NC, lm, iC2, s_q, s_F, sa, ca);
            grad_pairs(g0, g1, g2, g3, g4, g5, jG, jH, tl, lm, C2, iC2, CiD, CfD, xfar, sa, ca);
            // EDGE lines sit at least Cmin - (tile + 2 HC + 2 M) from every slice: by
            // nodes when that is >= JEDGE_MIN and beyond the core distance K
            const bool edge_nodes = (double)Cmin - (double)(256 * JP + 2 * HC + 2 * M + 64) >= (double)JEDGE_MIN &&
                                    K + 64 < JEDGE_MIN;
            if (edge_nodes) {
                grad_edge_nodes(a, it, g0, g1, g2, g3, g4, g5, jA, jB2, jE2, jF, tl, lm, C2, iC2, CiD, CfD, sa, ca);
            } else {
                grad_pairs(g0, g1, g2, g3, g4, g5, jA, jB2, tl, lm, C2, iC2, CiD, CfD, xfar, sa, ca);
                grad_pairs(g0, g1, g2, g3, g4, g5, jE2, jF, tl, lm, C2, iC2, CiD, CfD, xfar, sa, ca);
            }
        } else {
            grad_pairs(g0, g1, g2, g3, g4, g5, jA, jF, tl, lm, C2, iC2, CiD, CfD, xfar, sa, ca);
        }
#pragma unroll
        for (int p = 0; p < JP; ++p) {
            const int64_t i = i0 + t + 256 * p;
            if (i >= a.n_hi) continue;
            double sg = (double)sa[0][p] - (double)ca[0][p], gT = (double)sa[1][p] - (double)ca[1][p];
            double gP = (double)sa[2][p] - (double)ca[2][p], gV = (double)sa[3][p] - (double)ca[3][p];
            if (sg <= 0.0) { sg = 0.0; gT = 0.0; gP = 0.0; gV = 0.0; }
            if (lm) {
                sg += (double)sa[4][p] - (double)ca[4][p];
                gT += (double)sa[5][p] - (double)ca[5][p];
                gP += (double)sa[6][p] - (double)ca[6][p];
                if (sg <= 0.0) { sg = 0.0; gT = 0.0; gP = 0.0; }
            }
            const int64_t o = (int64_t)layer * a.n_hi + i;
            if (a.lm0 && s == a.lm_s) {
                // relaxation-matrix perturbation and its central differences in (T, p),
                // then the forward's max(sigma + dsigma, 0) clamp
                // (cross_section_jacobian.jl:292-319)
                sg += (double)a.lm0[o];
                gT += (double)a.lmT[o];
                gP += (double)a.lmP[o];
                if (sg <= 0.0) { sg = 0.0; gT = 0.0; gP = 0.0; }
            }
            a.tau[o] += (float)sg;
            a.dtlo[o] += (float)(gT * it.wT_lo);
            a.dthi[o] += (float)(gT * it.wT_hi);
            if (it.retr >= 0) {
                const int64_t ov = ((int64_t)it.retr * a.n_layers + layer) * a.n_hi + i;
                a.vlo[ov] += (float)(sg * it.gv_lo + gP * it.gp_lo + gT * it.gt_lo + gV * it.gs_lo);
                a.vhi[ov] += (float)(sg * it.gv_hi + gP * it.gp_hi + gT * it.gt_hi + gV * it.gs_hi);
            }
        }
    }
}

// Central differences of the line-mixing perturbation of one layer: d [5][n] at
// (T, p), (T + hT, p), (T - hT, p), (T, p + hp), (T, p - hp); coef folded in.
__global__ void __launch_bounds__(256) lm_fd_kernel(const double *d, int64_t n, double coef, double hT, double hp,
                                                    float *lm0, float *lmT, float *lmP) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    lm0[i] = (float)(coef * d[i]);
    lmT[i] = (float)(coef * (d[n + i] - d[2 * n + i]) / (2.0 * hT));
    lmP[i] = (float)(coef * (d[3 * n + i] - d[4 * n + i]) / (2.0 * hp));
}

hipError_t launch_lm_fd(const double *d, int64_t n, double coef, double hT, double hp, float *lm0, float *lmT,
                        float *lmP, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lm_fd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n, coef, hT, hp, lm0,
                       lmT, lmP);
    return hipGetLastError();
}

hipError_t launch_grad_linesum(const JacSumArgs &a, hipStream_t s) {
    const unsigned tiles = (unsigned)((a.n_hi + 256 * JP - 1) / (256 * JP));
    hipLaunchKernelGGL(grad_linesum_kernel, dim3(tiles, a.n_layers), dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// RTE Jacobian (FP64), one thread per grid point.  Transmittances come from the
// FP64 optical-depth sum from the top (T_k = exp(-sum_{j>=k} a_j)).
__device__ __forceinline__ double cimf(double s, double e) { return s < 0.06 ? s / 6.0 : 1.0 - 2.0 * (e / (e - 1.0) + 1.0 / s); }
__device__ __forceinline__ double cimd(double s, double e) { return s < 0.06 ? 1.0 / 6.0 : 2.0 / (s * s) - 2.0 * e / ((e - 1.0) * (e - 1.0)); }
__device__ __forceinline__ double litf(double s, double e) { return s < 1e-4 ? s * (0.5 - s * (1.0 / 3.0 - s * 0.125)) : (1.0 - e) / s - e; }
__device__ __forceinline__ double litd(double s, double e) { return s < 1e-4 ? 0.5 - s * (2.0 / 3.0 - s * 0.375) : (e * (s + 1.0) - 1.0) / (s * s) + e; }
__device__ __forceinline__ double dbdt64(double nu, double T) {
    const double x = IRS_C2_PL * nu / T, ex = exp(x);
    return IRS_C1_PL * IRS_C2_PL * nu * nu * nu * nu * ex / (T * T * (ex - 1.0) * (ex - 1.0));
}

__global__ void __launch_bounds__(256) rte_jac_kernel(const RteJacArgs a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int L = a.n_layers;
    const int64_t n = a.n;
    const double nu = (a.an + (double)i * a.sn) / a.den;
    const double invmu = 1.0 / a.mu, eps = a.eps, refl = 1.0 - eps;
    double Stot = 0.0;
    for (int k = 0; k < L; ++k) Stot += (double)a.tau[(int64_t)k * n + i] * invmu;
    const double t1 = exp(-Stot);
    const double bs = planck64(nu, a.T_sfc);
    double I = eps * bs * t1;
    double deps = bs * t1;
    const double dtsfc = eps * dbdt64(nu, a.T_sfc) * t1;
    double below = eps * bs * t1;
    double Aabove = Stot;                          // OD from level m to TOA
    for (int m = 0; m < L; ++m) {
        const double am = (double)a.tau[(int64_t)m * n + i] * invmu;
        const double w = exp(-am), u = 1.0 - w;
        Aabove -= am;
        const double ttop = exp(-fmax(Aabove, 0.0));
        double em, dem;
        if (a.source_function == 1) {
            const double bavg = planck64(nu, a.T_ave[m]), btop = planck64(nu, a.T_lev[m + 1]);
            const double f = cimf(am, w), fp = cimd(am, w);
            const double core = bavg + (btop - bavg) * f;
            em = u * core;
            dem = w * core + u * (btop - bavg) * fp;
            const double cavg = ttop * u * (1.0 - f) * dbdt64(nu, a.T_ave[m]);
            a.dTlev[(int64_t)m * n + i] += cavg * (1.0 - a.frac[m]);
            a.dTlev[(int64_t)(m + 1) * n + i] += cavg * a.frac[m] + ttop * u * f * dbdt64(nu, a.T_lev[m + 1]);
        } else {
            const double bbot = planck64(nu, a.T_lev[m]), btop = planck64(nu, a.T_lev[m + 1]);
            const double fl = litf(am, w), flp = litd(am, w);
            em = btop * u + (bbot - btop) * fl;
            dem = btop * w + (bbot - btop) * flp;
            a.dTlev[(int64_t)m * n + i] += ttop * fl * dbdt64(nu, a.T_lev[m]);
            a.dTlev[(int64_t)(m + 1) * n + i] += ttop * (u - fl) * dbdt64(nu, a.T_lev[m + 1]);
        }
        a.dtau[(int64_t)m * n + i] = invmu * (-below + ttop * dem);
        I += ttop * em;
        below += ttop * em;
    }
    if (eps < 1.0) {
        double g = 1.0, down = 0.0;
        for (int k = 0; k < L; ++k) {
            const double ak = (double)a.tau[(int64_t)k * n + i] * invmu;
            const double w = exp(-ak), u = 1.0 - w;
            double em;
            if (a.source_function == 1) {
                const double bavg = planck64(nu, a.T_ave[k]), bbot = planck64(nu, a.T_lev[k]);
                const double f = cimf(ak, w);
                em = u * (bavg + (bbot - bavg) * f);
                const double cavg = refl * t1 * g * u * (1.0 - f) * dbdt64(nu, a.T_ave[k]);
                a.dTlev[(int64_t)k * n + i] += cavg * (1.0 - a.frac[k]) + refl * t1 * g * u * f * dbdt64(nu, a.T_lev[k]);
                a.dTlev[(int64_t)(k + 1) * n + i] += cavg * a.frac[k];
            } else {
                const double bbot = planck64(nu, a.T_lev[k]), btop = planck64(nu, a.T_lev[k + 1]);
                const double fl = litf(ak, w);
                em = bbot * u + (btop - bbot) * fl;
                a.dTlev[(int64_t)k * n + i] += refl * t1 * g * (u - fl) * dbdt64(nu, a.T_lev[k]);
                a.dTlev[(int64_t)(k + 1) * n + i] += refl * t1 * g * fl * dbdt64(nu, a.T_lev[k + 1]);
            }
            down += g * em;
            g *= w;
        }
        I += refl * down * t1;
        deps += -down * t1;
        // backward sweep: g_m recomputed from the OD below level m
        double above = 0.0, Abelow = Stot;      // Abelow: OD from surface to level m+1 after update
        double odtop = 0.0;                      // OD from level m+1 to TOA
        for (int m = L - 1; m >= 0; --m) {
            const double am = (double)a.tau[(int64_t)m * n + i] * invmu;
            Abelow = Stot - odtop - am;          // OD from surface to level m
            const double gm = exp(-fmax(Abelow, 0.0));
            const double w = exp(-am), u = 1.0 - w;
            double em, dem;
            if (a.source_function == 1) {
                const double bavg = planck64(nu, a.T_ave[m]), bbot = planck64(nu, a.T_lev[m]);
                const double f = cimf(am, w), fp = cimd(am, w);
                const double core = bavg + (bbot - bavg) * f;
                em = u * core;
                dem = w * core + u * (bbot - bavg) * fp;
            } else {
                const double bbot = planck64(nu, a.T_lev[m]), btop = planck64(nu, a.T_lev[m + 1]);
                const double fl = litf(am, w), flp = litd(am, w);
                em = bbot * u + (btop - bbot) * fl;
                dem = bbot * w + (btop - bbot) * flp;
            }
            a.dtau[(int64_t)m * n + i] += refl * invmu * t1 * (-above + gm * dem - down);
            above += gm * em;
            odtop += am;
        }
    }
    a.I[i] = I;
    a.dtsfc[i] = dtsfc;
    a.deps[i] = deps;
}

hipError_t launch_rte_jac(const RteJacArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(rte_jac_kernel, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Column assembly: one thread per grid point, one column per blockIdx.y.
__global__ void __launch_bounds__(256) jac_columns_kernel(const JacColArgs a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int c = blockIdx.y;
    const JacColDesc d = a.cols[c];
    const int64_t n = a.n;
    const int L = a.n_layers;
    double v = 0.0;
    if (d.kind == 0) {            // temperature level j
        const int j = d.level;
        v = a.dTlev[(int64_t)j * n + i];
        if (j < L) v += a.dtau[(int64_t)j * n + i] * (double)a.dtlo[(int64_t)j * n + i];
        if (j >= 1) v += a.dtau[(int64_t)(j - 1) * n + i] * (double)a.dthi[(int64_t)(j - 1) * n + i];
    } else if (d.kind == 1) {     // VMR level(s) of retrieved species r (weight per level)
        for (int j = d.level; j <= d.level_end; ++j) {
            const double wj = a.level_w[(int64_t)d.r * (L + 1) + j];
            if (wj == 0.0) continue;
            double x = 0.0;
            if (j < L) x += a.dtau[(int64_t)j * n + i] * (double)a.vlo[((int64_t)d.r * L + j) * n + i];
            if (j >= 1) x += a.dtau[(int64_t)(j - 1) * n + i] * (double)a.vhi[((int64_t)d.r * L + j - 1) * n + i];
            v += wj * x;
        }
    } else if (d.kind == 2) {
        v = a.dtsfc[i];
    } else {
        v = a.deps[i];
    }
    a.out[(int64_t)c * n + i] = v;
}

hipError_t launch_jac_columns(const JacColArgs &a, int n_cols, hipStream_t s) {
    hipLaunchKernelGGL(jac_columns_kernel, dim3((unsigned)((a.n + 255) / 256), n_cols), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Multi-column channel ILS: block per (channel, column), FP64 reduction, times
// the per-channel observable factor (dBT/dR or 1).
__global__ void __launch_bounds__(256) jac_ils_kernel(const JacIlsArgs a) {
    __shared__ double red[4];
    const int64_t ch = blockIdx.x;
    const int c = blockIdx.y;
    const double *col = a.cols + (int64_t)c * a.n;
    const int64_t ic = a.ch_idx[ch];
    const double w = a.ch_w[ch];
    double out = 0.0;
    for (int side = 0; side < (w != 0.0 ? 2 : 1); ++side) {
        const int64_t i = ic + side;
        double acc = 0.0;
        const int64_t nt = 2 * a.half + 1;
        for (int64_t k = threadIdx.x; k < nt; k += 256) {
            const int64_t jj = i + k - a.half;
            if (jj >= 0 && jj < a.n) acc += col[jj] * a.taps[k];
        }
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        const double r = (red[0] + red[1] + red[2] + red[3]) * a.dnu;
        __syncthreads();
        out += side == 0 ? r * (w != 0.0 ? 1.0 - w : 1.0) : r * w;
    }
    if (threadIdx.x == 0) a.K[(int64_t)c * a.n_ch + ch] = out * a.factor[ch];
}

hipError_t launch_jac_ils(const JacIlsArgs &a, int n_cols, hipStream_t s) {
    hipLaunchKernelGGL(jac_ils_kernel, dim3((unsigned)a.n_ch, n_cols), dim3(256), 0, s, a);
    return hipGetLastError();
}
