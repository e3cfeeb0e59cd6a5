// Per-(layer, species, line) line parameters and DPTMIN decisions, FP64.
//
// One thread per (item, line).  Restates, per line, the host loop of
// compute_voigt_cross_sections (src/CrossSections/voigt.jl:402-414,448-454),
// the DPTMIN test of _reject_weak_lines (voigt.jl:333-355), the VP_Y line
// weights of _append_dispersive_lines! (src/CrossSections/line_mixing.jl:1193-1220)
// and the line-shape formulas of src/HITRAN/broadening.jl:40-160 and
// src/HITRAN/partition.jl:10-41.  Results are packed into three float4
// records per line, in grid-step units, for the FP32 line-sum kernel (plus the
// optional mid-core records of midcore_coeffs):
//   rec0 = {centre index (int bits), centre fraction, g^2, A}
//   rec1 = {B, P, fu, y}
//   rec2 = {Sn, fu_d, y_d, YSn}
// with g = gamma_L/dnu, A = coef S g/(pi dnu) (Lorentz wing amplitude),
// P = coef Snorm H(x_cut, y) (pedestal), fu = dnu f, y = gamma_L f,
// Sn = coef S f/sqrt(pi); the _d fields and B = coef Y p S_T/(pi dnu),
// YSn = coef Y p S_T f_d/sqrt(pi) carry the dispersive term.
#include "irs_devmath.h"
#include "irs_internal.h"

#define TIPS_T_MIN 70
#define TIPS_NT 331
#define T_REF 296.0
#define C2_LINE 1.4387752   // broadening.jl:10
#define C2_LM 1.4388        // line_mixing.jl:31

__device__ __forceinline__ double tips_lookup(const double *v, double T) {
    long lo = (long)floor(T) - TIPS_T_MIN;
    long hi = lo + 1;
    lo = lo < 0 ? 0 : (lo > TIPS_NT - 1 ? TIPS_NT - 1 : lo);
    hi = hi < 0 ? 0 : (hi > TIPS_NT - 1 ? TIPS_NT - 1 : hi);
    const double frac = T - floor(T);
    return v[lo] * (1.0 - frac) + v[hi] * frac;
}

__device__ __forceinline__ double mol_mass(int mol) {   // broadening.jl:131-148
    switch (mol) {
        case 1: return 18.010;  case 2: return 44.010;  case 3: return 47.998;
        case 4: return 44.013;  case 5: return 28.010;  case 6: return 16.043;
        case 7: return 31.999;  case 8: return 34.015;  case 9: return 64.065;
        case 10: return 36.461; case 11: return 17.031;
        default: return 30.0;
    }
}

__device__ __forceinline__ double lm_mass(int mol, int iso) {   // line_mixing.jl:743-753
    if (mol == 2) {
        switch (iso) {
            case 1: return 43.98983; case 2: return 44.99318; case 3: return 45.99398;
            case 4: return 44.99403; case 5: return 45.99706; case 6: return 44.99706;
            case 7: return 47.99832; case 8: return 47.00134; case 9: return 48.00196;
            case 0: case 10: return 47.00427;
            default: return 44.0;
        }
    }
    if (mol == 6) {
        switch (iso) {
            case 1: return 16.031300; case 2: return 17.034655;
            case 3: return 17.037475; case 4: return 18.040830;
            default: return 44.0;
        }
    }
    return 44.0;
}

// Mid-core record of one line (three float4, see the CORE range of the line
// sum): coefficients of the asymptotic series of w(x + iy) for |z| >= IRS_Z1,
// regrouped in powers of v = 1/(x^2 + y^2) (Y = y^2):
//   Re w = (y/sqrt(pi)) v (1 + 3/2 v + (15/4 - 2Y) v^2 + (105/8 - 15Y) v^3
//          + (945/16 - 105Y + 12Y^2) v^4 + (210Y^2 - 1575Y/2) v^5 + (2835Y^2 - 120Y^3) v^6
//          - 3780Y^3 v^7 + 1680Y^4 v^8)
//   Im w = (x/sqrt(pi)) v (1 + 1/2 v + (3/4 - 2Y) v^2 + (15/8 - 9Y) v^3
//          + (105/16 - 45Y + 12Y^2) v^4 + (150Y^2 - 525Y/2) v^5 + (1575Y^2 - 120Y^3) v^6
//          - 2940Y^3 v^7 + 1680Y^4 v^8)
// (w ~ (i/sqrt(pi)) sum_{n<=4} (2n-1)!!/(2^n z^{2n+1}); Im z^{2n+1} / y and
// Re z^{2n+1} / x expanded with x^2 = 1/v - Y).  Layout:
//   c[0] = {k3, k4, k5, k6}, c[1] = {k7, k8, k9, Y}, c[2] = {amp, KO, KI, 0}
// with amp the factor in front of v (...): Sn y/sqrt(pi) for Re w, YSn/sqrt(pi)
// for the dispersive record (whose series is further multiplied by x), and
// the integer distances (grid steps) KO: |du| > KO implies |x| > x_far, KI:
// |du| > KI implies |z| >= IRS_Z1 (KI = -1 when y >= IRS_Z1).  x_far < 0 marks
// a line with no core (KO = KI = -1).
__device__ void midcore_coeffs(double y, double s, double amp, double x_far, bool disp, float4 *c) {
    const double Y = y * y;
    double k[10];
    if (!disp) {
        k[3] = 3.75 - 2.0 * Y;
        k[4] = 13.125 - 15.0 * Y;
        k[5] = 59.0625 + Y * (12.0 * Y - 105.0);
        k[6] = Y * (210.0 * Y - 787.5);
        k[7] = Y * Y * (2835.0 - 120.0 * Y);
    } else {
        k[3] = 0.75 - 2.0 * Y;
        k[4] = 1.875 - 9.0 * Y;
        k[5] = 6.5625 + Y * (12.0 * Y - 45.0);
        k[6] = Y * (150.0 * Y - 262.5);
        k[7] = Y * Y * (1575.0 - 120.0 * Y);
    }
    k[8] = (disp ? -2940.0 : -3780.0) * Y * Y * Y;
    k[9] = 1680.0 * Y * Y * Y * Y;
    const double lim = 1.0e9;
    int KO = -1, KI = -1;
    if (x_far >= 0.0) {
        KO = (int)fmin(ceil(x_far / s) + 1.0, lim);
        if (Y < IRS_Z1 * IRS_Z1) KI = (int)fmin(ceil(sqrt(IRS_Z1 * IRS_Z1 - Y) / s) + 1.0, lim);
    }
    c[0] = make_float4((float)k[3], (float)k[4], (float)k[5], (float)k[6]);
    c[1] = make_float4((float)k[7], (float)k[8], (float)k[9], (float)Y);
    c[2] = make_float4((float)amp, __int_as_float(KO), __int_as_float(KI), 0.0f);
}

__global__ void __launch_bounds__(256) prep_kernel(const PrepArgs a) {
    const int item = blockIdx.y;
    const ItemDev it = a.items[item];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    unsigned long long my_active = 0, my_pairs = 0, my_dpairs = 0;
    double ped_P = 0.0, ped_c = 0.0;           // the record's pedestal amplitude and centre (grid steps)
    if (it.active && j < it.n) {
        const LineListDev L = a.lists[it.list];
        const double T = it.T, p = it.p_atm, vs = it.vmr_self, coef = it.coef;
        const double nu = L.nu[j];
        const int mol = L.mol[j], iso = L.iso[j];
        const int row = L.tips_row[j];
        double qr;
        if (row >= 0) {
            const double *v = a.tips_q + (int64_t)row * TIPS_NT;
            qr = tips_lookup(v, T_REF) / tips_lookup(v, T);
        } else {
            qr = T_REF / T;
        }
        // temperature_scaled_intensity, broadening.jl:89-99
        const double stim = (1.0 - exp(-C2_LINE * nu / T)) / (1.0 - exp(-C2_LINE * nu / T_REF));
        const double boltz = exp(-C2_LINE * L.elower[j] * (1.0 / T - 1.0 / T_REF));
        const double S = L.S[j] * qr * boltz * stim;
        // pressure_broadened_width, broadening.jl:40-53
        const double n_air = (double)L.nair[j];
        const double n_self = mol == 1 ? 0.95 : (mol == 2 ? 0.75 : n_air);
        const double tr = T_REF / T;
        const double gl = pow(tr, n_air) * (double)L.gair[j] * (1.0 - vs) * p +
                          pow(tr, n_self) * (double)L.gself[j] * vs * p;
        double gd = nu * 3.58126e-7 * sqrt(T / mol_mass(mol));
        gd = fmax(gd, 1e-10);
        const double f = IRS_SQRT_LN2 / gd;
        const double y = gl * f;
        const double snorm = S * f * IRS_INV_SQRT_PI;
        // pressure_shift, broadening.jl:156-160
        const double dself = mol == 1 ? 0.002 : (mol == 2 ? -0.003 : 0.0);
        const double c = nu + ((double)L.dair[j] * (1.0 - vs) + dself * vs) * p;
        // DPTMIN: peak = Snorm * erfcx(y) (voigt.jl:340-350); skipped on the VP_Y path
        bool keep = true;
        if (it.dptmn > 0.0 && !it.lm) keep = (snorm * erfcx(y)) * coef > it.dptmn;
        // pedestal value at the cutoff (voigt.jl:448-454)
        const double xc = it.cutoff * f;
        double hc;
        if (fabs(xc) > it.x_far) hc = y * IRS_INV_SQRT_PI / (xc * xc + y * y);
        else { double wi; w64(xc, y, hc, wi); }
        const double u = (c - a.nu_lo) / a.dnu;
        const double cn = floor(u);
        const double g = gl / a.dnu;
        float4 r0, r1, r2;
        r0.x = __int_as_float((int)cn);
        r0.y = (float)(u - cn);
        if (keep) {
            r0.z = (float)(g * g);
            r0.w = (float)(coef * S * g / (IRS_PI * a.dnu));
            r1.y = (float)(coef * snorm * hc);
            r1.z = (float)(a.dnu * f);
            r1.w = (float)y;
            r2.x = (float)(coef * snorm);
        } else {
            r0.z = 1.0f; r0.w = 0.0f; r1.y = 0.0f; r1.z = 1.0f; r1.w = 1.0f; r2.x = 0.0f;
        }
        r1.x = 0.0f; r2.y = 1.0f; r2.z = 1.0f; r2.w = 0.0f;
        // VP_Y dispersive weights with the S-file conventions (line_mixing.jl:1193-1220)
        bool disp = false;
        double yd = 0.0, sd = 1.0, ysn = 0.0, fd_ = 0.0, ysn1 = 0.0;
        if (it.lm && L.ylm && L.ylm[j] != 0.0) {
            const double Y = L.ylm[j];
            const double md = lm_mass(mol, iso);
            const double gdd = 3.5812e-7 * nu * sqrt(T / md);
            const double gld = (double)L.gair[j] * pow(296.0 / T, n_air) * p;
            const double fd = IRS_SQRT_LN2 / fmax(gdd, 1e-10);
            const double st0 = 1.0 - exp(-C2_LM * nu / 296.0);
            const double stt = 1.0 - exp(-C2_LM * nu / T);
            const double st = L.S[j] * qr * exp(-C2_LM * L.elower[j] * (1.0 / T - 1.0 / 296.0)) * stt / st0;
            if (st > 0.0) {
                r1.x = (float)(coef * Y * p * st / (IRS_PI * a.dnu));
                r2.y = (float)(a.dnu * fd);
                r2.z = (float)(gld * fd);
                r2.w = (float)(coef * Y * p * st * fd * IRS_INV_SQRT_PI);
                disp = true;
                fd_ = fd;
                ysn1 = Y * p * st * fd * IRS_INV_SQRT_PI;
                yd = gld * fd;
                sd = a.dnu * fd;
                ysn = coef * Y * p * st * fd * IRS_INV_SQRT_PI;
            }
        }
        const int64_t o = it.rec_off + j;
        a.rec0[o] = r0;
        a.rec1[o] = r1;
        a.rec2[o] = r2;
        if (a.rec64) {
            double *q = a.rec64 + o * 8;
            q[0] = c;
            q[1] = f;
            q[2] = y;
            q[3] = keep ? snorm : 0.0;
            q[4] = hc;
            q[5] = disp ? ysn1 : 0.0;             // per molecule (the FP64 sum multiplies by coef)
            q[6] = disp ? fd_ : 0.0;
            q[7] = disp ? yd : 0.0;
        }
        ped_P = (double)r1.y;
        ped_c = (double)__float_as_int(r0.x) + (double)r0.y;
        if (a.rec3) {
            float4 *c = a.rec3 + 6 * o;
            if (keep) midcore_coeffs(y, a.dnu * f, coef * snorm * y * IRS_INV_SQRT_PI, it.x_far, false, c);
            else midcore_coeffs(0.0, 1.0, 0.0, -1.0, false, c);
            if (it.lm) {
                if (disp) midcore_coeffs(yd, sd, ysn * IRS_INV_SQRT_PI, it.x_far, true, c + 3);
                else midcore_coeffs(0.0, 1.0, 0.0, -1.0, true, c + 3);
            }
        }
        if (keep) {
            my_active = 1;
            const double lo = ceil((c - it.cutoff - a.nu_lo) / a.dnu);
            const double hi = floor((c + it.cutoff - a.nu_lo) / a.dnu);
            const double l2 = fmax(lo, 0.0), h2 = fmin(hi, (double)(a.n_hi - 1));
            if (h2 >= l2) my_pairs = (unsigned long long)(h2 - l2 + 1.0);
        }
        if (disp) {
            const double lc = (double)it.Clm * a.dnu;
            const double lo = ceil((c - lc - a.nu_lo) / a.dnu);
            const double hi = floor((c + lc - a.nu_lo) / a.dnu);
            const double l2 = fmax(lo, 0.0), h2 = fmin(hi, (double)(a.n_hi - 1));
            if (h2 >= l2) my_dpairs = (unsigned long long)(h2 - l2 + 1.0);
        }
    }
    // pedestal prefix sums, first pass: inclusive scan of P, P c, P c^2 over the
    // block's 256 lines (the record values, so the moments match the records);
    // block totals go to pblk and are carried across blocks by the offsets pass
    if (a.pre) {
        __shared__ double wtot[3][4];
        double s0 = ped_P, s1 = ped_P * ped_c, s2 = ped_P * ped_c * ped_c;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (int o = 1; o < 64; o <<= 1) {
            const double w0 = __shfl_up(s0, o, 64), w1 = __shfl_up(s1, o, 64), w2 = __shfl_up(s2, o, 64);
            if (lane >= o) { s0 += w0; s1 += w1; s2 += w2; }
        }
        if (lane == 63) { wtot[0][wv] = s0; wtot[1][wv] = s1; wtot[2][wv] = s2; }
        __syncthreads();
        for (int w = 0; w < wv; ++w) { s0 += wtot[0][w]; s1 += wtot[1][w]; s2 += wtot[2][w]; }
        if (it.active && j < it.n) {
            double *p0 = a.pre + it.rec_off + j;
            p0[0] = s0;
            p0[a.n_rec] = s1;
            p0[2 * a.n_rec] = s2;
        }
        const int64_t nb = it.active ? ((int64_t)it.n + 255) / 256 : 0;
        if (threadIdx.x < 3 && (int64_t)blockIdx.x < nb)
            a.pblk[(it.pb_off + blockIdx.x) * 3 + threadIdx.x] =
                wtot[threadIdx.x][0] + wtot[threadIdx.x][1] + wtot[threadIdx.x][2] + wtot[threadIdx.x][3];
    }
    // block reduction of the work counters, one atomic per counter and block
    __shared__ unsigned long long red[3][4];
    for (int o = 32; o > 0; o >>= 1) {
        my_active += __shfl_xor(my_active, o, 64);
        my_pairs += __shfl_xor(my_pairs, o, 64);
        my_dpairs += __shfl_xor(my_dpairs, o, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = my_active;
        red[1][wv] = my_pairs;
        red[2][wv] = my_dpairs;
    }
    __syncthreads();
    if (threadIdx.x < 3 && a.counters) {
        const unsigned long long v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                                     red[threadIdx.x][3];
        const int slot = (int)((blockIdx.x + blockIdx.y) % IRS_CNT_SLOTS);
        if (v) atomicAdd(&a.counters[8 * slot + threadIdx.x], v);
    }
}

// floor((nu_j - nu_lo)/dnu) for one list: the sorted search key of the line sum.
__global__ void __launch_bounds__(256) keys_kernel(const double *nu, int64_t n, double nu_lo, double dnu,
                                                   int32_t *keys) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) {
        double k = floor((nu[j] - nu_lo) / dnu);
        k = fmin(fmax(k, -2.0e9), 2.0e9);
        keys[j] = (int32_t)k;
    }
}

// Pedestal prefix sums, second and third passes.  prep_kernel leaves in pre
// the inclusive scan of P, P c, P c^2 (c the centre in grid steps, FP64) within
// each 256-line block of an item and the block totals in pblk; the offsets pass
// turns an item's block totals into exclusive prefix sums (one workgroup per
// item, 256 blocks per step), and the fix-up pass adds each block's offset, so
// that pre holds per-item inclusive prefix sums over the item's lines: a tile
// obtains the three pedestal moments of any line range by two lookups.
__global__ void __launch_bounds__(256) pedestal_offsets_kernel(const PrepArgs a) {
    __shared__ double wsum[3][4];
    const ItemDev it = a.items[blockIdx.x];
    if (!it.active || it.n == 0) return;
    const int64_t nb = ((int64_t)it.n + 255) / 256;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double *pb = a.pblk + it.pb_off * 3;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int64_t b0 = 0; b0 < nb; b0 += 256) {
        const int64_t b = b0 + t;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (b < nb) { v0 = pb[3 * b]; v1 = pb[3 * b + 1]; v2 = pb[3 * b + 2]; }
        double u0 = v0, u1 = v1, u2 = v2;
        for (int o = 1; o < 64; o <<= 1) {
            const double w0 = __shfl_up(u0, o, 64), w1 = __shfl_up(u1, o, 64), w2 = __shfl_up(u2, o, 64);
            if (lane >= o) { u0 += w0; u1 += w1; u2 += w2; }
        }
        if (lane == 63) { wsum[0][wv] = u0; wsum[1][wv] = u1; wsum[2][wv] = u2; }
        __syncthreads();
        double e0 = c0 + (u0 - v0), e1 = c1 + (u1 - v1), e2 = c2 + (u2 - v2);   // exclusive
        for (int w = 0; w < wv; ++w) { e0 += wsum[0][w]; e1 += wsum[1][w]; e2 += wsum[2][w]; }
        if (b < nb) { pb[3 * b] = e0; pb[3 * b + 1] = e1; pb[3 * b + 2] = e2; }
        c0 += wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
        c1 += wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
        c2 += wsum[2][0] + wsum[2][1] + wsum[2][2] + wsum[2][3];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) pedestal_fixup_kernel(const PrepArgs a) {
    const ItemDev it = a.items[blockIdx.y];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (!it.active || blockIdx.x == 0 || j >= it.n) return;      // block 0 has offset 0
    const double *pb = a.pblk + (it.pb_off + blockIdx.x) * 3;
    double *p0 = a.pre + it.rec_off + j;
    p0[0] += pb[0];
    p0[a.n_rec] += pb[1];
    p0[2 * a.n_rec] += pb[2];
}

hipError_t launch_prep(const PrepArgs &a, int max_lines, hipStream_t s) {
    if (a.n_items <= 0 || max_lines <= 0) return hipSuccess;
    if (a.pre && !a.pblk) return hipErrorInvalidValue;
    dim3 grid((max_lines + 255) / 256, a.n_items);
    hipLaunchKernelGGL(prep_kernel, grid, dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.pre) return e;
    hipLaunchKernelGGL(pedestal_offsets_kernel, dim3(a.n_items), dim3(256), 0, s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (max_lines > 256) hipLaunchKernelGGL(pedestal_fixup_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// Key table of one list: out[t] = first line with key >= lo + t (lower bound in
// the sorted keys), so that the line sum replaces its window searches by one
// load per bound.
__global__ void __launch_bounds__(256) keytab_kernel(const int32_t *keys, int n, int64_t lo, int64_t count,
                                                     int32_t *out) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    const int64_t v = lo + t;
    int a = 0, b = n;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if ((int64_t)keys[mid] < v) a = mid + 1; else b = mid;
    }
    out[t] = a;
}

hipError_t launch_keytab(const int32_t *keys, int n, int64_t lo, int64_t count, int32_t *out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(keytab_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, keys, n, lo, count,
                       out);
    return hipGetLastError();
}

hipError_t launch_keys(const double *nu, int64_t n, double nu_lo, double dnu, int32_t *keys,
                       hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nu, n, nu_lo,
                       dnu, keys);
    return hipGetLastError();
}

// w(x + iy) with the FP32 routines of the line sums, one thread per point:
// mode 0 is the rule of the CORE range (w32_core inside |z| < 8, else w32),
// mode 1 the rule of the per-pair ranges and the gradient sum (w32).
__global__ void __launch_bounds__(256) faddeeva32_kernel(const double *x, const double *y, int64_t n, int mode,
                                                         double *re, double *im) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float xf = (float)x[i], yf = (float)y[i];
    float wr, wi;
    if (mode == 0 && fmaf(xf, xf, yf * yf) < (float)(IRS_Z1 * IRS_Z1)) w32_core(xf, yf, wr, wi);
    else w32(xf, yf, wr, wi);
    re[i] = (double)wr;
    im[i] = (double)wi;
}

hipError_t launch_faddeeva32(const double *x, const double *y, int64_t n, int mode, double *re, double *im,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(faddeeva32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n, mode, re, im);
    return hipGetLastError();
}
