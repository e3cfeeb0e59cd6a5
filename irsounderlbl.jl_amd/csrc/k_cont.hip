// Continuum optical depth per layer on the internal grid (FP64), added to the
// FP32 layer optical depths of the line sum.
//
// Restates src/CrossSections/continuum.jl: _radfn (:150-155), h2o_continuum
// (:163-190), co2_continuum (:209-232), co2_cia / n2_cia / o2_cia (:249-327)
// with _cia_sigma (:336-363), each times dz (src/CrossSections/
// optical_depth.jl:123-130) at the layer's mid-level VMR, p and T as
// src/Sensor/sounder.jl:310-322 does.  The Jacobian variant also writes the
// derivatives of tau_c with respect to T_mid and the H2O / CO2 mid-layer VMR,
// chained to the two bounding levels by the mid-layer interpolation weight
// (the terms src/Estimation/vmr_jacobian.jl:313-349,451-454,492-495 obtain by
// central differences; here analytic).
//
// One thread per grid point: the table interpolation weights depend on nu only
// and are formed once; the layer loop reads and writes tau coalesced.
#include "irs_cont.h"
#include "irs_internal.h"

#define CK_KB   1.380649e-23
#define CK_C2   1.4387769
#define CK_PREF 1013.0
#define CK_TREF 296.0
#define CK_TEFF 246.0

namespace {

// linear interpolant of one CIA block at nu (zero outside its knots)
__device__ double cia_block(const CiaDev &c, int b, double nu) {
    const int64_t o = c.off[b], m = c.off[b + 1] - o;
    const double *x = c.nu + o, *y = c.sig + o;
    if (m < 1 || nu < x[0] || nu > x[m - 1]) return 0.0;
    if (m == 1) return y[0];
    int64_t lo = 0, hi = m - 1;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (x[mid] <= nu) lo = mid; else hi = mid;
    }
    const double w = (nu - x[lo]) / (x[hi] - x[lo]);
    return y[lo] * (1.0 - w) + y[hi] * w;
}

// _cia_sigma for the layer temperature T: the applicable blocks (those whose
// knot range holds nu) bracketing T, linear in T between them, clamped to the
// nearest block outside.  dsdT is the slope in T (0 when clamped).
__device__ double cia_sigma(const CiaDev &c, double nu, double T, double &dsdT) {
    double T_lo = -1e300, T_hi = 1e300;
    int b_lo = -1, b_hi = -1;
    for (int b = 0; b < c.nb; ++b) {
        if (nu < c.numin[b] || nu > c.numax[b]) continue;
        const double Tb = c.T[b];
        if (Tb <= T && Tb > T_lo) { T_lo = Tb; b_lo = b; }
        if (Tb >= T && Tb < T_hi) { T_hi = Tb; b_hi = b; }
    }
    dsdT = 0.0;
    if (b_lo < 0 && b_hi < 0) return 0.0;
    if (b_lo < 0) return cia_block(c, b_hi, nu);
    if (b_hi < 0) return cia_block(c, b_lo, nu);
    const double s_lo = cia_block(c, b_lo, nu);
    if (T_lo == T_hi) return s_lo;
    const double s_hi = cia_block(c, b_hi, nu);
    const double w = (T - T_lo) / (T_hi - T_lo);
    dsdT = (s_hi - s_lo) / (T_hi - T_lo);
    return (1.0 - w) * s_lo + w * s_hi;
}

__global__ void __launch_bounds__(256) continuum_kernel(const ContArgs a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int q = blockIdx.y;
    const double nu = (a.an + (double)i * a.sn) / a.den;
    // MT-CKD H2O table on nu = 0, 10, .. (Flat beyond the ends)
    double Cs = 0, Cf = 0, tx = 0;
    if (a.continua & IRS_CONT_H2O) {
        const int N = 2001;
        const double *t = a.h2o;
        if (nu <= t[0]) { Cs = t[1]; Cf = t[2]; tx = t[3]; }
        else if (nu >= t[4 * (N - 1)]) { Cs = t[4 * (N - 1) + 1]; Cf = t[4 * (N - 1) + 2]; tx = t[4 * (N - 1) + 3]; }
        else {
            int lo = min(max((int)floor(nu / 10.0), 0), N - 2);
            const double *r0 = t + 4 * lo, *r1 = r0 + 4;
            const double w = (nu - r0[0]) / (r1[0] - r0[0]);
            Cs = r0[1] * (1.0 - w) + r1[1] * w;
            Cf = r0[2] * (1.0 - w) + r1[2] * w;
            tx = r0[3] * (1.0 - w) + r1[3] * w;
        }
    }
    // MT-CKD CO2 table on nu = -4, -2, ..: S*xfac and the bandhead exponent
    double Sx = 0, td = 0;
    if (a.continua & IRS_CONT_CO2) {
        const int N = 5003;
        const double *t = a.co2;
        if (nu <= -4.0) { Sx = t[0]; td = t[1]; }
        else if (nu >= -4.0 + 2.0 * (N - 1)) { Sx = t[2 * (N - 1)]; td = t[2 * (N - 1) + 1]; }
        else {
            int lo = min(max((int)floor((nu + 4.0) / 2.0), 0), N - 2);
            const double x0 = -4.0 + 2.0 * lo, x1 = x0 + 2.0;
            const double w = (nu - x0) / (x1 - x0);
            Sx = t[2 * lo] * (1.0 - w) + t[2 * lo + 2] * w;
            td = t[2 * lo + 1] * (1.0 - w) + t[2 * lo + 3] * w;
        }
    }
    const bool jac = a.dtlo || a.vlo_h || a.vlo_c;
    for (int k = 0; k < a.n_layers; ++k) {
        const ContLayer &cl = a.lay[(size_t)q * a.n_layers + k];
        const double T = cl.T_mid, p = cl.p_mid, vh = cl.vh2o, vc = cl.vco2;
        const double x = CK_C2 * nu / T;
        double rad, drad;                                   // _radfn and d/dT
        if (x <= 0.01) { rad = 0.5 * x * nu; drad = -rad / T; }
        else if (x <= 10.0) {
            const double e = exp(-x);
            rad = nu * (1.0 - e) / (1.0 + e);
            drad = -nu * 2.0 * e / ((1.0 + e) * (1.0 + e)) * x / T;
        } else { rad = nu; drad = 0.0; }
        const double rho = (p / CK_PREF) * (CK_TREF / T);
        const double A = p * 100.0 / (CK_KB * T) * 1e-6;    // molec/cm^3 per unit VMR
        double k_sum = 0.0, dk_dT = 0.0, dk_dvh = 0.0, dk_dvc = 0.0;
        if (a.continua & IRS_CONT_H2O) {
            const double n_h2o = vh * A;
            const double ks = Cs * pow(CK_TREF / T, tx) * vh * rho * rad * n_h2o;
            const double kf = Cf * (1.0 - vh) * rho * rad * n_h2o;
            k_sum += ks + kf;
            if (jac) {
                const double rr = drad / rad;
                dk_dT += ks * (rr - (tx + 2.0) / T) + kf * (rr - 2.0 / T);
                const double base = rho * rad * A;
                dk_dvh += Cs * pow(CK_TREF / T, tx) * base * 2.0 * vh + Cf * base * (1.0 - 2.0 * vh);
            }
        }
        if (a.continua & IRS_CONT_CO2) {
            const double n_co2 = vc * A;
            const double pref = 1e-20 * n_co2 * rho;
            const double tcor = td == 0.0 ? 1.0 : pow(T / CK_TEFF, td);
            const double kc = Sx * tcor * pref * rad;
            k_sum += kc;
            if (jac) {
                dk_dT += Sx * tcor * pref * drad + kc * ((td == 0.0 ? 0.0 : td / T) - 2.0 / T);
                dk_dvc += Sx * tcor * 1e-20 * A * rho * rad;
            }
        }
        if (a.continua & (IRS_CONT_CO2CIA | IRS_CONT_N2 | IRS_CONT_O2)) {
            for (int pr = 0; pr < 3; ++pr) {
                const int bit = pr == 0 ? IRS_CONT_CO2CIA : (pr == 1 ? IRS_CONT_N2 : IRS_CONT_O2);
                if (!(a.continua & bit) || a.cia[pr].nb == 0) continue;
                double ds;
                const double s = cia_sigma(a.cia[pr], nu, T, ds);
                if (pr == 0) {
                    const double nn = (vc * A) * A;             // n_CO2 n_air
                    k_sum += s * nn;
                    if (jac) { dk_dT += ds * nn - 2.0 * s * nn / T; dk_dvc += s * A * A; }
                } else {
                    const double f = pr == 1 ? 0.78084 : 0.20946;   // d vmr_x / d(1 - vmr_H2O)
                    const double nx = (pr == 1 ? cl.vn2 : cl.vo2) * A, nn = nx * nx;
                    k_sum += s * nn;
                    if (jac) { dk_dT += ds * nn - 2.0 * s * nn / T; dk_dvh += -2.0 * s * nx * f * A; }
                }
            }
        }
        const double tc = k_sum * cl.dz;
        const size_t o = ((size_t)q * a.n_layers + k) * (size_t)a.n + (size_t)i;
        if (a.tau64) a.tau64[o] = a.add64 ? a.tau64[o] + tc : tc;
        else if (a.tau) a.tau[o] += (float)tc;
        if (jac && q == 0) {
            const size_t oj = (size_t)k * (size_t)a.n + (size_t)i;
            const double wM = cl.wM;
            const double dT = dk_dT * cl.dz + tc / T;         // dz is proportional to T
            if (a.dtlo) {
                a.dtlo[oj] += (float)(dT * (1.0 - wM));
                a.dthi[oj] += (float)(dT * wM);
            }
            if (a.vlo_h) {
                a.vlo_h[oj] += (float)(dk_dvh * cl.dz * (1.0 - wM));
                a.vhi_h[oj] += (float)(dk_dvh * cl.dz * wM);
            }
            if (a.vlo_c) {
                a.vlo_c[oj] += (float)(dk_dvc * cl.dz * (1.0 - wM));
                a.vhi_c[oj] += (float)(dk_dvc * cl.dz * wM);
            }
        }
    }
}

}  // namespace

hipError_t launch_continuum(const ContArgs &a, hipStream_t s) {
    if (a.n <= 0 || a.n_layers <= 0 || !a.continua) return hipSuccess;
    dim3 grid((unsigned)((a.n + 255) / 256), (unsigned)std::max(a.n_prof, 1));
    hipLaunchKernelGGL(continuum_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}
