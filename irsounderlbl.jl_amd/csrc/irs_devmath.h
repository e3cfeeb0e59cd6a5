// Device math: Faddeeva w(z) in FP32 (line-sum core) and FP64 (line
// parameters), Planck in FP64.
//
// w(z) = exp(-z^2) erfc(-iz) = (i/pi) Int exp(-t^2)/(z - t) dt, Im z >= 0.
// Two published methods, chosen per region of s = |x| + y:
//  * s < S_TRAP: trapezoidal rule with the exact pole correction
//    (Chiarella & Reichel 1968; Matta & Reichel 1971) on nodes shifted so
//    that x lies midway between nodes (Hunter & Regan 1972).  With that shift
//    the pole term is 2 exp(-z^2)/(1 + exp(2 pi y/h)), so no term blows up
//    and Re w keeps full relative accuracy near the real axis.
//  * elsewhere: the Laplace continued fraction
//    w = (i/sqrt(pi)) / (z - (1/2)/(z - 1/(z - (3/2)/(z - ...)))).
// FP32: h = 0.7, nodes |t| <= 4.2, S_TRAP = 5.5, fraction depth 6/4/3/2 for
// s < 8/12/20/else.  Measured on the device against scipy.special.wofz over
// |x| <= 30, 1e-6 <= y <= 30 (relative error of Re w; of Im w against
// max(|Im w|, 1e-2 |w|)): w32 3.9e-6 / 5.4e-6 (largest at the s = 5.5 switch),
// w32_core inside |z| < 8 1.5e-6 / 2.1e-5.
// FP64: h = 0.5, |t| <= 6.1, S_TRAP = 8, depth 24: ~1e-14 relative.
#pragma once

#include <hip/hip_runtime.h>

#include "irs_internal.h"

__device__ __forceinline__ float irs_rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// |du| <= C for du = dI - frac (dI the exact integer offset of the grid point
// from the line's centre index, frac the centre's fraction) and C = Ci + Cf:
// decided on the exact integer parts and the FP32 fractions, so the window
// boundary is resolved to ~1e-7 grid steps (an FP32 |du| near a 25000-step
// cutoff resolves only 2e-3).  Used for the VP_Y dispersive window, whose
// term does not vanish at the boundary.
__device__ __forceinline__ bool in_window(int dI, float frac, int Ci, float Cf) {
    return (float)(dI - Ci) <= frac + Cf && (float)(dI + Ci) >= frac - Cf;
}

// Trapezoid sum over the nodes k = klo..khi (offsets off_k = (k + 1/2) h from
// x) of exp(-(x - off_k)^2) / (off_k^2 + y^2).  The Gaussian weights come from
// the node nearest x outward by the exact two-term recurrence
//   c_{k+1} = c_k rho_k,  rho_{k+1} = rho_k exp(-2 h^2),  rho_k = exp(2 h (x - off_k) - h^2)
// (and its mirror downward), so one evaluation takes three exponentials
// instead of one per node; starting at the largest weight keeps the rounding
// of the dominant terms at that of a direct evaluation.  The pole term uses the
// hardware sine and cosine: it is below 1e-7 of w wherever 2xy exceeds a few
// radians (there exp(y^2 - x^2) / (1 + exp(2 pi y / h)) is negligible).
__device__ __forceinline__ void w32_trap(float x, float y, float &re, float &im) {
    const float h = 0.7f, inv_h = 1.0f / 0.7f, tmax = 4.2f;
    const int klo = (int)ceilf((x - tmax) * inv_h - 0.5f);
    const int khi = (int)floorf((x + tmax) * inv_h - 0.5f);
    const float y2 = y * y;
    float sr = 0.0f, si = 0.0f;
    const int kc = min(max((int)rintf(x * inv_h - 0.5f), klo), khi);
    const float ttc = x - ((float)kc + 0.5f) * h;
    const float cc = __expf(-ttc * ttc);
    const float g = 0.37531109885139957f;             // exp(-2 h^2)
    float c = cc, rho = __expf(2.0f * h * ttc - h * h);
    for (int k = kc; k <= khi; ++k) {
        const float off = ((float)k + 0.5f) * h;
        const float r = irs_rcpf(fmaf(off, off, y2));
        sr = fmaf(c, r, sr);
        si = fmaf(c * off, r, si);
        c *= rho;
        rho *= g;
    }
    c = cc;
    rho = __expf(-2.0f * h * ttc - h * h);
    for (int k = kc - 1; k >= klo; --k) {
        c *= rho;
        rho *= g;
        const float off = ((float)k + 0.5f) * h;
        const float r = irs_rcpf(fmaf(off, off, y2));
        sr = fmaf(c, r, sr);
        si = fmaf(c * off, r, si);
    }
    const float hp = 0.7f / 3.14159265358979f;
    sr *= y * hp;
    si *= hp;
    const float e = 2.0f * __expf(y2 - x * x) / (1.0f + __expf((2.0f * 3.14159265358979f / 0.7f) * y));
    const float s2 = __sinf(2.0f * x * y), c2 = __cosf(2.0f * x * y);
    re = fmaf(e, c2, sr);
    im = fmaf(-e, s2, si);
}

__device__ __forceinline__ void w32_cf(float x, float y, float &re, float &im) {
    const float s = x + y;
    const int depth = s < 8.0f ? 6 : (s < 12.0f ? 4 : (s < 20.0f ? 3 : 2));
    float rr = 0.0f, ri = 0.0f;
    for (int k = depth; k >= 1; --k) {
        const float dr = x - rr, di = y - ri;
        const float q = (0.5f * (float)k) * irs_rcpf(fmaf(dr, dr, di * di));
        rr = dr * q;
        ri = -di * q;
    }
    const float dr = x - rr, di = y - ri;
    const float q = 0.56418958354775628f * irs_rcpf(fmaf(dr, dr, di * di));
    re = di * q;
    im = dr * q;
}

// w(x + iy) for y >= 0, any x.
__device__ __forceinline__ void w32(float x, float y, float &re, float &im) {
    const float ax = fabsf(x);
    float r, i;
    if (ax + y < 5.5f) w32_trap(ax, y, r, i);
    else w32_cf(ax, y, r, i);
    re = r;
    im = x < 0.0f ? -i : i;
}

// w(x + iy) for x^2 + y^2 < 64 (the inner core of the line sum): the
// trapezoid rule for y < 6 and the continued fraction (depth 6/4 by s) for
// y >= 6.  Within that disc the float32 model of each branch stays within
// 1.5e-6 (trapezoid, Re w) and 2.6e-7 (fraction) relative of w(z); the branch
// depends on y only, which is one value per line, so the lanes of a wave that
// evaluate one line's inner core do not split between the two methods.
__device__ __forceinline__ void w32_core(float x, float y, float &re, float &im) {
    const float ax = fabsf(x);
    float r, i;
    if (y < 6.0f) w32_trap(ax, y, r, i);
    else w32_cf(ax, y, r, i);
    re = r;
    im = x < 0.0f ? -i : i;
}

__device__ inline void w64(double x, double y, double &re, double &im) {
    const double ax = fabs(x);
    double r, i;
    if (ax + y < 8.0) {
        const double h = 0.5, tmax = 6.1;
        const long klo = (long)ceil((ax - tmax) / h - 0.5);
        const long khi = (long)floor((ax + tmax) / h - 0.5);
        double sr = 0.0, si = 0.0;
        const double y2 = y * y;
        for (long k = klo; k <= khi; ++k) {
            const double off = ((double)k + 0.5) * h;
            const double tt = ax - off;
            const double c = exp(-tt * tt);
            const double d = off * off + y2;
            sr += c / d;
            si += c * off / d;
        }
        sr *= y * h / IRS_PI;
        si *= h / IRS_PI;
        const double e = 2.0 * exp(y2 - ax * ax) / (1.0 + exp(2.0 * IRS_PI * y / h));
        r = sr + e * cos(2.0 * ax * y);
        i = si - e * sin(2.0 * ax * y);
    } else {
        double rr = 0.0, ri = 0.0;
        for (int k = 24; k >= 1; --k) {
            const double dr = ax - rr, di = y - ri;
            const double q = 0.5 * k / (dr * dr + di * di);
            rr = dr * q;
            ri = -di * q;
        }
        const double dr = ax - rr, di = y - ri;
        const double q = IRS_INV_SQRT_PI / (dr * dr + di * di);
        r = di * q;
        i = dr * q;
    }
    re = r;
    im = x < 0.0 ? -i : i;
}

// Planck radiance, mW m^-2 sr^-1 (cm^-1)^-1 (src/Solver/planck.jl:11-24)
#define IRS_C1_PL 1.191042953e-5
#define IRS_C2_PL 1.4387769
__device__ __forceinline__ double planck64(double nu, double T) {
    return IRS_C1_PL * nu * nu * nu / expm1(IRS_C2_PL * nu / T);
}
