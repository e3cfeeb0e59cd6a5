/*
 * refcpu — FP64 CPU restatement of the reference forward model (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product never uses it.
 *
 * Parallel structure mirrors the reference: OpenMP over grid points inside
 * each (layer, species) cross section (the KernelAbstractions CPU backend in
 * src/CrossSections/voigt.jl:455-457), a sequential layer x species loop
 * (src/Sensor/sounder.jl:283-307), and a serial ILS loop
 * (src/Sensor/ils.jl:142-151).
 */
#include "refcpu.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../irsounderlbl.jl_amd/data/tips2024_table.h"
#include "../irsounderlbl.jl_amd/data/mtckd_table.h"

#define SQRT_LN2    0.83255461115769775635   /* sqrt(log(2.0))  voigt.jl:4 */
#define INV_SQRT_PI 0.56418958354775628695   /* inv(sqrt(pi))   voigt.jl:5 */
#define PI_D        3.14159265358979323846

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ========================================================================== *
 * Faddeeva function w(z) = exp(-z^2) erfc(-iz), Im z >= 0.
 *
 * The reference calls SpecialFunctions.erfcx(::Complex) (voigt.jl:158-160,
 * line_mixing.jl:1182), i.e. S. G. Johnson's Faddeeva package, which is not
 * vendored.  This restatement uses two published methods of its own:
 *  - |x|+y < 8: trapezoidal rule for (i/pi) Int exp(-t^2)/(z-t) dt with the
 *    exact pole correction (Chiarella & Reichel 1968; Matta & Reichel 1971),
 *    on nodes shifted so that x sits midway between two nodes (Hunter &
 *    Regan 1972).  With that shift the pole term reduces to
 *    2 exp(-z^2) / (1 + exp(2 pi y / h)) and no cancellation occurs.
 *    h = 0.5, nodes |t| <= 6.1: truncation ~exp(-pi^2/h^2) ~ 1e-17.
 *  - elsewhere: Laplace continued fraction
 *    w = (i/sqrt(pi)) / (z - (1/2)/(z - 1/(z - (3/2)/(z - ...)))), depth 24.
 * Checked against scipy.special.wofz and mpmath (tests/golden/faddeeva_*).
 * ========================================================================== */
static void w_trapezoid(double x, double y, double *re, double *im) {
    const double h = 0.5, tmax = 6.1;
    long klo = (long)ceil((x - tmax) / h - 0.5);
    long khi = (long)floor((x + tmax) / h - 0.5);
    double sr = 0.0, si = 0.0;
    const double y2 = y * y;
    for (long k = klo; k <= khi; ++k) {
        double off = ((double)k + 0.5) * h;     /* x - t_k */
        double t = x - off;
        double c = exp(-t * t);
        double d = off * off + y2;
        sr += c / d;
        si += c * off / d;
    }
    sr *= y * h / PI_D;
    si *= h / PI_D;
    double e = 2.0 * exp(y2 - x * x) / (1.0 + exp(2.0 * PI_D * y / h));
    *re = sr + e * cos(2.0 * x * y);
    *im = si - e * sin(2.0 * x * y);
}

static void w_contfrac(double x, double y, double *re, double *im) {
    double complex z = x + I * y;
    double complex r = 0.0;
    for (int k = 24; k >= 1; --k) r = (0.5 * k) / (z - r);
    double complex w = (I * INV_SQRT_PI) / (z - r);
    *re = creal(w);
    *im = cimag(w);
}

void orc_faddeeva(double x, double y, double *re, double *im) {
    double ax = fabs(x);
    double r, i;
    if (ax + y < 8.0) w_trapezoid(ax, y, &r, &i);
    else w_contfrac(ax, y, &r, &i);
    *re = r;
    *im = (x < 0.0) ? -i : i;
}

double orc_erfcx(double y) {
    double r, i;
    orc_faddeeva(0.0, y, &r, &i);
    return r;
}

/* Re w(x+iy) = Re erfcx(y - ix)  (voigt.jl:158-160) */
static inline double faddeeva_voigt(double x, double y) {
    double r, i;
    orc_faddeeva(x, y, &r, &i);
    return r;
}

/* voigt.jl:170-172 */
static inline double voigt_H(double x, double y, double x_far) {
    return fabs(x) > x_far ? y * INV_SQRT_PI / (x * x + y * y) : faddeeva_voigt(x, y);
}

/* ========================================================================== *
 * Partition sums (src/HITRAN/partition.jl:6-66)
 * ========================================================================== */
#define T_REF 296.0

static const double *tips_row(int mol, int iso) {
    for (int r = 0; r < IRS_TIPS_NISO; ++r)
        if (irs_tips_keys[r][0] == mol && irs_tips_keys[r][1] == iso) return irs_tips_q[r];
    return NULL;
}

/* partition.jl:10-18 (1-based indices restated 0-based) */
static double tips_lookup(const double *v, double T) {
    long i_lo = (long)floor(T) - IRS_TIPS_T_MIN;          /* 0-based */
    long i_hi = i_lo + 1;
    long n = IRS_TIPS_NT;
    if (i_lo < 0) i_lo = 0;
    if (i_lo > n - 1) i_lo = n - 1;
    if (i_hi < 0) i_hi = 0;
    if (i_hi > n - 1) i_hi = n - 1;
    double frac = T - floor(T);
    return v[i_lo] * (1.0 - frac) + v[i_hi] * frac;
}

/* partition.jl:20-26 */
double orc_partition(int mol, int iso, double T) {
    const double *v = tips_row(mol, iso);
    if (!v) return T / T_REF;
    return tips_lookup(v, T);
}

/* partition.jl:35-41 */
double orc_q_ratio(int mol, int iso, double T) {
    const double *v = tips_row(mol, iso);
    if (!v) return T_REF / T;
    double q_t = tips_lookup(v, T);
    double q_ref = tips_lookup(v, T_REF);
    return q_ref / q_t;
}

/* partition.jl:59-66 */
double orc_partition_deriv(int mol, int iso, double T) {
    const double *v = tips_row(mol, iso);
    if (!v) return 1.0 / T_REF;
    long i_lo = (long)floor(T) - IRS_TIPS_T_MIN;
    long i_hi = i_lo + 1;
    if (i_lo < 0 || i_hi > IRS_TIPS_NT - 1) return 0.0;
    return v[i_hi] - v[i_lo];
}

/* ========================================================================== *
 * Line shape parameters (src/HITRAN/broadening.jl)
 * ========================================================================== */
#define C2_LINE 1.4387752   /* broadening.jl:10 */

static const double MOL_MASS_AMU[11] = {   /* broadening.jl:131-143 */
    18.010, 44.010, 47.998, 44.013, 28.010, 16.043, 31.999, 34.015, 64.065, 36.461, 17.031};

static double mol_mass(int mol) {          /* broadening.jl:145-148 */
    if (mol < 1 || mol > 11) return 30.0;
    return MOL_MASS_AMU[mol - 1];
}

static double n_self_of(int mol, double n_air) {   /* broadening.jl:17-20 */
    if (mol == 1) return 0.95;
    if (mol == 2) return 0.75;
    return n_air;
}

static double delta_self_of(int mol) {             /* broadening.jl:21-24 */
    if (mol == 1) return 0.002;
    if (mol == 2) return -0.003;
    return 0.0;
}

/* broadening.jl:40-53 */
void orc_broadened_width(const orc_linelist *ll, int64_t j, double p_atm, double T,
                         double vmr_self, double *gamma_l, double *gamma_d) {
    int mol = ll->mol[j];
    double n_air = (double)ll->n_air[j];
    double n_self = n_self_of(mol, n_air);
    double t_ratio = T_REF / T;
    double gl = pow(t_ratio, n_air) * (double)ll->gamma_air[j] * (1.0 - vmr_self) * p_atm +
                pow(t_ratio, n_self) * (double)ll->gamma_self[j] * vmr_self * p_atm;
    double m = mol_mass(mol);
    double gd = ll->nu[j] * 3.58126e-7 * sqrt(T / m);
    *gamma_l = gl;
    *gamma_d = gd;
}

/* broadening.jl:89-99 */
double orc_scaled_intensity(const orc_linelist *ll, int64_t j, double T) {
    double nu0 = ll->nu[j];
    double e = ll->elower[j];
    double qr = orc_q_ratio(ll->mol[j], ll->iso[j], T);
    double stim = (1.0 - exp(-C2_LINE * nu0 / T)) / (1.0 - exp(-C2_LINE * nu0 / T_REF));
    double boltz = exp(-C2_LINE * e * (1.0 / T - 1.0 / T_REF));
    return ll->S[j] * qr * boltz * stim;
}

/* broadening.jl:156-160 */
double orc_pressure_shift(const orc_linelist *ll, int64_t j, double p_atm, double vmr_self) {
    double ds = delta_self_of(ll->mol[j]);
    return ll->nu[j] + ((double)ll->delta_air[j] * (1.0 - vmr_self) + ds * vmr_self) * p_atm;
}

/* ========================================================================== *
 * DPTMIN line rejection (src/CrossSections/voigt.jl:333-355)
 * ========================================================================== */
int64_t orc_reject_weak_lines(const orc_linelist *ll, double T, double p_atm, double coef,
                              double vmr_self, double dptmn, uint8_t *keep) {
    int64_t nkeep = 0;
    if (dptmn <= 0.0) {
        for (int64_t j = 0; j < ll->n; ++j) keep[j] = 1;
        return ll->n;
    }
    for (int64_t j = 0; j < ll->n; ++j) {
        double S = orc_scaled_intensity(ll, j, T);
        double gl, gd;
        orc_broadened_width(ll, j, p_atm, T, vmr_self, &gl, &gd);
        gd = fmax(gd, 1e-10);
        double f = SQRT_LN2 / gd;
        double peak = (S * f * INV_SQRT_PI) * faddeeva_voigt(0.0, gl * f);
        keep[j] = (peak * coef > dptmn) ? 1 : 0;
        nkeep += keep[j];
    }
    return nkeep;
}

/* ========================================================================== *
 * Binary searches (voigt.jl:220-236), restated 0-based.
 * lower_bound: first j with a[j] >= v (n if none)
 * upper_bound: last  j with a[j] <= v (-1 if none)
 * The arrays are the SHIFTED centres of a list sorted by unshifted nu, exactly
 * as in the reference (parity trap: the searches run on a nearly sorted array).
 * ========================================================================== */
static int64_t lower_bound(const double *a, double v, int64_t n) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int64_t upper_bound(const double *a, double v, int64_t n) {
    int64_t lo = -1, hi = n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;   /* rounds up; lo+hi+1 >= 0 */
        if (a[mid] > v) hi = mid - 1; else lo = mid;
    }
    return lo;
}

/* ========================================================================== *
 * Uniform grid with Julia float-range semantics (wavenumber_grid.jl:23-27).
 * Julia's a:step:b recognises a rational step n/d and forms each element as
 * the correctly rounded (a*d + i*n)/d; we do the same when a*d, b*d and step*d
 * are integers (true for every grid in this path), else a + i*step.
 * ========================================================================== */
typedef struct { double a, step, an, sn, den; int rat; } grid_t;

/* continued-fraction convergents of the step, as Julia's rat() does */
static grid_t make_grid(double a, double step) {
    grid_t g = {a, step, 0.0, 0.0, 1.0, 0};
    double x = step, h0 = 0, h1 = 1, k0 = 1, k1 = 0;
    for (int it = 0; it < 40; ++it) {
        double ai = floor(x);
        double h2 = ai * h1 + h0, k2 = ai * k1 + k0;
        h0 = h1; h1 = h2; k0 = k1; k1 = k2;
        if (k1 > 1e9) break;
        if (fabs(h1 / k1 - step) <= 2.220446049250313e-16 * fabs(step)) {
            double ad = a * k1;
            if (h1 != 0.0 && fabs(ad - nearbyint(ad)) <= 1e-7 * fmax(1.0, fabs(ad))) {
                g.an = nearbyint(ad); g.sn = h1; g.den = k1; g.rat = 1;
            }
            break;
        }
        double fr = x - ai;
        if (fr == 0.0) break;
        x = 1.0 / fr;
    }
    return g;
}

static inline double grid_at(const grid_t *g, int64_t i) {
    return g->rat ? (g->an + (double)i * g->sn) / g->den : g->a + (double)i * g->step;
}

int64_t orc_grid_len(double a, double step, double b) {
    grid_t g = make_grid(a, step);
    if (g.rat) {
        double bn = b * g.den;
        if (fabs(bn - nearbyint(bn)) <= 1e-7 * fmax(1.0, fabs(bn)))
            return (int64_t)floor((nearbyint(bn) - g.an) / g.sn) + 1;
    }
    return (int64_t)floor((b - a) / step * (1.0 + 1e-14)) + 1;
}

double orc_grid_node(double a, double step, int64_t i) {
    grid_t g = make_grid(a, step);
    return grid_at(&g, i);
}

/* ========================================================================== *
 * Voigt cross sections: host precompute voigt.jl:391-414,445-454 and the
 * per-grid-point kernel voigt.jl:243-277 (FullFaddeeva, FP64).
 * ========================================================================== */
void orc_voigt_xsec(const orc_linelist *ll, const uint8_t *keep,
                    double nu0g, double dnu, int64_t n_nu,
                    double T, double p_atm, double vmr_self,
                    double cutoff, double x_far, double *sigma) {
    int64_t nl = 0;
    for (int64_t j = 0; j < ll->n; ++j) nl += (!keep || keep[j]) ? 1 : 0;
    double *c0 = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    double *fa = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    double *ya = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    double *sn = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    double *hc = (double *)malloc(sizeof(double) * (nl ? nl : 1));
    int64_t m = 0;
    for (int64_t j = 0; j < ll->n; ++j) {
        if (keep && !keep[j]) continue;
        double S = orc_scaled_intensity(ll, j, T);
        double gl, gd;
        orc_broadened_width(ll, j, p_atm, T, vmr_self, &gl, &gd);
        gd = fmax(gd, 1e-10);
        double f = SQRT_LN2 / gd;
        c0[m] = orc_pressure_shift(ll, j, p_atm, vmr_self);
        fa[m] = f;
        ya[m] = gl * f;
        sn[m] = S * f * INV_SQRT_PI;
        double xc = cutoff * f;
        hc[m] = fabs(xc) > x_far ? ya[m] * INV_SQRT_PI / (xc * xc + ya[m] * ya[m])
                                 : faddeeva_voigt(xc, ya[m]);
        ++m;
    }
    const double rc = 1.0 / cutoff;
    const grid_t gg = make_grid(nu0g, dnu);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n_nu; ++i) {
        double nu = grid_at(&gg, i);
        int64_t jlo = lower_bound(c0, nu - cutoff, nl);
        int64_t jhi = upper_bound(c0, nu + cutoff, nl);
        double acc = 0.0;
        for (int64_t j = jlo; j <= jhi; ++j) {
            double dv = nu - c0[j];
            double x = dv * fa[j];
            double y = ya[j];
            double zr = dv * rc;
            double ped = (2.0 - zr * zr) * hc[j];
            double H = voigt_H(x, y, x_far) - ped;
            acc += sn[j] * H;
        }
        sigma[i] = acc > 0.0 ? acc : 0.0;
    }
    free(c0); free(fa); free(ya); free(sn); free(hc);
}

/* ========================================================================== *
 * VP_Y dispersive term.  The reference builds its dispersive lines from the
 * relaxation-matrix S-files (line_mixing.jl:1193-1220); here the line list
 * carries Y per line, and the S-file conventions are kept: c2 = 1.4388
 * (line_mixing.jl:31), gamma_D constant 3.5812e-7 (:755), isotopologue mass
 * (:743-753), no self terms, S0 = S_ref at 296 K, lines sorted by shifted
 * centre (:1223-1236), kernel :1166-1186 (no pedestal, no clamp).
 * ========================================================================== */
#define C2_LM 1.4388
static double lm_mass(int mol, int iso) {
    static const double co2[10] = {43.98983, 44.99318, 45.99398, 44.99403, 45.99706,
                                   44.99706, 47.99832, 47.00134, 48.00196, 47.00427};
    static const double ch4[4] = {16.031300, 17.034655, 17.037475, 18.040830};
    if (mol == 2 && iso >= 1 && iso <= 10) return co2[iso - 1];
    if (mol == 2 && iso == 0) return co2[9];
    if (mol == 6 && iso >= 1 && iso <= 4) return ch4[iso - 1];
    return 44.0;
}

typedef struct { double c, f, y, w; } lm_line;
static int cmp_lm(const void *a, const void *b) {
    double ca = ((const lm_line *)a)->c, cb = ((const lm_line *)b)->c;
    return (ca > cb) - (ca < cb);
}

void orc_lm_dispersive(const orc_linelist *ll, double nu0g, double dnu, int64_t n_nu,
                       double T, double p_atm, double cutoff, double x_far, double *dsigma) {
    lm_line *L = (lm_line *)malloc(sizeof(lm_line) * (ll->n ? ll->n : 1));
    int64_t m = 0;
    for (int64_t j = 0; ll->y_lm && j < ll->n; ++j) {
        double Y = ll->y_lm[j];
        if (Y == 0.0) continue;
        int mol = ll->mol[j], iso = ll->iso[j];
        double M = lm_mass(mol, iso);
        double nu = ll->nu[j];
        double qr = orc_partition(mol, iso == 10 ? 10 : iso, T_REF) / orc_partition(mol, iso, T);
        double c = nu + (double)ll->delta_air[j] * p_atm;
        double gd = 3.5812e-7 * nu * sqrt(T / M);
        double gl = (double)ll->gamma_air[j] * pow(296.0 / T, (double)ll->n_air[j]) * p_atm;
        double f = SQRT_LN2 / fmax(gd, 1e-10);
        double st0 = 1.0 - exp(-C2_LM * nu / 296.0);
        double stt = 1.0 - exp(-C2_LM * nu / T);
        double s0 = ll->S[j];
        double st = s0 * qr * exp(-C2_LM * ll->elower[j] * (1.0 / T - 1.0 / 296.0)) * stt / st0;
        if (st <= 0.0) continue;
        L[m].c = c;
        L[m].f = f;
        L[m].y = gl * f;
        L[m].w = Y * p_atm * st * f * INV_SQRT_PI;
        ++m;
    }
    qsort(L, (size_t)m, sizeof(lm_line), cmp_lm);
    double *cs = (double *)malloc(sizeof(double) * (m ? m : 1));
    for (int64_t j = 0; j < m; ++j) cs[j] = L[j].c;
    const grid_t gg = make_grid(nu0g, dnu);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n_nu; ++i) {
        double nu = grid_at(&gg, i);
        double acc = 0.0;
        if (m > 0) {
            int64_t jlo = lower_bound(cs, nu - cutoff, m);
            int64_t jhi = upper_bound(cs, nu + cutoff, m);
            for (int64_t j = jlo; j <= jhi; ++j) {
                double x = (nu - cs[j]) * L[j].f;
                double yk = L[j].y;
                double imw;
                if (fabs(x) > x_far) imw = x * INV_SQRT_PI / (x * x + yk * yk);
                else { double re; orc_faddeeva(x, yk, &re, &imw); }
                acc += L[j].w * imw;
            }
        }
        dsigma[i] = acc;
    }
    free(cs);
    free(L);
}

void orc_voigt_lm_xsec(const orc_linelist *ll, double nu0g, double dnu, int64_t n_nu,
                       double T, double p_atm, double cutoff, double lm_cutoff,
                       double x_far, double *sigma) {
    double *d = (double *)malloc(sizeof(double) * (n_nu ? n_nu : 1));
    orc_voigt_xsec(ll, NULL, nu0g, dnu, n_nu, T, p_atm, 0.0, cutoff, x_far, sigma);
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T, p_atm, lm_cutoff, x_far, d);
    for (int64_t i = 0; i < n_nu; ++i) {
        double v = sigma[i] + d[i];
        sigma[i] = v > 0.0 ? v : 0.0;
    }
    free(d);
}

/* ========================================================================== *
 * Atmosphere (src/Atmosphere/layers.jl:10-106, src/Utils/interpolation.jl)
 * ========================================================================== */
static int approx_eq(double a, double b) {   /* Julia isapprox defaults: rtol = sqrt(eps) */
    double tol = 1.4901161193847656e-08 * fmax(fabs(a), fabs(b));
    return fabs(a - b) <= tol;
}

/* layers.jl:36-45 */
double orc_cg_column_vmr(double v1, double v2, double p1, double p2) {
    if (v1 <= 0.0 || v2 <= 0.0) return 0.5 * (v1 + v2);
    if (approx_eq(v1, v2)) return v1;
    double a = log(v2 / v1) / log(p2 / p1);
    double r = p2 / p1;
    double b = a + 1.0;
    if (fabs(b) < 1e-8) return (v1 - v2) / log(v1 / v2);
    return v1 * p1 * (1.0 - pow(r, b)) / (b * fabs(p1 - p2));
}

/* layers.jl:60-72 */
double orc_cg_pressure(double v1, double v2, double p1, double p2) {
    if (v1 <= 0.0 || v2 <= 0.0) return 0.5 * (p1 + p2);
    if (approx_eq(v1, v2)) return 0.5 * (p1 + p2);
    double a = log(v2 / v1) / log(p2 / p1);
    double r = p2 / p1;
    double b1 = a + 1.0, b2 = a + 2.0;
    if (fabs(b1) < 1e-8) return fabs(p1 - p2) / log(p1 / p2);
    if (fabs(b2) < 1e-8) return p2 * log(p2 / p1) / (p2 / p1 - 1.0);
    return p1 * b1 / b2 * (1.0 - pow(r, b2)) / (1.0 - pow(r, b1));
}

/* layers.jl:93-106 (TBAR) */
double orc_cg_temperature_mass(double T1, double T2, double p1, double p2) {
    if (approx_eq(T1, T2)) return T1;
    if (approx_eq(p1, p2)) return 0.5 * (T1 + T2);
    double a = log(p2 / p1);
    double frac;
    if (fabs(a) < 1e-3) frac = 0.5 + a / 12.0;
    else {
        double b = p2 / p1;
        frac = ((a - 1.0) * b + 1.0) / (a * (b - 1.0));
    }
    return T1 + (T2 - T1) * frac;
}

/* interpolation.jl:10-24: linear in log p, Flat extrapolation */
void orc_interp_logp(const double *p_ref, const double *x_ref, int n_ref,
                     const double *p_tgt, int n_tgt, double *out) {
    double *lp = (double *)malloc(sizeof(double) * n_ref);
    double *xv = (double *)malloc(sizeof(double) * n_ref);
    int ascending = 1;
    for (int k = 0; k < n_ref; ++k) lp[k] = log(p_ref[k]);
    for (int k = 1; k < n_ref; ++k) if (lp[k] < lp[k - 1]) ascending = 0;
    if (ascending) {
        for (int k = 0; k < n_ref; ++k) xv[k] = x_ref[k];
    } else {   /* profiles are surface-first (decreasing p): reverse */
        for (int k = 0; k < n_ref; ++k) { lp[k] = log(p_ref[n_ref - 1 - k]); xv[k] = x_ref[n_ref - 1 - k]; }
    }
    for (int t = 0; t < n_tgt; ++t) {
        double q = log(p_tgt[t]);
        if (q <= lp[0]) { out[t] = xv[0]; continue; }
        if (q >= lp[n_ref - 1]) { out[t] = xv[n_ref - 1]; continue; }
        int lo = 0, hi = n_ref - 1;
        while (hi - lo > 1) { int mid = (lo + hi) / 2; if (lp[mid] <= q) lo = mid; else hi = mid; }
        double w = (q - lp[lo]) / (lp[hi] - lp[lo]);
        out[t] = xv[lo] * (1.0 - w) + xv[hi] * w;
    }
    free(lp);
    free(xv);
}

/* layers.jl:133-178 (+ TBAR for :cim at sounder.jl:329-336) */
void orc_layer_properties(const orc_profile *prof, int T_method, double *p_mid, double *T_mid,
                          double *dp, double *vmr_cg, double *p_cg, double *T_cg,
                          double *T_ave) {
    int nl = prof->n_levels - 1;
    const double *p = prof->p_hpa, *T = prof->T;
    for (int k = 0; k < nl; ++k) {
        p_mid[k] = 0.5 * (p[k] + p[k + 1]);
        dp[k] = fabs(p[k] - p[k + 1]);
        T_ave[k] = orc_cg_temperature_mass(T[k], T[k + 1], p[k], p[k + 1]);
    }
    orc_interp_logp(p, T, prof->n_levels, p_mid, nl, T_mid);
    for (int s = 0; s < prof->n_species; ++s) {
        const double *v = prof->vmr + (size_t)s * prof->n_levels;
        double *vc = vmr_cg + (size_t)s * nl, *pc = p_cg + (size_t)s * nl, *tc = T_cg + (size_t)s * nl;
        for (int k = 0; k < nl; ++k) {
            vc[k] = orc_cg_column_vmr(v[k], v[k + 1], p[k], p[k + 1]);
            pc[k] = orc_cg_pressure(v[k], v[k + 1], p[k], p[k + 1]);
        }
        if (T_method == 1) for (int k = 0; k < nl; ++k) tc[k] = T_ave[k];
        else orc_interp_logp(p, T, prof->n_levels, pc, nl, tc);
    }
}

/* ========================================================================== *
 * Continuum (src/CrossSections/continuum.jl:138-363; sounder.jl:310-322;
 * optical_depth.jl:123-130)
 * ========================================================================== */
#define CK_KB    1.380649e-23
#define CK_C2    1.4387769
#define CK_PREF  1013.0
#define CK_TREF  296.0
#define CK_TEFF  246.0

/* linear interpolation on a table's ascending knots with Flat extrapolation
 * (Interpolations.linear_interpolation, continuum.jl:70-72,93-94) */
static double tab_interp(const double (*t)[4], int n, int col, double nu, const double *scale_col) {
    if (nu <= t[0][0]) return t[0][col] * (scale_col ? t[0][(int)scale_col[0]] : 1.0);
    if (nu >= t[n - 1][0]) return t[n - 1][col] * (scale_col ? t[n - 1][(int)scale_col[0]] : 1.0);
    int lo = 0, hi = n - 1;                  /* searchsortedlast */
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (t[mid][0] <= nu) lo = mid; else hi = mid;
    }
    double v0 = t[lo][col], v1 = t[hi][col];
    if (scale_col) {
        v0 *= t[lo][(int)scale_col[0]];
        v1 *= t[hi][(int)scale_col[0]];
    }
    double w = (nu - t[lo][0]) / (t[hi][0] - t[lo][0]);
    return v0 * (1.0 - w) + v1 * w;
}

double orc_radfn(double nu, double T) {         /* continuum.jl:150-155 */
    double x = CK_C2 * nu / T;
    if (x <= 0.01) return 0.5 * x * nu;
    if (x <= 10.0) return nu * (1.0 - exp(-x)) / (1.0 + exp(-x));
    return nu;
}

double orc_h2o_continuum(double nu, double vmr, double p, double T) {   /* continuum.jl:163-190 */
    double rho_rat = (p / CK_PREF) * (CK_TREF / T);
    double n_h2o = vmr * p * 100.0 / (CK_KB * T) * 1e-6;
    double Cs = tab_interp(irs_ckd_h2o, IRS_CKD_H2O_N, 1, nu, NULL);
    double Cf = tab_interp(irs_ckd_h2o, IRS_CKD_H2O_N, 2, nu, NULL);
    double tx = tab_interp(irs_ckd_h2o, IRS_CKD_H2O_N, 3, nu, NULL);
    double rad = orc_radfn(nu, T);
    double ks = Cs * pow(CK_TREF / T, tx) * vmr * rho_rat * rad * n_h2o;
    double kf = Cf * (1.0 - vmr) * rho_rat * rad * n_h2o;
    return ks + kf;
}

double orc_co2_continuum(double nu, double vmr, double p, double T) {   /* continuum.jl:209-232 */
    static const double xfac_col = 2.0;         /* S x xfac folded at load (continuum.jl:90) */
    double rho_rat = (p / CK_PREF) * (CK_TREF / T);
    double n_co2 = vmr * p * 100.0 / (CK_KB * T) * 1e-6;
    double pref = 1e-20 * n_co2 * rho_rat;
    double S = tab_interp(irs_ckd_co2, IRS_CKD_CO2_N, 1, nu, &xfac_col);
    double td = tab_interp(irs_ckd_co2, IRS_CKD_CO2_N, 3, nu, NULL);
    double tcor = td == 0.0 ? 1.0 : pow(T / CK_TEFF, td);
    return S * tcor * pref * orc_radfn(nu, T);
}

typedef struct {
    int nb;
    double *T;
    int64_t *off;
    double *nu, *sig;
} cia_tab;
static cia_tab g_cia[3];

int orc_cia_set(int pair, int n_blocks, const double *T, const int64_t *off, const double *nu,
                const double *sigma) {
    if (pair < 0 || pair > 2 || n_blocks < 0) return -1;
    cia_tab *c = &g_cia[pair];
    free(c->T); free(c->off); free(c->nu); free(c->sig);
    memset(c, 0, sizeof(*c));
    if (n_blocks == 0) return 0;
    int64_t nk = off[n_blocks];
    c->T = malloc(sizeof(double) * n_blocks);
    c->off = malloc(sizeof(int64_t) * (n_blocks + 1));
    c->nu = malloc(sizeof(double) * nk);
    c->sig = malloc(sizeof(double) * nk);
    memcpy(c->T, T, sizeof(double) * n_blocks);
    memcpy(c->off, off, sizeof(int64_t) * (n_blocks + 1));
    memcpy(c->nu, nu, sizeof(double) * nk);
    for (int64_t i = 0; i < nk; ++i) c->sig[i] = sigma[i] > 0.0 ? sigma[i] : 0.0;   /* :134 */
    c->nb = n_blocks;
    return 0;
}

/* one block's interpolant, zero outside its knots (extrapolation_bc = 0, :123) */
static double cia_block(const cia_tab *c, int b, double nu) {
    const double *x = c->nu + c->off[b], *y = c->sig + c->off[b];
    int64_t n = c->off[b + 1] - c->off[b];
    if (n < 1 || nu < x[0] || nu > x[n - 1]) return 0.0;
    if (n == 1) return y[0];
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) / 2;
        if (x[mid] <= nu) lo = mid; else hi = mid;
    }
    double w = (nu - x[lo]) / (x[hi] - x[lo]);
    return y[lo] * (1.0 - w) + y[hi] * w;
}

double orc_cia_sigma(int pair, double nu, double T) {   /* _cia_sigma, continuum.jl:336-363 */
    const cia_tab *c = &g_cia[pair];
    double T_lo = -INFINITY, s_lo = 0.0, T_hi = INFINITY, s_hi = 0.0;
    int any = 0;
    for (int b = 0; b < c->nb; ++b) {
        const double *x = c->nu + c->off[b];
        int64_t n = c->off[b + 1] - c->off[b];
        if (n < 1 || nu < x[0] || nu > x[n - 1]) continue;
        any = 1;
        double s = cia_block(c, b, nu);
        if (c->T[b] <= T && c->T[b] > T_lo) { T_lo = c->T[b]; s_lo = s; }
        if (c->T[b] >= T && c->T[b] < T_hi) { T_hi = c->T[b]; s_hi = s; }
    }
    if (!any) return 0.0;
    if (!isfinite(T_lo)) return s_hi;
    if (!isfinite(T_hi)) return s_lo;
    if (T_lo == T_hi) return s_lo;
    double w = (T - T_lo) / (T_hi - T_lo);
    return (1.0 - w) * s_lo + w * s_hi;
}

double orc_cia_k(int pair, double nu, double vmr, double p, double T) {   /* continuum.jl:249-327 */
    if (g_cia[pair].nb == 0) return 0.0;
    double n_x = vmr * p * 100.0 / (CK_KB * T) * 1e-6;
    double nn;
    if (pair == 0) nn = n_x * (p * 100.0 / (CK_KB * T) * 1e-6);   /* n_CO2 n_air */
    else nn = n_x * n_x;
    return orc_cia_sigma(pair, nu, T) * nn;
}

double orc_dp_to_dz_cm(double dp, double p_mid, double T) {       /* optical_depth.jl:123-130 */
    double H = 8.314462 * T / (0.028964 * 9.80665);
    return H * dp / p_mid * 100.0;
}

void orc_continuum_layer(double nu0, double dnu, int64_t n_nu, int continua, double vh, double vc,
                         double pm, double Tm, double dp, double *tau_c) {
    double dz = orc_dp_to_dz_cm(dp, pm, Tm);
    double vdry = 1.0 - vh;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nu; ++i) {
        double nu = orc_grid_node(nu0, dnu, i), t = 0.0;
        if (continua & ORC_CONT_H2O) t += orc_h2o_continuum(nu, vh, pm, Tm) * dz;
        if (continua & ORC_CONT_CO2) t += orc_co2_continuum(nu, vc, pm, Tm) * dz;
        if (continua & ORC_CONT_CO2CIA) t += orc_cia_k(0, nu, vc, pm, Tm) * dz;
        if (continua & ORC_CONT_N2) t += orc_cia_k(1, nu, 0.78084 * vdry, pm, Tm) * dz;
        if (continua & ORC_CONT_O2) t += orc_cia_k(2, nu, 0.20946 * vdry, pm, Tm) * dz;
        tau_c[i] = t;
    }
}

/* ========================================================================== *
 * Planck / BT (src/Solver/planck.jl:11-71)
 * ========================================================================== */
#define C1_PL 1.191042953e-5
#define C2_PL 1.4387769

double orc_planck(double nu, double T) { return C1_PL * nu * nu * nu / expm1(C2_PL * nu / T); }

double orc_bt(double nu, double R) {
    if (R <= 0.0) return 0.0;
    return C2_PL * nu / log1p(C1_PL * nu * nu * nu / R);
}

double orc_dbdt(double nu, double T) {
    double x = C2_PL * nu / T, ex = exp(x);
    return C1_PL * C2_PL * nu * nu * nu * nu * ex / (T * T * (ex - 1.0) * (ex - 1.0));
}

/* ========================================================================== *
 * RTE (src/Solver/schwarzschild.jl:43-64, 130-209, 228-274;
 *      src/Solver/transmittance.jl:26-43)
 * ========================================================================== */
double orc_lit_correction(double s) {
    if (s < 1e-4) return s * (0.5 - s * (1.0 / 3.0 - s * 0.125));
    double e = exp(-s);
    return (1.0 - e) / s - e;
}

double orc_cim_correction(double s) {
    if (s < 0.06) return s / 6.0;
    double e = exp(-s);
    return 1.0 - 2.0 * (e / (e - 1.0) + 1.0 / s);
}

double orc_lit_correction_deriv(double s) {
    if (s < 1e-4) return 0.5 - s * (2.0 / 3.0 - s * 0.375);
    double e = exp(-s);
    return (e * (s + 1.0) - 1.0) / (s * s) + e;
}

double orc_cim_correction_deriv(double s) {
    if (s < 0.06) return 1.0 / 6.0;
    double w = exp(-s);
    return 2.0 / (s * s) - 2.0 * w / ((w - 1.0) * (w - 1.0));
}

int orc_schwarzschild_rte(double nu0g, double dnu, int64_t n_nu, const double *tau, int L,
                          const double *T_lev, double T_sfc, double mu, double eps,
                          int source_function, const double *T_ave, double *out_I) {
    if (!(mu > 0.0 && mu <= 1.0)) return -1;
    if (source_function != 0 && source_function != 1) return -2;
    if (source_function == 1 && !T_ave) return -3;
    const grid_t gg = make_grid(nu0g, dnu);
#pragma omp parallel
    {
        double *tr = (double *)malloc(sizeof(double) * (L + 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n_nu; ++i) {
            double nu = grid_at(&gg, i);
            tr[L] = 1.0;
            for (int k = L - 1; k >= 0; --k) tr[k] = tr[k + 1] * exp(-tau[(size_t)k * n_nu + i] / mu);
            double acc = eps * orc_planck(nu, T_sfc) * tr[0];
            for (int k = 0; k < L; ++k) {
                double s = tau[(size_t)k * n_nu + i] / mu;
                double dT = tr[k + 1] - tr[k];
                if (source_function == 0) {
                    double bb = orc_planck(nu, T_lev[k]), bt = orc_planck(nu, T_lev[k + 1]);
                    double ck = tr[k + 1] * orc_lit_correction(s);
                    acc += bt * dT + (bb - bt) * ck;
                } else {
                    double ba = orc_planck(nu, T_ave[k]), bt = orc_planck(nu, T_lev[k + 1]);
                    double ck = dT * orc_cim_correction(s);
                    acc += ba * dT + (bt - ba) * ck;
                }
            }
            if (eps < 1.0) {   /* schwarzschild.jl:203-206,228-274 */
                double down = 0.0, g = 1.0;
                for (int k = 0; k < L; ++k) {
                    double s = tau[(size_t)k * n_nu + i] / mu;
                    double em;
                    if (source_function == 0) {
                        double bb = orc_planck(nu, T_lev[k]), bt = orc_planck(nu, T_lev[k + 1]);
                        em = bb * (1.0 - exp(-s)) + (bt - bb) * orc_lit_correction(s);
                    } else {
                        double ba = orc_planck(nu, T_ave[k]), bb = orc_planck(nu, T_lev[k]);
                        em = (1.0 - exp(-s)) * (ba + (bb - ba) * orc_cim_correction(s));
                    }
                    down += g * em;
                    g *= exp(-s);
                }
                acc += (1.0 - eps) * down * tr[0];
            }
            out_I[i] = acc;
        }
        free(tr);
    }
    return 0;
}

/* ========================================================================== *
 * ILS (src/Sensor/ils.jl:24-153)
 * ========================================================================== */
#define ILS_HALFWIDTH_CM 16.0
static const double NB[3][3] = {       /* ils.jl:26-30: C0, C2, C4 (weak, medium, strong) */
    {0.384093, 0.087577, 0.528330},
    {0.152442, 0.136176, 0.711382},
    {0.045335, 0.554883, 0.399782}};

double orc_norton_beer(double x, int style) {   /* ils.jl:38-44, style 1..3 */
    if (fabs(x) > 1.0) return 0.0;
    const double *c = NB[style - 1];
    double t = 1.0 - x * x;
    return c[0] + c[1] * t * t + c[2] * t * t * t * t;
}

int64_t orc_ils_half(double dnu) { return (int64_t)ceil(ILS_HALFWIDTH_CM / dnu); }

int orc_ils_kernel(double dnu, double opd_max, double fwhm, int apodization, int n_pts,
                   double *out) {
    if (apodization < 0 || apodization > 3) return -1;
    double *A = (double *)malloc(sizeof(double) * n_pts);
    double *L = (double *)malloc(sizeof(double) * n_pts);
    double dL = opd_max / (n_pts - 1);
    double sig = fwhm / (2.0 * sqrt(2.0 * log(2.0)));
    for (int m = 0; m < n_pts; ++m) {
        double t = (double)m / (double)(n_pts - 1);       /* LinRange: (1-t)*0 + t*opd */
        L[m] = t * opd_max;
        A[m] = (apodization == 0) ? exp(-2.0 * PI_D * PI_D * sig * sig * L[m] * L[m])
                                  : orc_norton_beer(L[m] / opd_max, apodization);
    }
    int64_t nh = orc_ils_half(dnu);
    int64_t n = 2 * nh + 1;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double dv = (double)(k - nh) * dnu;
        double s = 0.0;
        for (int m = 0; m < n_pts; ++m) s += A[m] * cos(2.0 * PI_D * L[m] * dv);
        out[k] = 2.0 * s * dL;
    }
    double tot = 0.0;
    for (int64_t k = 0; k < n; ++k) tot += out[k];
    double norm = tot * dnu;
    for (int64_t k = 0; k < n; ++k) out[k] /= norm;
    free(A);
    free(L);
    return 0;
}

void orc_apply_ils_at(const double *spec, int64_t n_nu, double dnu, const double *kern,
                      int64_t n_ils, const int64_t *idx, int64_t n_idx, double *out) {
    int64_t half = (n_ils - 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < n_idx; ++q) {
        int64_t i = idx[q];
        double acc = 0.0;
        for (int64_t k = 0; k < n_ils; ++k) {
            int64_t j = i + k - half;
            if (j < 0 || j >= n_nu) continue;
            acc += spec[j] * kern[k];
        }
        out[q] = acc * dnu;
    }
}

/* geometry.jl:92-100 (sind/asind in degrees) */
double orc_scan_to_local_zenith(double scan_deg, double h_sat_km, double r_earth_km) {
    double s = (r_earth_km + h_sat_km) / r_earth_km * sin(fabs(scan_deg) * PI_D / 180.0);
    if (s > 1.0) return -1.0;
    return asin(s) * 180.0 / PI_D;
}

/* ========================================================================== *
 * forward_model (src/Sensor/sounder.jl:233-360)
 * ========================================================================== */
static void internal_grid(const orc_sounder *snd, const orc_options *opt, int *hrf_out,
                          double *dnu_hi, double *lo, double *hi, double *step) {
    int hrf;
    if (opt->high_res_factor > 0) hrf = opt->high_res_factor;
    else if (opt->internal_dnu > 0.0) {
        hrf = (int)ceil(snd->dnu / opt->internal_dnu);
        if (hrf < 1) hrf = 1;
    } else hrf = 4;
    double d = snd->dnu / hrf;                               /* sounder.jl:263-270 */
    double pad = opt->with_ils ? ceil(ILS_HALFWIDTH_CM / d) * d : 0.0;   /* :155-161 */
    *hrf_out = hrf;
    *dnu_hi = d;
    *lo = snd->nu_min - pad;
    *hi = snd->nu_max + pad;
    *step = d;
}

int orc_forward_dims(const orc_sounder *snd, const orc_options *opt, orc_forward_info *info) {
    int hrf;
    double d, lo, hi, st;
    internal_grid(snd, opt, &hrf, &d, &lo, &hi, &st);
    info->nu_lo = lo;
    info->dnu = d;
    info->n_hi = orc_grid_len(lo, d, hi);
    info->n_ch = orc_grid_len(snd->nu_min, snd->dnu, snd->nu_max);
    return 0;
}

static int forward_impl(const orc_profile *prof, const orc_sounder *snd, const orc_options *opt,
                        double *R, double *BT, double *tau_out, const uint8_t *layer_mask,
                        orc_forward_info *info, int lm_species, const double *lm_add);

int orc_forward(const orc_profile *prof, const orc_sounder *snd, const orc_options *opt,
                double *R, double *BT, double *tau_out, const uint8_t *layer_mask,
                orc_forward_info *info) {
    return forward_impl(prof, snd, opt, R, BT, tau_out, layer_mask, info, 0, NULL);
}

/* forward_model with a relaxation-matrix line-mixing perturbation (VPYLineMixing /
 * VPWLineMixing): lm_add [n_layers][n_hi] is the perturbation of species
 * lm_species at each layer's T_cg, p_cg; that species' cross section is
 * max(sigma_voigt + lm_add, 0) (line_mixing.jl:222-241,1093-1118,1351-1369) and
 * DPTMIN is skipped for CO2 (sounder.jl:295-299). */
int orc_forward_lm(const orc_profile *prof, const orc_sounder *snd, const orc_options *opt,
                   int lm_species, const double *lm_add, double *R, double *BT, orc_forward_info *info) {
    return forward_impl(prof, snd, opt, R, BT, NULL, NULL, info, lm_species, lm_add);
}

static int forward_impl(const orc_profile *prof, const orc_sounder *snd, const orc_options *opt,
                        double *R, double *BT, double *tau_out, const uint8_t *layer_mask,
                        orc_forward_info *info, int lm_species, const double *lm_add) {
    orc_forward_info inf;
    memset(&inf, 0, sizeof(inf));
    orc_forward_dims(snd, opt, &inf);
    const double lo = inf.nu_lo, d = inf.dnu;
    const int64_t n = inf.n_hi;
    const int L = prof->n_levels - 1;
    const int ns = prof->n_species;

    double *p_mid = malloc(sizeof(double) * L), *T_mid = malloc(sizeof(double) * L);
    double *dp = malloc(sizeof(double) * L), *T_ave = malloc(sizeof(double) * L);
    double *vc = malloc(sizeof(double) * L * ns), *pc = malloc(sizeof(double) * L * ns);
    double *tc = malloc(sizeof(double) * L * ns);
    orc_layer_properties(prof, opt->T_method, p_mid, T_mid, dp, vc, pc, tc, T_ave);

    double *tau = tau_out ? tau_out : (double *)malloc(sizeof(double) * (size_t)n * L);
    memset(tau, 0, sizeof(double) * (size_t)n * L);
    double *sig = (double *)malloc(sizeof(double) * n);
    int64_t maxn = 1;
    for (int s = 0; s < ns; ++s) if (prof->lines[s].n > maxn) maxn = prof->lines[s].n;
    uint8_t *keep = (uint8_t *)malloc(maxn);
    const double nair = 2.1209e22;                            /* sounder.jl:280 */
    double t0 = now_s();
    for (int k = 0; k < L; ++k) {
        if (layer_mask && !layer_mask[k]) continue;
        for (int s = 0; s < ns; ++s) {
            const orc_linelist *ll = &prof->lines[s];
            if (ll->n == 0) continue;
            double vmr = vc[(size_t)s * L + k];
            if (vmr == 0.0) continue;
            double p_atm = pc[(size_t)s * L + k] / 1013.25;
            double T = tc[(size_t)s * L + k];
            int sp = prof->species[s];
            double vmr_self = (sp == 1) ? vmr : 0.0;
            double coef = vmr * dp[k] * nair;
            int lm_co2 = (sp == 2 && opt->line_mixing == 1);
            if (lm_co2) {
                orc_voigt_lm_xsec(ll, lo, d, n, T, p_atm, opt->cutoff, opt->lm_cutoff, opt->x_far, sig);
                inf.active_lines += ll->n;
                for (int64_t j = 0; j < ll->n; ++j) keep[j] = 1;
            } else {
                const double dpt = (lm_add && sp == 2) ? 0.0 : opt->dptmn;
                int64_t nk = orc_reject_weak_lines(ll, T, p_atm, coef, vmr_self, dpt, keep);
                inf.active_lines += nk;
                orc_voigt_xsec(ll, keep, lo, d, n, T, p_atm, vmr_self, opt->cutoff, opt->x_far, sig);
                if (lm_add && sp == lm_species) {
                    const double *ak = lm_add + (size_t)k * n;
                    for (int64_t i = 0; i < n; ++i) {
                        const double v = sig[i] + ak[i];
                        sig[i] = v > 0.0 ? v : 0.0;
                    }
                }
            }
            /* active pairs: lines whose window [c-cut, c+cut] overlaps the grid */
            for (int64_t j = 0; j < ll->n; ++j) {
                if (!keep[j]) continue;
                double c = orc_pressure_shift(ll, j, p_atm, vmr_self);
                double a = ceil((c - opt->cutoff - lo) / d), b = floor((c + opt->cutoff - lo) / d);
                if (a < 0) a = 0;
                if (b > (double)(n - 1)) b = (double)(n - 1);
                if (b >= a) inf.active_pairs += (int64_t)(b - a + 1);
            }
            double *tk = tau + (size_t)k * n;
            for (int64_t i = 0; i < n; ++i) tk[i] += sig[i] * coef;
        }
    }
    if (opt->continua) {                                       /* sounder.jl:310-322 */
        double *vh_mid = malloc(sizeof(double) * L), *vc_mid = malloc(sizeof(double) * L);
        int sh = -1, sc = -1;
        for (int s = 0; s < ns; ++s) {
            if (prof->species[s] == 1) sh = s;
            if (prof->species[s] == 2) sc = s;
        }
        if (sh >= 0) orc_interp_logp(prof->p_hpa, prof->vmr + (size_t)sh * prof->n_levels, prof->n_levels, p_mid, L, vh_mid);
        else for (int k = 0; k < L; ++k) vh_mid[k] = 0.0;
        if (sc >= 0) orc_interp_logp(prof->p_hpa, prof->vmr + (size_t)sc * prof->n_levels, prof->n_levels, p_mid, L, vc_mid);
        else for (int k = 0; k < L; ++k) vc_mid[k] = 4.15e-4;
        for (int k = 0; k < L; ++k) {
            orc_continuum_layer(lo, d, n, opt->continua, vh_mid[k], vc_mid[k], p_mid[k], T_mid[k], dp[k], sig);
            double *tk = tau + (size_t)k * n;
            for (int64_t i = 0; i < n; ++i) tk[i] += sig[i];
        }
        free(vh_mid);
        free(vc_mid);
    }
    double t1 = now_s();
    double T_sfc = opt->T_sfc > 0.0 ? opt->T_sfc : prof->T[0];
    double *Ihi = (double *)malloc(sizeof(double) * n);
    int rc = orc_schwarzschild_rte(lo, d, n, tau, L, prof->T, T_sfc, opt->mu, opt->eps_sfc,
                                   opt->source_function, T_ave, Ihi);
    double t2 = now_s();
    int64_t nch = inf.n_ch;
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * nch);
    for (int64_t c = 0; c < nch; ++c) {
        double nuc = orc_grid_node(snd->nu_min, snd->dnu, c);
        int64_t i = (int64_t)llround((nuc - lo) / d);
        if (i < 0) i = 0;
        if (i > n - 1) i = n - 1;
        idx[c] = i;
    }
    if (opt->with_ils) {
        int64_t nh = orc_ils_half(d);
        double *kern = (double *)malloc(sizeof(double) * (2 * nh + 1));
        orc_ils_kernel(d, snd->opd_max, snd->fwhm, opt->apodization, 2048, kern);
        if (opt->ils_direct_full) {
            /* the reference convolves every internal point, single-threaded */
            double *full = (double *)malloc(sizeof(double) * n);
            int64_t half = nh;
            for (int64_t i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int64_t k = 0; k < 2 * nh + 1; ++k) {
                    int64_t j = i + k - half;
                    if (j < 0 || j >= n) continue;
                    acc += Ihi[j] * kern[k];
                }
                full[i] = acc * d;
            }
            for (int64_t c = 0; c < nch; ++c) R[c] = full[idx[c]];
            free(full);
        } else {
            orc_apply_ils_at(Ihi, n, d, kern, 2 * nh + 1, idx, nch, R);
        }
        free(kern);
    } else {
        for (int64_t c = 0; c < nch; ++c) R[c] = Ihi[idx[c]];
    }
    double t3 = now_s();
    for (int64_t c = 0; c < nch; ++c) BT[c] = orc_bt(orc_grid_node(snd->nu_min, snd->dnu, c), R[c]);
    inf.seconds_xsec = t1 - t0;
    inf.seconds_rte = t2 - t1;
    inf.seconds_ils = t3 - t2;
    if (info) *info = inf;
    free(idx); free(Ihi); free(sig); free(keep);
    if (!tau_out) free(tau);
    free(p_mid); free(T_mid); free(dp); free(T_ave); free(vc); free(pc); free(tc);
    return rc;
}

/* ========================================================================== *
 * Jacobians
 * ========================================================================== */

/* broadening.jl:112-128 */
static void scaled_intensity_deriv(const orc_linelist *ll, int64_t j, double T, double *S, double *dS) {
    double nu0 = ll->nu[j], e = ll->elower[j];
    int mol = ll->mol[j], iso = ll->iso[j];
    double qr = orc_q_ratio(mol, iso, T);
    double qt = orc_partition(mol, iso, T);
    double dqt = orc_partition_deriv(mol, iso, T);
    double stim_num = 1.0 - exp(-C2_LINE * nu0 / T);
    double stim = stim_num / (1.0 - exp(-C2_LINE * nu0 / T_REF));
    double boltz = exp(-C2_LINE * e * (1.0 / T - 1.0 / T_REF));
    double s = ll->S[j] * qr * boltz * stim;
    double a = C2_LINE * nu0;
    double dln = -dqt / qt + C2_LINE * e / (T * T) - (a / (T * T)) / (exp(a / T) - 1.0);
    *S = s;
    *dS = s * dln;
}

/* broadening.jl:62-78 */
static void width_deriv(const orc_linelist *ll, int64_t j, double p, double T, double vs,
                        double *gl, double *gd, double *dgl, double *dgd, double *acoef, double *bcoef) {
    int mol = ll->mol[j];
    double n_air = (double)ll->n_air[j];
    double n_self = n_self_of(mol, n_air);
    double tr = T_REF / T;
    double A = pow(tr, n_air) * (double)ll->gamma_air[j] * (1.0 - vs) * p;
    double B = pow(tr, n_self) * (double)ll->gamma_self[j] * vs * p;
    *gl = A + B;
    *dgl = -(n_air * A + n_self * B) / T;
    *gd = ll->nu[j] * 3.58126e-7 * sqrt(T / mol_mass(mol));
    *dgd = *gd / (2.0 * T);
    *acoef = pow(tr, n_air) * (double)ll->gamma_air[j];
    *bcoef = pow(tr, n_self) * (double)ll->gamma_self[j];
}

/* cross_section_jacobian.jl:43-56 */
static void voigt_H_grad(double x, double y, double x_far, double *H, double *Hx, double *Hy) {
    if (fabs(x) > x_far) {
        double d = x * x + y * y, d2 = d * d;
        *H = y * INV_SQRT_PI / d;
        *Hx = -2.0 * x * y * INV_SQRT_PI / d2;
        *Hy = INV_SQRT_PI * (x * x - y * y) / d2;
        return;
    }
    double K, Lw;
    orc_faddeeva(x, y, &K, &Lw);
    *H = K;
    *Hx = -2.0 * x * K + 2.0 * y * Lw;
    *Hy = 2.0 * (x * Lw + y * K) - 2.0 * INV_SQRT_PI;
}

/* cross_section_jacobian.jl:68-114 + 155-240 */
static void voigt_xsec_grad(const orc_linelist *ll, double nu0g, double dnu, int64_t n_nu, double T,
                            double p, double vs, double cutoff, double x_far, double *sig, double *dT,
                            double *dp, double *dv) {
    int64_t n = ll->n;
    size_t sz = sizeof(double) * (n ? n : 1);
    double *c0 = malloc(sz), *fa = malloc(sz), *dfa = malloc(sz), *ya = malloc(sz), *dyT = malloc(sz);
    double *dyp = malloc(sz), *dyv = malloc(sz), *d0p = malloc(sz), *d0v = malloc(sz), *sn = malloc(sz);
    double *dsnT = malloc(sz), *hc = malloc(sz), *dhcT = malloc(sz), *dhcp = malloc(sz), *dhcv = malloc(sz);
    for (int64_t j = 0; j < n; ++j) {
        int mol = ll->mol[j];
        double dair = (double)ll->delta_air[j], dself = delta_self_of(mol);
        c0[j] = orc_pressure_shift(ll, j, p, vs);
        d0p[j] = dair * (1.0 - vs) + dself * vs;
        d0v[j] = (dself - dair) * p;
        double S, dS, gl, gd, dgl, dgd, ac, bc;
        scaled_intensity_deriv(ll, j, T, &S, &dS);
        width_deriv(ll, j, p, T, vs, &gl, &gd, &dgl, &dgd, &ac, &bc);
        gd = fmax(gd, 1e-10);
        double f = SQRT_LN2 / gd;
        double df = -f * (dgd / gd);
        double dglp = gl / p;
        double dglv = (bc - ac) * p;
        fa[j] = f; dfa[j] = df; ya[j] = gl * f;
        dyT[j] = dgl * f + gl * df;
        dyp[j] = dglp * f;
        dyv[j] = dglv * f;
        sn[j] = S * f * INV_SQRT_PI;
        dsnT[j] = (dS * f + S * df) * INV_SQRT_PI;
        double Hv, Hx, Hy;
        voigt_H_grad(cutoff * f, ya[j], x_far, &Hv, &Hx, &Hy);
        hc[j] = Hv;
        dhcT[j] = Hx * (cutoff * df) + Hy * dyT[j];
        dhcp[j] = Hy * dyp[j];
        dhcv[j] = Hy * dyv[j];
    }
    const grid_t gg = make_grid(nu0g, dnu);
    const double invc = 1.0 / cutoff, invc2 = invc * invc;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n_nu; ++i) {
        double nu = grid_at(&gg, i);
        int64_t jlo = lower_bound(c0, nu - cutoff, n);
        int64_t jhi = upper_bound(c0, nu + cutoff, n);
        double acc = 0, aT = 0, ap = 0, av = 0;
        for (int64_t j = jlo; j <= jhi; ++j) {
            double dv_ = nu - c0[j], f = fa[j], x = dv_ * f;
            double H, Hx, Hy;
            voigt_H_grad(x, ya[j], x_far, &H, &Hx, &Hy);
            double dHt = Hx * (dv_ * dfa[j]) + Hy * dyT[j];
            double dHp = Hx * (-d0p[j] * f) + Hy * dyp[j];
            double dHv = Hx * (-d0v[j] * f) + Hy * dyv[j];
            double pedfac = 2.0 - (dv_ * invc) * (dv_ * invc);
            double dpf_p = 2.0 * dv_ * d0p[j] * invc2;
            double dpf_v = 2.0 * dv_ * d0v[j] * invc2;
            double dpedT = pedfac * dhcT[j];
            double dpedp = dpf_p * hc[j] + pedfac * dhcp[j];
            double dpedv = dpf_v * hc[j] + pedfac * dhcv[j];
            double hm = H - pedfac * hc[j];
            acc += sn[j] * hm;
            aT += dsnT[j] * hm + sn[j] * (dHt - dpedT);
            ap += sn[j] * (dHp - dpedp);
            av += sn[j] * (dHv - dpedv);
        }
        if (acc > 0.0) { sig[i] = acc; dT[i] = aT; dp[i] = ap; dv[i] = av; }
        else { sig[i] = 0; dT[i] = 0; dp[i] = 0; dv[i] = 0; }
    }
    free(c0); free(fa); free(dfa); free(ya); free(dyT); free(dyp); free(dyv); free(d0p); free(d0v);
    free(sn); free(dsnT); free(hc); free(dhcT); free(dhcp); free(dhcv);
}

void orc_species_grad(const orc_linelist *ll, int line_mixing, double nu0g, double dnu, int64_t n_nu,
                      double T, double p, double vs, double cutoff, double lm_cutoff, double x_far,
                      double *sig, double *dT, double *dp, double *dv) {
    int lm = line_mixing && ll->y_lm && ll->mol[0] == 2;
    voigt_xsec_grad(ll, nu0g, dnu, n_nu, T, p, lm ? 0.0 : vs, cutoff, x_far, sig, dT, dp, dv);
    if (!lm) return;
    /* cross_section_jacobian.jl:276-322: central differences of the dispersive term */
    double *d0 = malloc(sizeof(double) * n_nu), *da = malloc(sizeof(double) * n_nu),
           *db = malloc(sizeof(double) * n_nu);
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T, p, lm_cutoff, x_far, d0);
    const double hT = 0.02;
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T + hT, p, lm_cutoff, x_far, da);
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T - hT, p, lm_cutoff, x_far, db);
    for (int64_t i = 0; i < n_nu; ++i) dT[i] += (da[i] - db[i]) / (2.0 * hT);
    const double hp = fmax(p, 1e-6) * 1e-4;
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T, p + hp, lm_cutoff, x_far, da);
    orc_lm_dispersive(ll, nu0g, dnu, n_nu, T, p - hp, lm_cutoff, x_far, db);
    for (int64_t i = 0; i < n_nu; ++i) {
        dp[i] += (da[i] - db[i]) / (2.0 * hp);
        sig[i] += d0[i];
        if (sig[i] <= 0.0) { sig[i] = 0; dT[i] = 0; dp[i] = 0; }
    }
    free(d0); free(da); free(db);
}

/* schwarzschild_jacobian.jl:47-53 */
double orc_cg_mass_frac(double p1, double p2) {
    if (approx_eq(p1, p2)) return 0.5;
    double a = log(p2 / p1);
    if (fabs(a) < 1e-3) return 0.5 + a / 12.0;
    double b = p2 / p1;
    return ((a - 1.0) * b + 1.0) / (a * (b - 1.0));
}

/* vmr_jacobian.jl:81-107 */
void orc_cg_column_vmr_grad(double v1, double v2, double p1, double p2, double *d1, double *d2) {
    if (v1 <= 0.0 || v2 <= 0.0) { *d1 = 0.5; *d2 = 0.5; return; }
    if (approx_eq(v1, v2)) { double fr = orc_cg_mass_frac(p1, p2); *d1 = 1.0 - fr; *d2 = fr; return; }
    double Lg = log(p2 / p1), al = log(v2 / v1) / Lg, be = al + 1.0, dp = fabs(p1 - p2);
    if (fabs(be) < 1e-6) {
        double h1 = fmax(fabs(v1), 1e-300) * 1e-6, h2 = fmax(fabs(v2), 1e-300) * 1e-6;
        *d1 = (orc_cg_column_vmr(v1 + h1, v2, p1, p2) - orc_cg_column_vmr(v1 - h1, v2, p1, p2)) / (2 * h1);
        *d2 = (orc_cg_column_vmr(v1, v2 + h2, p1, p2) - orc_cg_column_vmr(v1, v2 - h2, p1, p2)) / (2 * h2);
        return;
    }
    double r = p2 / p1, rb = pow(r, be);
    double c = v1 * p1 * (1.0 - rb) / (be * dp);
    double J = (v1 * p1 / Lg) * ((rb - 1.0) / (be * be) - rb * Lg / be);
    *d1 = c / v1 - J / (v1 * dp);
    *d2 = J / (v2 * dp);
}

/* vmr_jacobian.jl:120-125 */
double orc_dbt_dr(double nu, double R) {
    if (R <= 0.0) return 0.0;
    double A = C1_PL * nu * nu * nu, Lg = log1p(A / R);
    return C2_PL * nu * A / (Lg * Lg * R * (R + A));
}

int orc_rte_jacobian(double nu0g, double dnu, int64_t n_nu, const double *tau, int L,
                     const double *T_lev, const double *p_lev, double T_sfc, double mu, double eps,
                     int sf, double *Iout, double *dtau, double *dTlev, double *dTsfc, double *deps) {
    if (!(mu > 0.0 && mu <= 1.0)) return -1;
    if (sf != 0 && sf != 1) return -2;
    if (sf == 1 && !p_lev) return -3;
    double *Tave = malloc(sizeof(double) * L), *fr = malloc(sizeof(double) * L);
    for (int k = 0; k < L; ++k) {
        if (sf == 1) {
            Tave[k] = orc_cg_temperature_mass(T_lev[k], T_lev[k + 1], p_lev[k], p_lev[k + 1]);
            fr[k] = orc_cg_mass_frac(p_lev[k], p_lev[k + 1]);
        } else { Tave[k] = 0; fr[k] = 0; }
    }
    memset(dTlev, 0, sizeof(double) * (size_t)(L + 1) * n_nu);
    const grid_t gg = make_grid(nu0g, dnu);
    const double invmu = 1.0 / mu, refl = 1.0 - eps;
#pragma omp parallel
    {
        double *tr = malloc(sizeof(double) * (L + 1));
        double *gk = malloc(sizeof(double) * L), *em = malloc(sizeof(double) * L), *dem = malloc(sizeof(double) * L);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n_nu; ++i) {
            double nu = grid_at(&gg, i);
            tr[L] = 1.0;
            for (int k = L - 1; k >= 0; --k) tr[k] = tr[k + 1] * exp(-tau[(size_t)k * n_nu + i] / mu);
            double bs = orc_planck(nu, T_sfc), t1 = tr[0];
            double Irad = eps * bs * t1;
            deps[i] = bs * t1;
            dTsfc[i] = eps * orc_dbdt(nu, T_sfc) * t1;
            double below = eps * bs * t1;
            for (int m = 0; m < L; ++m) {
                double a = tau[(size_t)m * n_nu + i] * invmu, w = exp(-a), u = 1.0 - w, ttop = tr[m + 1];
                double emv, dem_da;
                if (sf == 1) {
                    double bavg = orc_planck(nu, Tave[m]), btop = orc_planck(nu, T_lev[m + 1]);
                    double f = orc_cim_correction(a), fp = orc_cim_correction_deriv(a);
                    double core = bavg + (btop - bavg) * f;
                    emv = u * core;
                    dem_da = w * core + u * (btop - bavg) * fp;
                    double cavg = ttop * u * (1.0 - f) * orc_dbdt(nu, Tave[m]);
                    dTlev[(size_t)m * n_nu + i] += cavg * (1.0 - fr[m]);
                    dTlev[(size_t)(m + 1) * n_nu + i] += cavg * fr[m] + ttop * u * f * orc_dbdt(nu, T_lev[m + 1]);
                } else {
                    double bbot = orc_planck(nu, T_lev[m]), btop = orc_planck(nu, T_lev[m + 1]);
                    double fl = orc_lit_correction(a), flp = orc_lit_correction_deriv(a);
                    emv = btop * u + (bbot - btop) * fl;
                    dem_da = btop * w + (bbot - btop) * flp;
                    dTlev[(size_t)m * n_nu + i] += ttop * fl * orc_dbdt(nu, T_lev[m]);
                    dTlev[(size_t)(m + 1) * n_nu + i] += ttop * (u - fl) * orc_dbdt(nu, T_lev[m + 1]);
                }
                dtau[(size_t)m * n_nu + i] = invmu * (-below + ttop * dem_da);
                Irad += ttop * emv;
                below += ttop * emv;
            }
            if (eps < 1.0) {   /* schwarzschild_jacobian.jl:189-266 */
                double g = 1.0, down = 0.0;
                for (int k = 0; k < L; ++k) {
                    double a = tau[(size_t)k * n_nu + i] * invmu, w = exp(-a), u = 1.0 - w;
                    gk[k] = g;
                    if (sf == 1) {
                        double bavg = orc_planck(nu, Tave[k]), bbot = orc_planck(nu, T_lev[k]);
                        double f = orc_cim_correction(a), fp = orc_cim_correction_deriv(a);
                        double core = bavg + (bbot - bavg) * f;
                        em[k] = u * core;
                        dem[k] = w * core + u * (bbot - bavg) * fp;
                        double cavg = refl * t1 * g * u * (1.0 - f) * orc_dbdt(nu, Tave[k]);
                        dTlev[(size_t)k * n_nu + i] += cavg * (1.0 - fr[k]) + refl * t1 * g * u * f * orc_dbdt(nu, T_lev[k]);
                        dTlev[(size_t)(k + 1) * n_nu + i] += cavg * fr[k];
                    } else {
                        double bbot = orc_planck(nu, T_lev[k]), btop = orc_planck(nu, T_lev[k + 1]);
                        double fl = orc_lit_correction(a), flp = orc_lit_correction_deriv(a);
                        em[k] = bbot * u + (btop - bbot) * fl;
                        dem[k] = bbot * w + (btop - bbot) * flp;
                        dTlev[(size_t)k * n_nu + i] += refl * t1 * g * (u - fl) * orc_dbdt(nu, T_lev[k]);
                        dTlev[(size_t)(k + 1) * n_nu + i] += refl * t1 * g * fl * orc_dbdt(nu, T_lev[k + 1]);
                    }
                    down += g * em[k];
                    g *= w;
                }
                Irad += refl * down * t1;
                deps[i] += -down * t1;
                double above = 0.0;
                for (int m = L - 1; m >= 0; --m) {
                    dtau[(size_t)m * n_nu + i] += refl * invmu * t1 * (-above + gk[m] * dem[m] - down);
                    above += gk[m] * em[m];
                }
            }
            Iout[i] = Irad;
        }
        free(tr); free(gk); free(em); free(dem);
    }
    free(Tave); free(fr);
    return 0;
}

int orc_state_size(const orc_state_spec *s, int nlev) {
    int n = s->include_temperature ? nlev : 0;
    for (int r = 0; r < s->n_retrieved; ++r) n += (s->column_scale && s->column_scale[r]) ? 1 : nlev;
    return n + (s->include_tsfc ? 1 : 0) + (s->include_emissivity ? 1 : 0);
}

int orc_jacobian(const orc_profile *prof, const orc_sounder *snd, const orc_options *opt,
                 const orc_state_spec *spec, double *K, double *y0, double *x0) {
    orc_forward_info inf;
    memset(&inf, 0, sizeof(inf));
    orc_forward_dims(snd, opt, &inf);
    const double lo = inf.nu_lo, d = inf.dnu;
    const int64_t n = inf.n_hi, nch = inf.n_ch;
    const int nlev = prof->n_levels, L = nlev - 1, ns = prof->n_species;
    const double *P = prof->p_hpa, *Tl = prof->T;
    double *p_mid = malloc(sizeof(double) * L), *T_mid = malloc(sizeof(double) * L);
    double *dpk = malloc(sizeof(double) * L), *T_ave = malloc(sizeof(double) * L);
    double *vc = malloc(sizeof(double) * L * ns), *pc = malloc(sizeof(double) * L * ns);
    double *tc = malloc(sizeof(double) * L * ns);
    orc_layer_properties(prof, opt->T_method, p_mid, T_mid, dpk, vc, pc, tc, T_ave);
    const size_t NL = (size_t)n * L;
    double *tau = calloc(NL, sizeof(double));
    double *DTlo = calloc(NL, sizeof(double)), *DThi = calloc(NL, sizeof(double));
    const int nr = spec->n_retrieved;
    double *Vlo = calloc(NL * (nr ? nr : 1), sizeof(double)), *Vhi = calloc(NL * (nr ? nr : 1), sizeof(double));
    double *sg = malloc(sizeof(double) * n), *gT = malloc(sizeof(double) * n);
    double *gp = malloc(sizeof(double) * n), *gv = malloc(sizeof(double) * n);
    for (int k = 0; k < L; ++k) {
        for (int s = 0; s < ns; ++s) {
            const orc_linelist *ll = &prof->lines[s];
            double vmr = vc[(size_t)s * L + k];
            if (ll->n == 0 || vmr == 0.0) continue;
            int sp = prof->species[s];
            double p_atm = pc[(size_t)s * L + k] / 1013.25, T = tc[(size_t)s * L + k];
            double vs = sp == 1 ? vmr : 0.0, coef = vmr * dpk[k] * 2.1209e22;
            int lm = (sp == 2 && opt->line_mixing == 1);
            orc_species_grad(ll, lm, lo, d, n, T, p_atm, vs, opt->cutoff, opt->lm_cutoff, opt->x_far,
                             sg, gT, gp, gv);
            double p1 = P[k], p2 = P[k + 1];
            double wT = opt->T_method == 1 ? orc_cg_mass_frac(p1, p2) : log(pc[(size_t)s * L + k] / p1) / log(p2 / p1);
            double *tk = tau + (size_t)k * n;
            for (int64_t i = 0; i < n; ++i) {
                tk[i] += sg[i] * coef;
                DTlo[(size_t)k * n + i] += gT[i] * coef * (1.0 - wT);
                DThi[(size_t)k * n + i] += gT[i] * coef * wT;
            }
            for (int r = 0; r < nr; ++r) {
                if (spec->retrieved[r] != sp) continue;
                const double *vl = prof->vmr + (size_t)s * nlev;
                double v1 = vl[k], v2 = vl[k + 1], gV1, gV2, gP1 = 0, gP2 = 0, gT1 = 0, gT2 = 0;
                orc_cg_column_vmr_grad(v1, v2, p1, p2, &gV1, &gV2);
                if (spec->vmr_coupling) {
                    double h1 = fmax(fabs(v1), 1e-300) * 1e-6, h2 = fmax(fabs(v2), 1e-300) * 1e-6;
                    gP1 = (orc_cg_pressure(v1 + h1, v2, p1, p2) - orc_cg_pressure(v1 - h1, v2, p1, p2)) / (2 * h1);
                    gP2 = (orc_cg_pressure(v1, v2 + h2, p1, p2) - orc_cg_pressure(v1, v2 - h2, p1, p2)) / (2 * h2);
                    if (opt->T_method != 1) {
                        double dtdp = (Tl[k + 1] - Tl[k]) / (pc[(size_t)s * L + k] * log(p2 / p1));
                        gT1 = dtdp * gP1;
                        gT2 = dtdp * gP2;
                    }
                }
                double *lo_ = Vlo + ((size_t)r * L + k) * n, *hi_ = Vhi + ((size_t)r * L + k) * n;
                for (int64_t i = 0; i < n; ++i) {
                    double ts = sg[i] * coef;
                    double base_lo = ts * (gV1 / vmr), base_hi = ts * (gV2 / vmr);
                    if (spec->vmr_coupling) {
                        double dtp = gp[i] * (coef / 1013.25), dtT = gT[i] * coef, dtv = gv[i] * coef;
                        base_lo += dtp * gP1 + dtT * gT1 + (sp == 1 ? dtv * gV1 : 0.0);
                        base_hi += dtp * gP2 + dtT * gT2 + (sp == 1 ? dtv * gV2 : 0.0);
                    }
                    lo_[i] = base_lo;
                    hi_[i] = base_hi;
                }
            }
        }
    }
    /* continuum and its central-difference derivatives (vmr_jacobian.jl:288-349):
     * T part when temperature is retrieved, VMR parts for retrieved H2O / CO2 with
     * vmr_coupling; chained to the levels by the mid-layer weight wM */
    double *CVlo = calloc(NL * (nr ? nr : 1), sizeof(double)), *CVhi = calloc(NL * (nr ? nr : 1), sizeof(double));
    if (opt->continua) {
        double *vh_mid = malloc(sizeof(double) * L), *vco_mid = malloc(sizeof(double) * L);
        int sh = -1, sc = -1;
        for (int s = 0; s < ns; ++s) {
            if (prof->species[s] == 1) sh = s;
            if (prof->species[s] == 2) sc = s;
        }
        if (sh >= 0) orc_interp_logp(P, prof->vmr + (size_t)sh * nlev, nlev, p_mid, L, vh_mid);
        else for (int k = 0; k < L; ++k) vh_mid[k] = 0.0;
        if (sc >= 0) orc_interp_logp(P, prof->vmr + (size_t)sc * nlev, nlev, p_mid, L, vco_mid);
        else for (int k = 0; k < L; ++k) vco_mid[k] = 4.15e-4;
        double *a = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n);
        for (int k = 0; k < L; ++k) {
            const double vh = vh_mid[k], vcx = vco_mid[k], Tm = T_mid[k], pm = p_mid[k], dp_ = dpk[k];
            const double wM = log(pm / P[k]) / log(P[k + 1] / P[k]);
            orc_continuum_layer(lo, d, n, opt->continua, vh, vcx, pm, Tm, dp_, a);
            for (int64_t i = 0; i < n; ++i) tau[(size_t)k * n + i] += a[i];
            if (spec->include_temperature) {
                const double hT = 0.05;
                orc_continuum_layer(lo, d, n, opt->continua, vh, vcx, pm, Tm + hT, dp_, a);
                orc_continuum_layer(lo, d, n, opt->continua, vh, vcx, pm, Tm - hT, dp_, b);
                for (int64_t i = 0; i < n; ++i) {
                    double g = (a[i] - b[i]) / (2 * hT);
                    DTlo[(size_t)k * n + i] += g * (1.0 - wM);
                    DThi[(size_t)k * n + i] += g * wM;
                }
            }
            for (int r = 0; r < nr && spec->vmr_coupling; ++r) {
                const int spid = spec->retrieved[r];
                if (spid != 1 && spid != 2) continue;
                const double v = spid == 1 ? vh : vcx, h = fmax(fabs(v), 1e-30) * 1e-6;
                if (spid == 1) {
                    orc_continuum_layer(lo, d, n, opt->continua, vh + h, vcx, pm, Tm, dp_, a);
                    orc_continuum_layer(lo, d, n, opt->continua, vh - h, vcx, pm, Tm, dp_, b);
                } else {
                    orc_continuum_layer(lo, d, n, opt->continua, vh, vcx + h, pm, Tm, dp_, a);
                    orc_continuum_layer(lo, d, n, opt->continua, vh, vcx - h, pm, Tm, dp_, b);
                }
                for (int64_t i = 0; i < n; ++i) {
                    double g = (a[i] - b[i]) / (2 * h);
                    CVlo[((size_t)r * L + k) * n + i] = g * (1.0 - wM);
                    CVhi[((size_t)r * L + k) * n + i] = g * wM;
                }
            }
        }
        free(a); free(b); free(vh_mid); free(vco_mid);
    }
    double T_sfc = opt->T_sfc > 0.0 ? opt->T_sfc : Tl[0];
    double *Ihi = malloc(sizeof(double) * n), *dtau = malloc(sizeof(double) * NL);
    double *dTl = malloc(sizeof(double) * (NL + n)), *dTs = malloc(sizeof(double) * n), *de = malloc(sizeof(double) * n);
    int rc = orc_rte_jacobian(lo, d, n, tau, L, Tl, P, T_sfc, opt->mu, opt->eps_sfc, opt->source_function,
                              Ihi, dtau, dTl, dTs, de);
    if (rc) return rc;
    /* channel-space linear tail: ILS at the channel nodes, x dBT/dR */
    int64_t nh = orc_ils_half(d), nt = 2 * nh + 1;
    double *kern = malloc(sizeof(double) * nt);
    orc_ils_kernel(d, snd->opd_max, snd->fwhm, opt->apodization, 2048, kern);
    int64_t *idx = malloc(sizeof(int64_t) * nch);
    double *nuc = malloc(sizeof(double) * nch), *Rc = malloc(sizeof(double) * nch), *dbt = malloc(sizeof(double) * nch);
    for (int64_t c = 0; c < nch; ++c) {
        nuc[c] = orc_grid_node(snd->nu_min, snd->dnu, c);
        int64_t i = (int64_t)llround((nuc[c] - lo) / d);
        idx[c] = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    }
    if (opt->with_ils) orc_apply_ils_at(Ihi, n, d, kern, nt, idx, nch, Rc);
    else for (int64_t c = 0; c < nch; ++c) Rc[c] = Ihi[idx[c]];
    for (int64_t c = 0; c < nch; ++c) {
        dbt[c] = spec->observable == 0 ? orc_dbt_dr(nuc[c], Rc[c]) : 1.0;
        y0[c] = spec->observable == 0 ? orc_bt(nuc[c], Rc[c]) : Rc[c];
    }
    double *col = malloc(sizeof(double) * n), *ch = malloc(sizeof(double) * nch);
    int ci = 0;
#define EMIT(colptr)                                                                 \
    do {                                                                             \
        if (opt->with_ils) orc_apply_ils_at(colptr, n, d, kern, nt, idx, nch, ch);   \
        else for (int64_t c_ = 0; c_ < nch; ++c_) ch[c_] = (colptr)[idx[c_]];        \
        for (int64_t c_ = 0; c_ < nch; ++c_) K[(size_t)ci * nch + c_] = ch[c_] * dbt[c_]; \
    } while (0)
    if (spec->include_temperature) {
        for (int j = 0; j < nlev; ++j) {
            for (int64_t i = 0; i < n; ++i) {
                double v = dTl[(size_t)j * n + i];
                if (j < L) v += dtau[(size_t)j * n + i] * DTlo[(size_t)j * n + i];
                if (j >= 1) v += dtau[(size_t)(j - 1) * n + i] * DThi[(size_t)(j - 1) * n + i];
                col[i] = v;
            }
            EMIT(col);
            x0[ci] = Tl[j];
            ++ci;
        }
    }
    for (int r = 0; r < nr; ++r) {
        int sidx = -1;
        for (int s = 0; s < ns; ++s) if (prof->species[s] == spec->retrieved[r]) sidx = s;
        const double *vl = sidx >= 0 ? prof->vmr + (size_t)sidx * nlev : NULL;
        int cs = spec->column_scale && spec->column_scale[r];
        if (cs) memset(col, 0, sizeof(double) * n);
        for (int j = 0; j < nlev; ++j) {
            double vj = vl ? vl[j] : 0.0;
            double *tmp = cs ? malloc(sizeof(double) * n) : col;
            for (int64_t i = 0; i < n; ++i) {
                double v = 0.0;
                if (vl) {
                    if (j < L && vc[(size_t)sidx * L + j] != 0.0) v += dtau[(size_t)j * n + i] * Vlo[((size_t)r * L + j) * n + i];
                    if (j >= 1 && vc[(size_t)sidx * L + j - 1] != 0.0)
                        v += dtau[(size_t)(j - 1) * n + i] * Vhi[((size_t)r * L + j - 1) * n + i];
                }
                /* continuum term, not gated by the CG VMR (vmr_jacobian.jl:451-454) */
                if (j < L) v += dtau[(size_t)j * n + i] * CVlo[((size_t)r * L + j) * n + i];
                if (j >= 1) v += dtau[(size_t)(j - 1) * n + i] * CVhi[((size_t)r * L + j - 1) * n + i];
                tmp[i] = v;
            }
            if (cs) {
                for (int64_t i = 0; i < n; ++i) col[i] += vj * tmp[i];
                free(tmp);
            } else {
                if (spec->log_vmr) for (int64_t i = 0; i < n; ++i) col[i] *= vj;
                EMIT(col);
                x0[ci] = spec->log_vmr ? log(fmax(vj, 1e-30)) : vj;
                ++ci;
            }
        }
        if (cs) { EMIT(col); x0[ci] = 0.0; ++ci; }
    }
    if (spec->include_tsfc) { EMIT(dTs); x0[ci] = T_sfc; ++ci; }
    if (spec->include_emissivity) { EMIT(de); x0[ci] = opt->eps_sfc; ++ci; }
#undef EMIT
    free(col); free(ch); free(kern); free(idx); free(nuc); free(Rc); free(dbt);
    free(Ihi); free(dtau); free(dTl); free(dTs); free(de);
    free(tau); free(DTlo); free(DThi); free(Vlo); free(Vhi); free(CVlo); free(CVhi);
    free(sg); free(gT); free(gp); free(gv);
    free(p_mid); free(T_mid); free(dpk); free(T_ave); free(vc); free(pc); free(tc);
    return 0;
}
