// Relaxation-matrix line mixing (src/CrossSections/line_mixing.jl).
//
// Host: the T-independent coupling table of every band at upload; per layer
// (T, p) the sum-rule-corrected W(T), Y(T) for the first-order (VP_Y) bands and
// the eigenmodes of D + i p W for the VP_W bands, all turned into one sorted list
// of terms  a Re w(z) + b Im w(z),  z = (nu - c) f + i y.
// Device: lm_terms_kernel sums the terms of every (profile, layer) over the
// internal grid -- one workgroup per 1024-point tile, terms staged through LDS,
// w(z) in FP32 (VP_Y terms) or FP64 (VP_W terms, whose full and no-coupling sums
// nearly cancel), the far-wing form beyond |x| > x_far -- and either stores the
// perturbation or folds it into the optical depths of the line-mixing species:
// tau += max(tau_species + coef * dsigma, 0).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <tuple>

#include "irs_devmath.h"
#include "irs_relmat.h"

double relmat_partition(int mol, int iso, double T);   // TIPS-2024 (irs_api.hip)

namespace {

using cd = std::complex<double>;

constexpr double kT0 = 296.0;          // line_mixing.jl:30
constexpr double kCT = 1.4388;         // line_mixing.jl:31
constexpr double kCTGAMD = 3.5812e-7;  // line_mixing.jl:755

double lm_mass(int mol, int iso) {     // line_mixing.jl:743-753,772
    static const double co2[10] = {43.98983, 44.99318, 45.99398, 44.99403, 45.99706,
                                   44.99706, 47.99832, 47.00134, 48.00196, 47.00427};
    static const double ch4[4] = {16.031300, 17.034655, 17.037475, 18.040830};
    if (mol == 2 && iso >= 1 && iso <= 10) return co2[iso - 1];
    if (mol == 6 && iso >= 1 && iso <= 4) return ch4[iso - 1];
    return 44.0;
}

bool sym_restrict(int isot) { return isot > 2 && isot != 7 && isot != 10; }   // line_mixing.jl:538

int branch_type(int j, int jf) { return j > jf ? 0 : (j == jf ? 1 : 2); }     // p, q, r

double q_ratio(const RelmatBandHost &b, double T) {   // line_mixing.jl:615,1198
    return relmat_partition(b.molecule, b.isot, kT0) / relmat_partition(b.molecule, b.isot, T);
}

}  // namespace

std::string relmat_build(const irs_relmat *rm, RelmatHost &out) {
    if (!rm || rm->n_bands < 0 || (rm->n_bands > 0 && (!rm->li || !rm->lf || !rm->isot || !rm->molecule ||
                                                        !rm->line_off)))
        return "null relaxation-matrix field";
    const int64_t nl = rm->n_bands > 0 ? rm->line_off[rm->n_bands] : 0;
    if (nl < 0 || (rm->n_bands > 0 && rm->line_off[0] != 0)) return "line_off must start at 0";
    if (nl > 0 && (!rm->Ji || !rm->branch || !rm->nu || !rm->gV_air || !rm->E_lower || !rm->Dipo0 ||
                   !rm->PopuT0 || !rm->DipoT || !rm->n_air || !rm->shift))
        return "null line field";
    if (rm->n_tables < 0 || (rm->n_tables > 0 && (!rm->table_lli || !rm->table_llf || !rm->entry_off)))
        return "null table field";
    const int64_t ne = rm->n_tables > 0 ? rm->entry_off[rm->n_tables] : 0;
    if (ne > 0 && (!rm->entry_sub || !rm->entry_ji || !rm->entry_jip || !rm->entry_w0 || !rm->entry_b0))
        return "null table entry field";
    out = RelmatHost{};
    out.Ji.assign(rm->Ji, rm->Ji + nl);
    out.branch.assign(rm->branch, rm->branch + nl);
    out.nu.assign(rm->nu, rm->nu + nl);
    out.gV_air.assign(rm->gV_air, rm->gV_air + nl);
    out.E_lower.assign(rm->E_lower, rm->E_lower + nl);
    out.Dipo0.assign(rm->Dipo0, rm->Dipo0 + nl);
    out.PopuT0.assign(rm->PopuT0, rm->PopuT0 + nl);
    out.DipoT.assign(rm->DipoT, rm->DipoT + nl);
    out.n_air.assign(rm->n_air, rm->n_air + nl);
    out.shift.assign(rm->shift, rm->shift + nl);
    // W0/B0 lookup: (table, sub-table, Ji, Jip) -> (W0R, B0R); a later duplicate wins, as
    // the reference's Dict assignment does (line_mixing.jl:472)
    std::map<std::tuple<int, int, int, int>, std::pair<double, double>> wt;
    std::map<std::pair<int, int>, int> table_of;
    for (int t = 0; t < rm->n_tables; ++t) {
        table_of[{rm->table_lli[t], rm->table_llf[t]}] = t;
        for (int64_t e = rm->entry_off[t]; e < rm->entry_off[t + 1]; ++e) {
            if (rm->entry_sub[e] < 0 || rm->entry_sub[e] > 8) return "entry_sub must be 0..8";
            wt[{t, rm->entry_sub[e], rm->entry_ji[e], rm->entry_jip[e]}] = {rm->entry_w0[e], rm->entry_b0[e]};
        }
    }
    for (int bi = 0; bi < rm->n_bands; ++bi) {
        RelmatBandHost b;
        b.li = rm->li[bi];
        b.lf = rm->lf[bi];
        b.isot = rm->isot[bi];
        b.molecule = rm->molecule[bi];
        b.line0 = rm->line_off[bi];
        b.n = rm->line_off[bi + 1] - b.line0;
        if (b.n < 0) return "line_off must be non-decreasing";
        auto it = table_of.find({std::min(b.li, b.lf), std::max(b.li, b.lf)});
        b.table = it == table_of.end() ? -1 : it->second;
        // coupling pairs (line_mixing.jl:530-578): outer i, inner j != i with Ji_j <= Ji_i,
        // the isotopologue symmetry rule, a W0/B0 entry under the swapped roles when li > lf
        b.row_ptr.assign((size_t)b.n + 1, 0);
        const bool sym = sym_restrict(b.isot);
        for (int64_t i = 0; i < b.n; ++i) {
            b.row_ptr[i] = (int32_t)b.col.size();
            if (b.table < 0) continue;
            const int jji = out.Ji[b.line0 + i], jjf = jji + out.branch[b.line0 + i];
            for (int64_t j = 0; j < b.n; ++j) {
                if (j == i) continue;
                const int jjip = out.Ji[b.line0 + j], jjfp = jjip + out.branch[b.line0 + j];
                if (jjip > jji) continue;
                if (sym && std::abs(jji - jjip) % 2 != 0) continue;
                int a0 = jji, a1 = jjf, c0 = jjip, c1 = jjfp;
                if (b.li > b.lf) { a0 = jjf; a1 = jji; c0 = jjfp; c1 = jjip; }
                const int sub = 3 * branch_type(a0, a1) + branch_type(c0, c1);
                auto e = wt.find({b.table, sub, a0, c0});
                if (e == wt.end()) continue;
                b.col.push_back((int32_t)j);
                b.w0r.push_back(e->second.first);
                b.b0r.push_back(e->second.second);
            }
        }
        b.row_ptr[b.n] = (int32_t)b.col.size();
        out.bands.push_back(std::move(b));
    }
    return "";
}

// line_mixing.jl:607-710
void relmat_build_W(const RelmatHost &rm, const RelmatBandHost &b, double T, std::vector<double> &W,
                    std::vector<double> &rho) {
    const int64_t n = b.n, o = b.line0;
    W.assign((size_t)(n * n), 0.0);
    rho.resize((size_t)n);
    const double qr = q_ratio(b, T);
    for (int64_t i = 0; i < n; ++i)
        rho[i] = rm.PopuT0[o + i] * qr * std::exp(-kCT * rm.E_lower[o + i] * (1.0 / T - 1.0 / kT0));
    auto gam = [&](int64_t i) { return rm.gV_air[o + i] * std::pow(kT0 / T, (double)rm.n_air[o + i]); };
    if (b.li > 8 || b.lf > 8 || std::abs(b.li - b.lf) > 1) {
        for (int64_t i = 0; i < n; ++i) W[i * n + i] = gam(i);
        return;
    }
    std::vector<double> seff((size_t)n);
    for (int64_t i = 0; i < n; ++i) seff[i] = rm.nu[o + i] * rho[i] * rm.DipoT[o + i] * rm.DipoT[o + i];
    std::vector<int64_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return seff[x] > seff[y]; });
    const double dl = std::log(kT0 / T);
    // W(j, i) stored at W[j * n + i]; outer order = descending intensity (later writers win)
    for (int64_t ir = 0; ir < n; ++ir) {
        const int64_t i = ord[ir];
        for (int32_t t = b.row_ptr[i]; t < b.row_ptr[i + 1]; ++t) {
            const int64_t j = b.col[t];
            const double y = std::exp(b.w0r[t] - b.b0r[t] * dl);
            W[j * n + i] = y;
            W[i * n + j] = y * rho[i] / rho[j];
        }
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
            if (i != j) W[i * n + j] = -std::fabs(W[i * n + j]);
    for (int64_t i = 0; i < n; ++i) W[i * n + i] = gam(i);
    // sum rule: sum_j |Dipo0_j| W(j, i) = 0, the lower (weaker) part rescaled in sorted order
    const bool sym = sym_restrict(b.isot);
    for (int64_t ir = 0; ir < n; ++ir) {
        const int64_t i = ord[ir];
        const int ji = rm.Ji[o + i];
        double lw = 0.0, up = 0.0;
        for (int64_t irp = 0; irp < n; ++irp) {
            const int64_t j = ord[irp];
            if (sym && std::abs(ji - (int)rm.Ji[o + j]) % 2 != 0) continue;
            const double v = std::fabs(rm.Dipo0[o + j]) * W[j * n + i];
            if (irp > ir) lw += v; else up += v;
        }
        if (lw == 0.0) continue;
        const double r = -up / lw;
        for (int64_t irp = ir + 1; irp < n; ++irp) {
            const int64_t j = ord[irp];
            W[j * n + i] *= r;
            W[i * n + j] = W[j * n + i] * rho[i] / rho[j];
        }
    }
}

// ---- complex eigen-decomposition ---------------------------------------------------
// Householder reduction to upper Hessenberg form, explicitly shifted QR with
// Wilkinson shifts and deflation to the Schur form T = Q^H A Q, eigenvectors of T
// by back-substitution (perturbed denominators as in LAPACK's ztrevc), V = Q Y.
bool complex_eig(int n, std::vector<cd> &A, std::vector<cd> &w, std::vector<cd> &V) {
    auto a = [&](int i, int j) -> cd & { return A[(size_t)i * n + j]; };
    std::vector<cd> Q((size_t)n * n, cd(0.0));
    auto q = [&](int i, int j) -> cd & { return Q[(size_t)i * n + j]; };
    for (int i = 0; i < n; ++i) q(i, i) = 1.0;
    w.assign(n, cd(0.0));
    double anorm = 0.0;
    for (auto &v : A) anorm = std::max(anorm, std::abs(v));
    if (n == 0) { V.clear(); return true; }
    // Hessenberg reduction
    std::vector<cd> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        double xn = 0.0;
        for (int i = k + 1; i < n; ++i) xn += std::norm(a(i, k));
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const cd x0 = a(k + 1, k);
        const cd ph = std::abs(x0) > 0.0 ? x0 / std::abs(x0) : cd(1.0);
        const cd alpha = -ph * xn;
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        v[k + 1] = x0 - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = a(i, k);
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
        if (vn == 0.0) continue;
        const double beta = 2.0 / vn;
        // A = (I - beta v v^H) A
        for (int j = 0; j < n; ++j) {
            cd s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * a(i, j);
            s *= beta;
            for (int i = k + 1; i < n; ++i) a(i, j) -= v[i] * s;
        }
        // A = A (I - beta v v^H), Q = Q (I - beta v v^H)
        for (int i = 0; i < n; ++i) {
            cd s = 0.0, t = 0.0;
            for (int j = k + 1; j < n; ++j) {
                s += a(i, j) * v[j];
                t += q(i, j) * v[j];
            }
            s *= beta;
            t *= beta;
            for (int j = k + 1; j < n; ++j) {
                a(i, j) -= s * std::conj(v[j]);
                q(i, j) -= t * std::conj(v[j]);
            }
        }
        for (int i = k + 2; i < n; ++i) a(i, k) = 0.0;
    }
    // shifted QR
    const double eps = 2.220446049250313e-16;
    std::vector<double> gc(n);
    std::vector<cd> gs(n);
    int hi = n - 1, iter = 0, total = 0;
    while (hi >= 0) {
        int l = hi;
        while (l > 0) {
            double s = std::abs(a(l - 1, l - 1)) + std::abs(a(l, l));
            if (s == 0.0) s = anorm;
            if (std::abs(a(l, l - 1)) <= eps * s) {
                a(l, l - 1) = 0.0;
                break;
            }
            --l;
        }
        if (l == hi) {
            w[hi] = a(hi, hi);
            --hi;
            iter = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        ++iter;
        cd mu;
        if (iter % 11 == 10) {
            mu = a(hi, hi) + std::abs(a(hi, hi - 1).real()) + std::abs(a(hi, hi - 1).imag());   // exceptional
        } else {
            const cd p = a(hi - 1, hi - 1), r = a(hi - 1, hi), c = a(hi, hi - 1), d = a(hi, hi);
            const cd half = 0.5 * (p - d);
            const cd disc = std::sqrt(half * half + r * c);
            const cd m1 = d - r * c / (half + disc), m2 = d - r * c / (half - disc);
            const bool ok1 = std::abs(half + disc) > 0.0, ok2 = std::abs(half - disc) > 0.0;
            if (ok1 && ok2) mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
            else mu = ok1 ? m1 : (ok2 ? m2 : d);
        }
        for (int k = l; k <= hi; ++k) a(k, k) -= mu;
        for (int k = l; k < hi; ++k) {
            const cd x = a(k, k), y = a(k + 1, k);
            const double ax = std::abs(x), ay = std::abs(y);
            const double nr = std::hypot(ax, ay);
            double c;
            cd s;
            if (nr == 0.0) { c = 1.0; s = 0.0; }
            else if (ax == 0.0) { c = 0.0; s = std::conj(y) / ay; }
            else { c = ax / nr; s = (x / ax) * std::conj(y) / nr; }
            gc[k] = c;
            gs[k] = s;
            for (int j = k; j < n; ++j) {
                const cd u = a(k, j), t = a(k + 1, j);
                a(k, j) = c * u + s * t;
                a(k + 1, j) = -std::conj(s) * u + c * t;
            }
        }
        for (int k = l; k < hi; ++k) {
            const double c = gc[k];
            const cd s = gs[k];
            const int top = std::min(k + 1, hi);
            for (int i = 0; i <= top; ++i) {
                const cd u = a(i, k), t = a(i, k + 1);
                a(i, k) = u * c + t * std::conj(s);
                a(i, k + 1) = -u * s + t * c;
            }
            for (int i = 0; i < n; ++i) {
                const cd u = q(i, k), t = q(i, k + 1);
                q(i, k) = u * c + t * std::conj(s);
                q(i, k + 1) = -u * s + t * c;
            }
        }
        for (int k = l; k <= hi; ++k) a(k, k) += mu;
    }
    // eigenvectors of the triangular T, then V = Q Y
    const double smlnum = 1e-300 * n / eps;
    std::vector<cd> Y((size_t)n * n, cd(0.0));
    std::vector<cd> y(n);
    for (int k = n - 1; k >= 0; --k) {
        const cd lam = a(k, k);
        const double smin = std::max(eps * (std::abs(lam.real()) + std::abs(lam.imag())), smlnum);
        for (int i = 0; i < n; ++i) y[i] = 0.0;
        y[k] = 1.0;
        for (int i = k - 1; i >= 0; --i) {
            cd s = 0.0;
            for (int j = i + 1; j <= k; ++j) s += a(i, j) * y[j];
            cd d = a(i, i) - lam;
            if (std::abs(d.real()) + std::abs(d.imag()) < smin) d = smin;
            y[i] = -s / d;
        }
        for (int i = 0; i <= k; ++i) Y[(size_t)i * n + k] = y[i];
    }
    V.assign((size_t)n * n, cd(0.0));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            cd s = 0.0;
            for (int j = 0; j <= k; ++j) s += q(i, j) * Y[(size_t)j * n + k];
            V[(size_t)i * n + k] = s;
        }
    return true;
}

namespace {

// Solves X b = r (LU with partial pivoting, X row-major, destroyed).
bool complex_solve(int n, std::vector<cd> &X, std::vector<cd> &r) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(X[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(X[(size_t)i * n + k]) > best) { best = std::abs(X[(size_t)i * n + k]); p = i; }
        if (best == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(X[(size_t)k * n + j], X[(size_t)p * n + j]);
            std::swap(r[k], r[p]);
        }
        const cd piv = X[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const cd m = X[(size_t)i * n + k] / piv;
            if (m == cd(0.0)) continue;
            for (int j = k; j < n; ++j) X[(size_t)i * n + j] -= m * X[(size_t)k * n + j];
            r[i] -= m * r[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        cd s = r[i];
        for (int j = i + 1; j < n; ++j) s -= X[(size_t)i * n + j] * r[j];
        r[i] = s / X[(size_t)i * n + i];
    }
    return true;
}

// Y(T) (line_mixing.jl:712-738) and the dispersive lines (line_mixing.jl:1193-1221)
void vpy_terms(const RelmatHost &rm, const RelmatBandHost &b, double T, double p, std::vector<LmTerm> &out) {
    const int64_t n = b.n, o = b.line0;
    if (b.li > 8 || b.lf > 8 || std::abs(b.li - b.lf) > 1) return;       // Y = 0
    std::vector<double> W, rho;
    relmat_build_W(rm, b, T, W, rho);
    const double M = lm_mass(b.molecule, b.isot);
    const double qr = q_ratio(b, T);
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double d = rm.nu[o + i] - rm.nu[o + j];
            if (std::fabs(d) < 1e-4) d = d >= 0.0 ? 1e-4 : -1e-4;
            s += 2.0 * std::fabs(rm.DipoT[o + j]) / std::fabs(rm.DipoT[o + i]) * (1.0 / d) * W[j * n + i];
        }
        const double Y = s;
        if (Y == 0.0) continue;
        const double nu = rm.nu[o + i];
        const double gd = kCTGAMD * nu * std::sqrt(T / M);
        const double gl = rm.gV_air[o + i] * std::pow(kT0 / T, (double)rm.n_air[o + i]) * p;
        const double f = IRS_SQRT_LN2 / std::max(gd, 1e-10);
        const double s0t = 1.0 - std::exp(-kCT * nu / kT0), stt = 1.0 - std::exp(-kCT * nu / T);
        const double S0 = rm.DipoT[o + i] * rm.DipoT[o + i] * rm.PopuT0[o + i] * nu * s0t;
        const double ST = S0 * qr * std::exp(-kCT * rm.E_lower[o + i] * (1.0 / T - 1.0 / kT0)) * stt / s0t;
        if (!(ST > 0.0)) continue;
        LmTerm t{};
        t.c = nu + (double)rm.shift[o + i] * p;
        t.f = f;
        t.y = gl * f;
        t.a = 0.0;
        t.b = Y * p * ST * f * IRS_INV_SQRT_PI;
        t.fp64 = 0;
        out.push_back(t);
    }
}

// eigenmodes (line_mixing.jl:892-970) minus the no-coupling limit (line_mixing.jl:1030-1091)
std::string vpw_terms(const RelmatHost &rm, const RelmatBandHost &b, double T, double p, double keep,
                      std::vector<LmTerm> &out) {
    const int n = (int)b.n;
    const int64_t o = b.line0;
    if (n == 0) return "";
    std::vector<double> W, rho;
    relmat_build_W(rm, b, T, W, rho);
    std::vector<cd> d(n), Mx((size_t)n * n), lam, X;
    for (int i = 0; i < n; ++i) {
        const double nu = rm.nu[o + i];
        d[i] = rm.DipoT[o + i] * std::sqrt(nu * (1.0 - std::exp(-kCT * nu / T)));
        for (int j = 0; j < n; ++j)
            Mx[(size_t)i * n + j] = i == j ? cd(nu, p * W[(size_t)i * n + j]) : cd(0.0, p * W[(size_t)i * n + j]);
    }
    if (!complex_eig(n, Mx, lam, X)) return "VP_W eigen-decomposition did not converge";
    std::vector<cd> aa(n, cd(0.0)), bb(n);
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) aa[k] += d[i] * X[(size_t)i * n + k];
    for (int j = 0; j < n; ++j) bb[j] = rho[j] * d[j];
    std::vector<cd> Xc = X;
    if (!complex_solve(n, Xc, bb)) return "VP_W eigenvectors are singular";
    double nsum = 0.0;
    for (int i = 0; i < n; ++i) nsum += rm.nu[o + i];
    const double gd = kCTGAMD * (nsum / n) * std::sqrt(T / lm_mass(b.molecule, b.isot));
    const double f = IRS_SQRT_LN2 / std::max(gd, 1e-10);
    std::vector<cd> A(n);
    double amax = 0.0;
    for (int k = 0; k < n; ++k) {
        A[k] = aa[k] * bb[k];
        if (lam[k].imag() > 0.0) amax = std::max(amax, std::abs(A[k]));
    }
    for (int k = 0; k < n; ++k) {
        if (!(lam[k].imag() > 0.0)) continue;
        if (keep > 0.0 && std::abs(A[k]) < keep * amax) continue;
        LmTerm t{};
        t.c = lam[k].real();
        t.f = f;
        t.y = lam[k].imag() * f;
        t.a = A[k].real() * f * IRS_INV_SQRT_PI;
        t.b = A[k].imag() * f * IRS_INV_SQRT_PI;
        t.fp64 = 1;
        out.push_back(t);
    }
    const double qr = q_ratio(b, T);
    for (int i = 0; i < n; ++i) {
        const double nu = rm.nu[o + i];
        const double gl = rm.gV_air[o + i] * std::pow(kT0 / T, (double)rm.n_air[o + i]) * p;
        const double pop = rm.PopuT0[o + i] * qr * std::exp(-kCT * rm.E_lower[o + i] * (1.0 / T - 1.0 / kT0));
        const double ST = rm.DipoT[o + i] * rm.DipoT[o + i] * pop * nu * (1.0 - std::exp(-kCT * nu / T));
        LmTerm t{};
        t.c = nu;
        t.f = f;
        t.y = gl * f;
        t.a = -ST * f * IRS_INV_SQRT_PI;
        t.b = 0.0;
        t.fp64 = 1;
        out.push_back(t);
    }
    return "";
}

}  // namespace

std::string relmat_terms(const RelmatHost &rm, const uint8_t *band_mode, double T, double p_atm,
                         double keep_threshold, std::vector<LmTerm> &out) {
    out.clear();
    if (!(T > 0.0)) return "T must be > 0";
    for (size_t bi = 0; bi < rm.bands.size(); ++bi) {
        const int mode = band_mode ? band_mode[bi] : IRS_RELMAT_VPY;
        const RelmatBandHost &b = rm.bands[bi];
        if (mode == IRS_RELMAT_SKIP || b.table < 0 || b.li > 8) continue;
        if (mode == IRS_RELMAT_VPY) {
            vpy_terms(rm, b, T, p_atm, out);
        } else if (mode == IRS_RELMAT_VPW) {
            std::string e = vpw_terms(rm, b, T, p_atm, keep_threshold, out);
            if (!e.empty()) return e;
        } else {
            return "relmat_band_mode must be 0, 1 or 2";
        }
    }
    std::stable_sort(out.begin(), out.end(), [](const LmTerm &x, const LmTerm &y) { return x.c < y.c; });
    return "";
}

extern "C" int irs_relmat_terms(const irs_relmat *rm, const uint8_t *band_mode, double T, double p_atm,
                                double keep_threshold, int64_t capacity, double *terms, int64_t *n_terms) {
    if (!rm || !n_terms || (capacity > 0 && !terms)) return IRS_EINVAL;
    RelmatHost h;
    if (!relmat_build(rm, h).empty()) return IRS_EINVAL;
    std::vector<LmTerm> t;
    if (!relmat_terms(h, band_mode, T, p_atm, keep_threshold, t).empty()) return IRS_EINVAL;
    *n_terms = (int64_t)t.size();
    for (int64_t k = 0; k < std::min<int64_t>(capacity, (int64_t)t.size()); ++k) {
        terms[5 * k + 0] = t[k].c;
        terms[5 * k + 1] = t[k].f;
        terms[5 * k + 2] = t[k].y;
        terms[5 * k + 3] = t[k].a;
        terms[5 * k + 4] = t[k].b;
    }
    return IRS_OK;
}

// ---- device: term sums ---------------------------------------------------------------
#define LMT_THREADS 256
#define LMT_P 4
#define LMT_CHUNK 256

__global__ void __launch_bounds__(LMT_THREADS) lm_terms_kernel(const LmTermArgs a) {
    __shared__ LmTerm s_t[LMT_CHUNK];
    __shared__ int64_t s_rng[2];
    const int t = threadIdx.x;
    const int u = blockIdx.y;
    const int64_t i0 = (int64_t)blockIdx.x * (LMT_THREADS * LMT_P);
    const int64_t i1 = min(i0 + (int64_t)(LMT_THREADS * LMT_P), a.n);
    const int64_t t0 = a.off[u], t1 = a.off[u + 1];
    if (t == 0) {
        // terms with c in [nu(i0) - cutoff, nu(i1 - 1) + cutoff]: the window test below
        // decides per point; the bounds only narrow the loop
        const double lo = (a.an + (double)i0 * a.sn) / a.den - a.cutoff - 1e-9;
        const double hi = (a.an + (double)(i1 - 1) * a.sn) / a.den + a.cutoff + 1e-9;
        int64_t l = t0, h = t1;
        while (l < h) {
            const int64_t m = (l + h) >> 1;
            if (a.terms[m].c < lo) l = m + 1; else h = m;
        }
        int64_t l2 = l, h2 = t1;
        while (l2 < h2) {
            const int64_t m = (l2 + h2) >> 1;
            if (a.terms[m].c <= hi) l2 = m + 1; else h2 = m;
        }
        s_rng[0] = l;
        s_rng[1] = l2;
    }
    __syncthreads();
    const int64_t ja = s_rng[0], jb = s_rng[1];
    double nu[LMT_P], acc[LMT_P];
#pragma unroll
    for (int p = 0; p < LMT_P; ++p) {
        const int64_t i = i0 + t + LMT_THREADS * p;
        nu[p] = (a.an + (double)i * a.sn) / a.den;
        acc[p] = 0.0;
    }
    const float xf = (float)a.x_far;
    for (int64_t c0 = ja; c0 < jb; c0 += LMT_CHUNK) {
        const int cn = (int)min((int64_t)LMT_CHUNK, jb - c0);
        __syncthreads();
        if (t < cn) s_t[t] = a.terms[c0 + t];
        __syncthreads();
        for (int k = 0; k < cn; ++k) {
            const LmTerm tm = s_t[k];
            if (tm.fp64) {
#pragma unroll
                for (int p = 0; p < LMT_P; ++p) {
                    const double dn = nu[p] - tm.c;
                    if (fabs(dn) > a.cutoff) continue;
                    const double x = dn * tm.f;
                    double re, im;
                    if (fabs(x) > a.x_far) {
                        const double q = IRS_INV_SQRT_PI / (x * x + tm.y * tm.y);
                        re = tm.y * q;
                        im = x * q;
                    } else {
                        w64(x, tm.y, re, im);
                    }
                    acc[p] += tm.a * re + tm.b * im;
                }
            } else {
                const float y = (float)tm.y, fa = (float)tm.a, fb = (float)tm.b;
#pragma unroll
                for (int p = 0; p < LMT_P; ++p) {
                    const double dn = nu[p] - tm.c;
                    if (fabs(dn) > a.cutoff) continue;
                    const float x = (float)(dn * tm.f);
                    float re, im;
                    if (fabsf(x) > xf) {
                        const float q = 0.56418958354775628f / fmaf(x, x, y * y);
                        re = y * q;
                        im = x * q;
                    } else {
                        w32(x, y, re, im);
                    }
                    acc[p] += (double)fmaf(fa, re, fb * im);
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < LMT_P; ++p) {
        const int64_t i = i0 + t + LMT_THREADS * p;
        if (i >= i1) continue;
        const size_t ix = (size_t)u * (size_t)a.n + (size_t)i;
        if (a.out64) {
            a.out64[ix] = acc[p];
        } else {
            const double v = (double)a.tau_lm[ix] + a.coef[u] * acc[p];
            a.tau[ix] += (float)(v > 0.0 ? v : 0.0);
        }
    }
}

hipError_t launch_lm_terms(const LmTermArgs &a, hipStream_t s) {
    if (a.n <= 0 || a.n_units <= 0) return hipSuccess;
    const int64_t tiles = (a.n + LMT_THREADS * LMT_P - 1) / (LMT_THREADS * LMT_P);
    hipLaunchKernelGGL(lm_terms_kernel, dim3((unsigned)tiles, (unsigned)a.n_units), dim3(LMT_THREADS), 0, s, a);
    return hipGetLastError();
}
