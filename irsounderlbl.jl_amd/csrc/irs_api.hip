// C ABI of libirs (include/irs.h): context, line-list upload, layer
// properties and the forward-model orchestration on one HIP stream.
//
// Host work per profile is O(levels x species) FP64 arithmetic restated from
// src/Atmosphere/layers.jl:10-178 and src/Utils/interpolation.jl:10-39; all
// O(lines), O(grid) and O(grid x layers) work runs in the HIP kernels of
// k_prep.hip, k_linesum.hip and k_rte_ils.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/irs.h"
#include "../data/mtckd_table.h"
#include "../data/tips2024_table.h"
#include "irs_cont.h"
#include "irs_internal.h"
#include "irs_jac.h"
#include "irs_relmat.h"

hipError_t launch_keys(const double *nu, int64_t n, double nu_lo, double dnu, int32_t *keys,
                       hipStream_t s);

// Device scratch of irs_xsec_layer_noalloc: every device buffer of one
// cross-section evaluation (list descriptor, work item, counters, keys, line
// records, pedestal sums, the FP32 output) laid out in this fixed array, so
// that the call allocates nothing.  Each GPU holds its own copy for the life of
// the process; host code uses the address hipGetSymbolAddress returns for the
// current device (irs_ctx::scratch), never the array's name.
static __device__ __align__(256) unsigned char scratch[IRS_NOALLOC_SCRATCH_BYTES];

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct ListHost {
    bool used = false;
    int64_t n = 0;
    LineListDev dev{};
    void *block = nullptr;          // one device allocation for all fields
    double nu_min = 0, nu_max = 0, max_abs_dair = 0;
    double dop_coef = 0;            // max over lines of the Doppler coefficient / sqrt(T)
    bool has_lm = false;
    std::vector<double> nu_host;    // for host-side geometry
};

// Grow-only pinned host buffer (staging for graph-captured copies).
struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Everything the device sequence of one profile group depends on.  Graphs are
// cached by the byte image of this struct (zero-filled before use).
struct GroupPlan {
    const void *items, *rec0, *rec1, *rec2, *rec3, *keys, *ktab, *tau, *rhi, *counters, *queue, *lists, *tips, *order;
    const void *prof, *rch, *bt, *chidx, *chw, *chnu, *taps, *h_items, *h_prof, *h_out, *pre, *pblk;
    int64_t n, nch, half, n_items, max_lines, items_bytes, prof_bytes, n_rec, ktab_lo, ktab_n;
    int64_t ch_idx0, ch_stride;    // equally spaced exact-node channels (stride 0: general ILS path)
    int gn, L, ns, P, pblocks, tiles, source_function, with_ils, far_d, far_hc;
    double nu_lo, dnu, an, sn, den;
    int mode;                      // 0 radiance + BT, 1 channel optical depths (2: + CIM weights),
                                   // 3 level channel transmittances + CIM sums
    const void *chod, *chlev, *chg;
    const void *chrows, *chsums;   // banded channel-OD scratch of one profile (exact-node channels)
    const void *chmom, *chmsums;   // ILS first moments [gn][L][nch] and their tap sums + moment taps
    double chod_floor;
    int continua;                  // IRS_CONT_* added to tau after the line sum (0: none)
    int cia_gen;                   // CIA table generation (graphs capture the table pointers)
    // relaxation-matrix line mixing: the line sum runs once without the line-mixing
    // species (items_lm1) and once with it alone (items_lm2, into tau_lm); the term sum
    // then adds max(tau_lm + coef * dsigma, 0) to tau.  Never captured in a graph.
    const void *items_lm1, *items_lm2, *tau_lm, *lm_terms, *lm_off, *lm_coef;
    double lm_cutoff, x_far;
};

struct GraphEntry {
    GroupPlan plan;
    hipGraphExec_t exec = nullptr;
};

struct TapsKey {
    double dnu, opd, fwhm;
    int apod, n_pts;
    bool operator<(const TapsKey &o) const {
        return std::tie(dnu, opd, fwhm, apod, n_pts) < std::tie(o.dnu, o.opd, o.fwhm, o.apod, o.n_pts);
    }
};

}  // namespace

struct irs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;
    std::vector<ListHost> lists;
    DevBuf d_lists, d_tips, d_items, d_rec0, d_rec1, d_rec2, d_rec3, d_keys, d_ktab, d_tau, d_rhi, d_counters;
    int64_t ktab_lo = 0, ktab_n = 0;   // key tables built by build_keys (0: none)
    DevBuf d_tmp0, d_tmp1, d_tmp2, d_tmp3, d_tmp4, d_tmp5, d_chidx, d_chw, d_chnu, d_rch, d_bt;
    DevBuf j_items, j_g, j_tau, j_dt, j_v, j_rte, j_cols, j_desc, j_w, j_K, j_fac;
    std::map<TapsKey, DevBuf> taps;
    irs_stats stats{};
    irs_schedule sched{4, 1, 0, 0, 1.5, 0, 0};      // persistent heaviest-first queue: 26.7 vs 27.9 ms per full-IASI step
    std::map<std::string, std::pair<int, int>> tuned;     // shape -> (P, persistent)
    std::map<std::string, DevBuf> orders;                 // shape -> heaviest-first tile order
    std::map<int, int> capacity;                          // P -> resident persistent workgroups
    DevBuf d_queue, d_prof, d_chod, d_chlev, d_chg, d_pre;
    DevBuf d_pblk;                     // pedestal scan block sums [sum_items ceil(n/256)][3]
    DevBuf d_chrows, d_chsums;         // channel optical depths on exact-node channels: per-point rows, tap sums
    DevBuf d_chmom, d_chmsums;         // ILS first moments of the level transmittance steps
    DevBuf d_rec64, d_tau64;           // every-pair FP64 path: line records, FP64 optical depths
    DevBuf d_ckd;                      // MT-CKD tables: H2O [2001][4], then CO2 [5003][2] = {S*xfac, tdep}
    DevBuf d_cia[3];                   // per CIA pair: T, numin, numax [nb], off [nb+1], nu, sig [knots]
    int cia_nb[3] = {0, 0, 0};
    int64_t cia_nk[3] = {0, 0, 0};
    int cia_gen = 0;
    unsigned char *scratch = nullptr;  // this device's copy of the no-allocation scratch array
    std::vector<std::unique_ptr<RelmatHost>> relmats;   // uploaded relaxation matrices (host data)
    DevBuf d_items_lm, d_tau_lm, d_lm_terms, d_lm_off, d_lm_coef, d_lm_out;
    PinBuf h_items, h_prof, h_out;
    std::vector<GraphEntry> graphs;
};

namespace {

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            return fail(ctx, IRS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                            \
    } while (0)

int fail(irs_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}


hipError_t ensure(DevBuf &b, size_t bytes) {
    if (bytes <= b.cap) return hipSuccess;
    b.release();
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.cap = want;
    return hipSuccess;
}

// ---- grids with Julia float-range semantics (src/Utils/wavenumber_grid.jl:23-27)
struct Grid {
    double a = 0, step = 0, an = 0, sn = 0, den = 1;
    bool rat = false;
    int64_t n = 0;
    double at(int64_t i) const { return rat ? (an + (double)i * sn) / den : a + (double)i * step; }
};

Grid make_grid(double a, double step, double b) {
    Grid g;
    g.a = a;
    g.step = step;
    double x = step, h0 = 0, h1 = 1, k0 = 1, k1 = 0;
    for (int it = 0; it < 40; ++it) {
        double ai = std::floor(x);
        double h2 = ai * h1 + h0, k2 = ai * k1 + k0;
        h0 = h1; h1 = h2; k0 = k1; k1 = k2;
        if (k1 > 1e9) break;
        if (std::fabs(h1 / k1 - step) <= 2.220446049250313e-16 * std::fabs(step)) {
            double ad = a * k1, bd = b * k1;
            if (h1 != 0.0 && std::fabs(ad - std::nearbyint(ad)) <= 1e-7 * std::max(1.0, std::fabs(ad)) &&
                std::fabs(bd - std::nearbyint(bd)) <= 1e-7 * std::max(1.0, std::fabs(bd))) {
                g.an = std::nearbyint(ad);
                g.sn = h1;
                g.den = k1;
                g.rat = true;
                g.n = (int64_t)std::floor((std::nearbyint(bd) - g.an) / g.sn) + 1;
                return g;
            }
            break;
        }
        double fr = x - ai;
        if (fr == 0.0) break;
        x = 1.0 / fr;
    }
    g.n = (int64_t)std::floor((b - a) / step * (1.0 + 1e-14)) + 1;
    g.an = a;
    g.sn = step;
    g.den = 1.0;
    return g;
}

// internal grid of forward_model (src/Sensor/sounder.jl:155-161,263-271)
Grid internal_grid(const irs_sounder *snd, const irs_fm_options *opt, int *hrf_out) {
    int hrf;
    if (opt->high_res_factor > 0) hrf = opt->high_res_factor;
    else if (opt->internal_dnu > 0.0) hrf = std::max(1, (int)std::ceil(snd->dnu / opt->internal_dnu));
    else hrf = 4;
    double d = snd->dnu / hrf;
    double pad = opt->with_ils ? std::ceil(16.0 / d) * d : 0.0;
    if (hrf_out) *hrf_out = hrf;
    return make_grid(snd->nu_min - pad, d, snd->nu_max + pad);
}

// ---- atmosphere (src/Atmosphere/layers.jl, src/Utils/interpolation.jl)
bool approx_eq(double a, double b) { return std::fabs(a - b) <= 1.4901161193847656e-08 * std::max(std::fabs(a), std::fabs(b)); }

double cg_column_vmr(double v1, double v2, double p1, double p2) {   // layers.jl:36-45
    if (v1 <= 0.0 || v2 <= 0.0) return 0.5 * (v1 + v2);
    if (approx_eq(v1, v2)) return v1;
    double al = std::log(v2 / v1) / std::log(p2 / p1), r = p2 / p1, be = al + 1.0;
    if (std::fabs(be) < 1e-8) return (v1 - v2) / std::log(v1 / v2);
    return v1 * p1 * (1.0 - std::pow(r, be)) / (be * std::fabs(p1 - p2));
}

double cg_pressure(double v1, double v2, double p1, double p2) {     // layers.jl:60-72
    if (v1 <= 0.0 || v2 <= 0.0) return 0.5 * (p1 + p2);
    if (approx_eq(v1, v2)) return 0.5 * (p1 + p2);
    double al = std::log(v2 / v1) / std::log(p2 / p1), r = p2 / p1, b1 = al + 1.0, b2 = al + 2.0;
    if (std::fabs(b1) < 1e-8) return std::fabs(p1 - p2) / std::log(p1 / p2);
    if (std::fabs(b2) < 1e-8) return p2 * std::log(p2 / p1) / (p2 / p1 - 1.0);
    return p1 * b1 / b2 * (1.0 - std::pow(r, b2)) / (1.0 - std::pow(r, b1));
}

double cg_temperature_mass(double T1, double T2, double p1, double p2) {   // layers.jl:93-106
    if (approx_eq(T1, T2)) return T1;
    if (approx_eq(p1, p2)) return 0.5 * (T1 + T2);
    double al = std::log(p2 / p1), frac;
    if (std::fabs(al) < 1e-3) frac = 0.5 + al / 12.0;
    else {
        double be = p2 / p1;
        frac = ((al - 1.0) * be + 1.0) / (al * (be - 1.0));
    }
    return T1 + (T2 - T1) * frac;
}

// linear in log p with Flat extrapolation (interpolation.jl:10-24)
double interp_logp(const double *p, const double *x, int n, double pt) {
    std::vector<double> lp(n), xv(n);
    bool asc = true;
    for (int k = 0; k < n; ++k) lp[k] = std::log(p[k]);
    for (int k = 1; k < n; ++k) if (lp[k] < lp[k - 1]) asc = false;
    for (int k = 0; k < n; ++k) {
        int s = asc ? k : n - 1 - k;
        lp[k] = std::log(p[s]);
        xv[k] = x[s];
    }
    double q = std::log(pt);
    if (q <= lp[0]) return xv[0];
    if (q >= lp[n - 1]) return xv[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (lp[mid] <= q) lo = mid; else hi = mid;
    }
    double w = (q - lp[lo]) / (lp[hi] - lp[lo]);
    return xv[lo] * (1.0 - w) + xv[hi] * w;
}

int tips_row_of(int mol, int iso) {
    for (int r = 0; r < IRS_TIPS_NISO; ++r)
        if (irs_tips_keys[r][0] == mol && irs_tips_keys[r][1] == iso) return r;
    return -1;
}

double mol_mass_h(int mol) {
    static const double m[11] = {18.010, 44.010, 47.998, 44.013, 28.010, 16.043, 31.999, 34.015, 64.065, 36.461, 17.031};
    return (mol < 1 || mol > 11) ? 30.0 : m[mol - 1];
}

double lm_mass_h(int mol, int iso) {
    static const double co2[11] = {47.00427, 43.98983, 44.99318, 45.99398, 44.99403, 45.99706,
                                   44.99706, 47.99832, 47.00134, 48.00196, 47.00427};
    static const double ch4[5] = {44.0, 16.031300, 17.034655, 17.037475, 18.040830};
    if (mol == 2 && iso >= 0 && iso <= 10) return co2[iso];
    if (mol == 6 && iso >= 1 && iso <= 4) return ch4[iso];
    return 44.0;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// One item per (layer, species) of a profile.
struct LayerProps {
    std::vector<double> p_mid, dp, T_ave, T_mid;
    std::vector<std::vector<double>> vmr_cg, p_cg, T_cg, vmr_mid;
};

int layer_properties(irs_ctx *ctx, const irs_profile *pr, int T_method, LayerProps &lp) {
    const int nlev = pr->n_levels, L = nlev - 1;
    if (nlev < 2) return fail(ctx, IRS_EINVAL, "profile needs at least 2 levels");
    if (T_method != IRS_TMETHOD_LOGP_AT_PCG && T_method != IRS_TMETHOD_MASS_WEIGHTED)
        return fail(ctx, IRS_EINVAL, "T_method must be :logp_at_pcg or :mass_weighted");
    const double *p = pr->p_hpa, *T = pr->T;
    for (int k = 0; k < nlev; ++k)
        if (!(p[k] > 0.0) || !(T[k] > 0.0)) return fail(ctx, IRS_EINVAL, "non-positive pressure or temperature");
    lp.p_mid.resize(L);
    lp.dp.resize(L);
    lp.T_ave.resize(L);
    for (int k = 0; k < L; ++k) {
        lp.p_mid[k] = 0.5 * (p[k] + p[k + 1]);
        lp.dp[k] = std::fabs(p[k] - p[k + 1]);
        lp.T_ave[k] = cg_temperature_mass(T[k], T[k + 1], p[k], p[k + 1]);
    }
    lp.T_mid.resize(L);
    for (int k = 0; k < L; ++k) lp.T_mid[k] = interp_logp(p, T, nlev, lp.p_mid[k]);   // layers.jl:142
    lp.vmr_cg.assign(pr->n_species, std::vector<double>(L));
    lp.p_cg.assign(pr->n_species, std::vector<double>(L));
    lp.T_cg.assign(pr->n_species, std::vector<double>(L));
    lp.vmr_mid.assign(pr->n_species, std::vector<double>(L));
    for (int s = 0; s < pr->n_species; ++s) {
        const double *v = pr->vmr + (size_t)s * nlev;
        for (int k = 0; k < L; ++k) {
            lp.vmr_mid[s][k] = interp_logp(p, v, nlev, lp.p_mid[k]);                      // layers.jl:145-148
            lp.vmr_cg[s][k] = cg_column_vmr(v[k], v[k + 1], p[k], p[k + 1]);
            lp.p_cg[s][k] = cg_pressure(v[k], v[k + 1], p[k], p[k + 1]);
            lp.T_cg[s][k] = T_method == IRS_TMETHOD_MASS_WEIGHTED ? lp.T_ave[k]
                                                                  : interp_logp(p, T, nlev, lp.p_cg[s][k]);
        }
    }
    return IRS_OK;
}

// Continuum state of each layer (sounder.jl:311-315; optical_depth.jl:123-130):
// mid-layer VMRs (H2O 0 and CO2 415 ppm when the profile has no such species),
// p_mid, T_mid, dz and the mid-layer weight of the upper level.
void cont_layers(const irs_profile *pr, const LayerProps &lp, ContLayer *out) {
    const int L = pr->n_levels - 1;
    int sh = -1, sc = -1;
    for (int s = 0; s < pr->n_species; ++s) {
        if (pr->species[s] == 1) sh = s;
        if (pr->species[s] == 2) sc = s;
    }
    for (int k = 0; k < L; ++k) {
        ContLayer &c = out[k];
        std::memset(&c, 0, sizeof(c));
        c.vh2o = sh >= 0 ? lp.vmr_mid[sh][k] : 0.0;
        c.vco2 = sc >= 0 ? lp.vmr_mid[sc][k] : 4.15e-4;
        c.vn2 = 0.78084 * (1.0 - c.vh2o);                 // sounder.jl:315,319-320
        c.vo2 = 0.20946 * (1.0 - c.vh2o);
        c.p_mid = lp.p_mid[k];
        c.T_mid = lp.T_mid[k];
        const double H = 8.314462 * c.T_mid / (0.028964 * 9.80665);
        c.dz = H * lp.dp[k] / c.p_mid * 100.0;
        c.wM = std::log(lp.p_mid[k] / pr->p_hpa[k]) / std::log(pr->p_hpa[k + 1] / pr->p_hpa[k]);
    }
}

// Continuum launch arguments for the context's tables; the caller sets tau,
// the grid, the layer states and the optional derivative outputs.
ContArgs cont_args(const irs_ctx *ctx, int continua) {
    ContArgs a{};
    a.continua = continua;
    a.h2o = (const double *)ctx->d_ckd.p;
    a.co2 = a.h2o + (size_t)IRS_CKD_H2O_N * 4;
    for (int pr = 0; pr < 3; ++pr) {
        CiaDev &c = a.cia[pr];
        c.nb = ctx->cia_nb[pr];
        if (!c.nb) continue;
        const double *b = (const double *)ctx->d_cia[pr].p;
        c.T = b;
        c.numin = b + c.nb;
        c.numax = b + 2 * c.nb;
        c.off = (const int64_t *)(b + 3 * c.nb);
        c.nu = b + 4 * c.nb + 1;
        c.sig = c.nu + ctx->cia_nk[pr];
    }
    return a;
}

int check_list(irs_ctx *ctx, int32_t h) {
    if (h < 0 || h >= (int32_t)ctx->lists.size() || !ctx->lists[h].used)
        return fail(ctx, IRS_ENOTFOUND, "unknown line-list handle " + std::to_string(h));
    return IRS_OK;
}

// Fill one work item.
ItemDev make_item(const ListHost &lh, int32_t list, double T, double p_atm, double vmr_self, double coef,
                  double dptmn, double cutoff, double lm_cutoff, double x_far, bool lm, double dnu) {
    ItemDev it{};
    it.n = (int32_t)lh.n;
    it.list = list;
    it.lm = lm ? 1 : 0;
    it.active = lh.n > 0 ? 1 : 0;
    double shift = (lh.max_abs_dair + 0.003) * p_atm;   // |delta_air| + |delta_self| bound
    it.M = (int32_t)std::ceil(shift / dnu) + 2;
    it.C = (float)(cutoff / dnu);
    it.Clm = (float)(lm_cutoff / dnu);
    it.kcore_per_nu = x_far * lh.dop_coef * std::sqrt(T) / (IRS_SQRT_LN2 * dnu) * (1.0 + 1e-6);
    it.T = T;
    it.p_atm = p_atm;
    it.vmr_self = vmr_self;
    it.coef = coef;
    it.dptmn = dptmn;
    it.cutoff = cutoff;
    it.x_far = x_far;
    it.lm_cutoff = lm_cutoff;
    return it;
}

int ensure_common(irs_ctx *ctx) {
    if (!ctx->d_lists.p || ctx->d_lists.cap < sizeof(LineListDev) * std::max<size_t>(1, ctx->lists.size())) {
        HIPCHK(ensure(ctx->d_lists, sizeof(LineListDev) * std::max<size_t>(16, 2 * ctx->lists.size())));
    }
    std::vector<LineListDev> v(ctx->lists.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = ctx->lists[i].dev;
    if (!v.empty())
        HIPCHK(hipMemcpyAsync(ctx->d_lists.p, v.data(), sizeof(LineListDev) * v.size(), hipMemcpyHostToDevice,
                              ctx->stream));
    return IRS_OK;
}

// Keys for each used list; returns per-list key offsets.  With n_hi > 0 it
// also builds each list's key table over v in [-OFF, n_hi + OFF], OFF = the
// larger window in grid steps + 64 (the line sum's window bounds fall in that
// range unless the pressure shift exceeds 62 grid steps; outside it the line
// sum searches the keys), and returns the tables' offsets.
int build_keys(irs_ctx *ctx, const std::vector<int32_t> &lists, double nu_lo, double dnu,
               std::map<int32_t, int32_t> &key_off, int64_t n_hi = 0, double cut_max = 0.0,
               std::map<int32_t, int64_t> *ktab_off = nullptr) {
    int64_t tot = 0;
    for (int32_t h : lists)
        if (!key_off.count(h)) {
            key_off[h] = (int32_t)tot;
            tot += ctx->lists[h].n;
        }
    HIPCHK(ensure(ctx->d_keys, sizeof(int32_t) * std::max<int64_t>(tot, 1)));
    for (auto &kv : key_off) {
        const ListHost &lh = ctx->lists[kv.first];
        HIPCHK(launch_keys(lh.dev.nu, lh.n, nu_lo, dnu, (int32_t *)ctx->d_keys.p + kv.second, ctx->stream));
    }
    ctx->ktab_lo = ctx->ktab_n = 0;
    if (ktab_off && n_hi > 0 && cut_max / dnu < 1.0e8) {
        const int64_t off = (int64_t)std::ceil(cut_max / dnu) + 64;
        const int64_t cnt = n_hi + 2 * off + 1;
        HIPCHK(ensure(ctx->d_ktab, sizeof(int32_t) * (size_t)cnt * key_off.size()));
        int64_t o = 0;
        for (auto &kv : key_off) {
            (*ktab_off)[kv.first] = o;
            HIPCHK(launch_keytab((const int32_t *)ctx->d_keys.p + kv.second, (int)ctx->lists[kv.first].n, -off,
                                 cnt, (int32_t *)ctx->d_ktab.p + o, ctx->stream));
            o += cnt;
        }
        ctx->ktab_lo = -off;
        ctx->ktab_n = cnt;
    }
    return IRS_OK;
}

hipError_t ensure_pinned(PinBuf &b, size_t bytes) {
    if (bytes <= b.cap) return hipSuccess;
    b.release();
    const size_t want = std::max<size_t>(bytes, 4096);
    hipError_t e = hipHostMalloc(&b.p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.cap = want;
    return hipSuccess;
}

std::string shape_key(const std::vector<const ItemDev *> &per_species, int64_t n_hi, double nu_lo, double dnu,
                      int P) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%lld|%.17g|%.17g|%d", (long long)n_hi, nu_lo, dnu, P);
    std::string k(buf);
    for (const ItemDev *it : per_species) {
        std::snprintf(buf, sizeof buf, "|%d:%d:%d:%.9g:%.9g", it->list, it->lm, it->n, (double)it->C,
                      (double)it->Clm);
        k += buf;
    }
    return k;
}

// Key of a tuned schedule that outlives the process: the device (architecture and
// compute-unit count), the grid, and per species the line list's content (count and
// an FNV-1a hash of the line centres, not the session's list handle), line mixing and
// the window widths.
std::string persistent_tune_key(const irs_ctx *ctx, const std::vector<const ItemDev *> &per_species, int64_t n_hi,
                                double nu_lo, double dnu) {
    hipDeviceProp_t pr{};
    std::string dev = "unknown";
    if (hipGetDeviceProperties(&pr, ctx->device) == hipSuccess)
        dev = std::string(pr.gcnArchName) + ":" + std::to_string(pr.multiProcessorCount);
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s|%lld|%.17g|%.17g", dev.c_str(), (long long)n_hi, nu_lo, dnu);
    std::string k(buf);
    for (const ItemDev *it : per_species) {
        const ListHost &lh = ctx->lists[it->list];
        uint64_t h = 1469598103934665603ull;
        const unsigned char *b = (const unsigned char *)lh.nu_host.data();
        for (size_t i = 0; i < lh.nu_host.size() * sizeof(double); ++i) h = (h ^ b[i]) * 1099511628211ull;
        std::snprintf(buf, sizeof buf, "|%lld:%016llx:%d:%.9g:%.9g", (long long)lh.n, (unsigned long long)h, it->lm,
                      (double)it->C, (double)it->Clm);
        k += buf;
    }
    for (char &c : k)
        if (c == '\t' || c == '\n') c = ' ';
    return k;
}

// Tuned schedules kept in the file named by IRS_TUNE_FILE (one "key<TAB>P<TAB>persistent"
// line per shape), so a new process skips the timing runs.
bool tune_file_lookup(const std::string &key, int &P, int &pers) {
    const char *path = std::getenv("IRS_TUNE_FILE");
    if (!path || !*path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    bool found = false;
    std::string line;
    int c;
    while (true) {
        line.clear();
        while ((c = std::fgetc(f)) != EOF && c != '\n') line.push_back((char)c);
        const size_t t1 = line.find('\t');
        const size_t t2 = t1 == std::string::npos ? t1 : line.find('\t', t1 + 1);
        if (t2 != std::string::npos && line.compare(0, t1, key) == 0 && t1 == key.size()) {
            const int p = std::atoi(line.c_str() + t1 + 1), q = std::atoi(line.c_str() + t2 + 1);
            if ((p == 4 || p == 8 || p == 16) && (q == 0 || q == 1)) {
                P = p;
                pers = q;
                found = true;          // the last entry for a key wins
            }
        }
        if (c == EOF) break;
    }
    std::fclose(f);
    return found;
}

void tune_file_store(const std::string &key, int P, int pers) {
    const char *path = std::getenv("IRS_TUNE_FILE");
    if (!path || !*path) return;
    if (FILE *f = std::fopen(path, "a")) {
        std::fprintf(f, "%s\t%d\t%d\n", key.c_str(), P, pers);
        std::fclose(f);
    }
}

// Heaviest-first order of the spectral tiles: the cost of a tile is the number
// of lines whose window reaches it, weighted 1.6 on line-mixing lists (the
// dispersive term), from host copies of the line centres.
int tile_order(irs_ctx *ctx, const std::vector<const ItemDev *> &per_species, int64_t n_hi, double nu_lo,
               double dnu, int P, const int32_t **order) {
    const std::string key = shape_key(per_species, n_hi, nu_lo, dnu, P);
    auto f = ctx->orders.find(key);
    if (f != ctx->orders.end()) {
        *order = (const int32_t *)f->second.p;
        return IRS_OK;
    }
    const int64_t W = IRS_LS_THREADS * (int64_t)P;
    const int tiles = (int)((n_hi + W - 1) / W);
    std::vector<double> cost(tiles, 0.0);
    for (const ItemDev *it : per_species) {
        const ListHost &lh = ctx->lists[it->list];
        const int64_t C = (int64_t)std::ceil(std::max((double)it->C, it->lm ? (double)it->Clm : 0.0)) + it->M;
        // histogram of line keys over [-C, n_hi + C)
        const int64_t span = n_hi + 2 * C + 1;
        std::vector<int32_t> pre(span + 1, 0);
        for (double nu : lh.nu_host) {
            const double k = std::floor((nu - nu_lo) / dnu);
            if (k < (double)-C || k >= (double)(n_hi + C)) continue;
            pre[(int64_t)k + C + 1] += 1;
        }
        for (int64_t i = 1; i <= span; ++i) pre[i] += pre[i - 1];
        const double w = it->lm ? 1.6 : 1.0;
        for (int t = 0; t < tiles; ++t) {
            const int64_t a = t * W, b = std::min(a + W, n_hi);   // lines with key in [a - C, b + C)
            cost[t] += w * (double)(pre[b + 2 * C] - pre[a]);
        }
    }
    std::vector<int32_t> ord(tiles);
    for (int t = 0; t < tiles; ++t) ord[t] = t;
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    DevBuf &b = ctx->orders[key];
    HIPCHK(ensure(b, sizeof(int32_t) * tiles));
    HIPCHK(hipMemcpy(b.p, ord.data(), sizeof(int32_t) * tiles, hipMemcpyHostToDevice));
    *order = (const int32_t *)b.p;
    return IRS_OK;
}

// Coarse far-wing parameters for a grid step and tile: node spacing HC (a power
// of two dividing the tile, HC dnu <= 0.032 cm^-1) and distance D (grid steps).
void far_params(const irs_ctx *ctx, double dnu, int P, int &far_d, int &far_hc) {
    far_d = far_hc = 0;
    const double fw = ctx->sched.far_wing_cm;
    if (!(fw > 0.0)) return;
    int hc = 1;
    while (hc * 2 <= IRS_LS_THREADS * P / 8 && (hc * 2) * dnu <= 0.032 * (1.0 + 1e-9)) hc *= 2;
    if (hc < 2) return;
    far_hc = hc;
    far_d = (int)std::ceil(fw / dnu);
}

int persistent_blocks(irs_ctx *ctx, int P) {
    auto f = ctx->capacity.find(P);
    if (f != ctx->capacity.end()) return f->second;
    const int c = linesum_persistent_capacity(P);
    ctx->capacity[P] = c;
    return c;
}

// prep + line sum for a set of items laid out [n_prof][n_layers][n_species]
// (eager; used by the layer-level export).
int run_linesum(irs_ctx *ctx, std::vector<ItemDev> &items, int n_prof, int n_layers, int n_species,
                double nu_lo, double dnu, int64_t n_hi, float *out) {
    int64_t rec = 0;
    int max_lines = 0;
    int64_t pb = 0;
    for (auto &it : items) {
        it.rec_off = rec;
        rec += it.n;
        it.pb_off = pb;
        pb += (it.n + 255) / 256;
        max_lines = std::max(max_lines, it.n);
    }
    HIPCHK(ensure(ctx->d_pblk, sizeof(double) * 3 * std::max<int64_t>(pb, 1)));
    HIPCHK(ensure(ctx->d_rec0, sizeof(float4) * std::max<int64_t>(rec, 1)));
    HIPCHK(ensure(ctx->d_rec1, sizeof(float4) * std::max<int64_t>(rec, 1)));
    HIPCHK(ensure(ctx->d_rec2, sizeof(float4) * std::max<int64_t>(rec, 1)));
    HIPCHK(ensure(ctx->d_rec3, sizeof(float4) * 6 * std::max<int64_t>(rec, 1)));
    HIPCHK(ensure(ctx->d_items, sizeof(ItemDev) * std::max<size_t>(items.size(), 1)));
    HIPCHK(ensure(ctx->d_counters, sizeof(unsigned long long) * IRS_NCNT));
    HIPCHK(ensure(ctx->d_queue, 64));
    HIPCHK(ensure(ctx->d_pre, sizeof(double) * 3 * std::max<int64_t>(rec, 1)));
    HIPCHK(hipMemcpyAsync(ctx->d_items.p, items.data(), sizeof(ItemDev) * items.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(launch_reset_counters((unsigned *)ctx->d_queue.p, (unsigned long long *)ctx->d_counters.p, IRS_NCNT,
                                 ctx->stream));
    PrepArgs pa{};
    pa.items = (const ItemDev *)ctx->d_items.p;
    pa.lists = (const LineListDev *)ctx->d_lists.p;
    pa.tips_q = (const double *)ctx->d_tips.p;
    pa.n_items = (int)items.size();
    pa.nu_lo = nu_lo;
    pa.dnu = dnu;
    pa.rec0 = (float4 *)ctx->d_rec0.p;
    pa.rec1 = (float4 *)ctx->d_rec1.p;
    pa.rec2 = (float4 *)ctx->d_rec2.p;
    pa.rec3 = (float4 *)ctx->d_rec3.p;
    pa.counters = (unsigned long long *)ctx->d_counters.p;
    pa.n_hi = n_hi;
    pa.pre = (double *)ctx->d_pre.p;
    pa.pblk = (double *)ctx->d_pblk.p;
    pa.n_rec = std::max<int64_t>(rec, 1);
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(launch_prep(pa, max_lines, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    LinesumArgs la{};
    la.items = (const ItemDev *)ctx->d_items.p;
    la.rec0 = pa.rec0;
    la.rec1 = pa.rec1;
    la.rec2 = pa.rec2;
    la.keys = (const int32_t *)ctx->d_keys.p;
    la.ktab = ctx->ktab_n > 0 ? (const int32_t *)ctx->d_ktab.p : nullptr;
    la.ktab_lo = ctx->ktab_lo;
    la.ktab_n = ctx->ktab_n;
    la.out = out;
    la.n_hi = n_hi;
    la.n_layers = n_layers;
    la.n_species = n_species;
    la.nu_lo = nu_lo;
    la.dnu = dnu;
    la.n_prof = n_prof;
    const int P = ctx->sched.tile_points;
    la.tiles = (int)((n_hi + (int64_t)IRS_LS_THREADS * P - 1) / ((int64_t)IRS_LS_THREADS * P));
    la.tile_order = nullptr;
    la.queue = (unsigned *)ctx->d_queue.p;
    far_params(ctx, dnu, P, la.far_d, la.far_hc);
    la.work = (unsigned long long *)ctx->d_counters.p + 3;
    la.pre = pa.pre;
    la.n_rec = pa.n_rec;
    la.rec3 = pa.rec3;
    HIPCHK(launch_linesum(la, P, ctx->sched.persistent ? persistent_blocks(ctx, P) : 0, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    ctx->stats.linesum_launches += 1;
    ctx->stats.tile_points = P;
    ctx->stats.persistent = ctx->sched.persistent;
    return IRS_OK;
}

void read_counters(irs_ctx *ctx) {
    unsigned long long cs[IRS_NCNT], c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (hipMemcpy(cs, ctx->d_counters.p, sizeof(cs), hipMemcpyDeviceToHost) == hipSuccess) {
        for (int i = 0; i < IRS_NCNT; ++i) c[i % 8] += cs[i];
        ctx->stats.active_lines += (int64_t)c[0];
        ctx->stats.active_pairs += (int64_t)c[1];
        ctx->stats.dispersive_pairs += (int64_t)c[2];
        ctx->stats.fine_pairs += (int64_t)c[4];
        ctx->stats.coarse_pairs += (int64_t)c[5];
        ctx->stats.fine_pairs_lm += (int64_t)c[6];
        ctx->stats.coarse_pairs_lm += (int64_t)c[7];
        ctx->stats.expanded_line_slices += (int64_t)c[3];
    }
}

int get_taps(irs_ctx *ctx, double dnu, double opd, double fwhm, int apod, int n_pts, double **taps_dev,
             int64_t *half_out) {
    if (apod < 0 || apod > 3)
        return fail(ctx, IRS_EINVAL, "Unknown apodization; expected :gaussian or a Norton-Beer style");
    if (n_pts < 2) return fail(ctx, IRS_EINVAL, "n_pts must be >= 2");
    if (!(dnu > 0.0)) return fail(ctx, IRS_EINVAL, "dnu must be > 0");
    const int64_t half = (int64_t)std::ceil(16.0 / dnu);
    *half_out = half;
    TapsKey key{dnu, opd, fwhm, apod, n_pts};
    auto itm = ctx->taps.find(key);
    if (itm != ctx->taps.end()) {
        *taps_dev = (double *)itm->second.p;
        return IRS_OK;
    }
    // OPD-domain taper (ils.jl:86-97), taps on the device, area normalisation here (:110-111)
    static const double NB[3][3] = {{0.384093, 0.087577, 0.528330},
                                    {0.152442, 0.136176, 0.711382},
                                    {0.045335, 0.554883, 0.399782}};
    std::vector<double> A(n_pts), Lm(n_pts);
    const double dL = opd / (n_pts - 1);
    const double sig = fwhm / (2.0 * std::sqrt(2.0 * std::log(2.0)));
    for (int m = 0; m < n_pts; ++m) {
        double t = (double)m / (double)(n_pts - 1);
        Lm[m] = t * opd;
        if (apod == IRS_APOD_GAUSSIAN) A[m] = std::exp(-2.0 * IRS_PI * IRS_PI * sig * sig * Lm[m] * Lm[m]);
        else {
            double x = Lm[m] / opd;
            if (std::fabs(x) > 1.0) A[m] = 0.0;
            else {
                const double *c = NB[apod - 1];
                double tt = 1.0 - x * x;
                A[m] = c[0] + c[1] * tt * tt + c[2] * tt * tt * tt * tt;
            }
        }
    }
    const int64_t nt = 2 * half + 1;
    HIPCHK(ensure(ctx->d_tmp0, sizeof(double) * n_pts));
    HIPCHK(ensure(ctx->d_tmp1, sizeof(double) * n_pts));
    DevBuf buf;
    HIPCHK(ensure(buf, sizeof(double) * nt));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp0.p, A.data(), sizeof(double) * n_pts, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp1.p, Lm.data(), sizeof(double) * n_pts, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_ils_taps((const double *)ctx->d_tmp0.p, (const double *)ctx->d_tmp1.p, n_pts, dL, dnu, half,
                           (double *)buf.p, ctx->stream));
    std::vector<double> h(nt);
    HIPCHK(hipMemcpyAsync(h.data(), buf.p, sizeof(double) * nt, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    for (int64_t k = 0; k < nt; ++k) tot += h[k];
    const double norm = tot * dnu;
    for (int64_t k = 0; k < nt; ++k) h[k] /= norm;
    HIPCHK(hipMemcpy(buf.p, h.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
    ctx->taps[key] = buf;
    *taps_dev = (double *)buf.p;
    return IRS_OK;
}

// channel sample positions on an internal grid (src/Sensor/sounder.jl:372-377)
void channel_positions(const Grid &gi, const Grid &gc, std::vector<int64_t> &idx, std::vector<double> &w,
                       std::vector<double> &nu) {
    idx.resize(gc.n);
    w.resize(gc.n);
    nu.resize(gc.n);
    for (int64_t c = 0; c < gc.n; ++c) {
        double v = gc.at(c);
        nu[c] = v;
        double x = (v - gi.at(0)) / gi.step;
        int64_t k = (int64_t)std::llround(x);
        if (k >= 0 && k < gi.n && gi.at(k) == v) { idx[c] = k; w[c] = 0.0; continue; }
        if (v <= gi.at(0)) { idx[c] = 0; w[c] = 0.0; continue; }
        if (v >= gi.at(gi.n - 1)) { idx[c] = gi.n - 1; w[c] = 0.0; continue; }
        k = (int64_t)std::floor(x);
        while (k > 0 && gi.at(k) > v) --k;
        while (k + 1 < gi.n - 1 && gi.at(k + 1) <= v) ++k;
        idx[c] = k;
        w[c] = (v - gi.at(k)) / (gi.at(k + 1) - gi.at(k));
    }
}

// Stride of the channel nodes when every channel is an exact internal node and
// the nodes are equally spaced (the case of every forward_model grid); 0 otherwise.
int64_t channel_stride(const std::vector<int64_t> &idx, const std::vector<double> &w) {
    if (idx.empty()) return 0;
    for (double x : w)
        if (x != 0.0) return 0;
    if (idx.size() == 1) return 1;
    const int64_t S = idx[1] - idx[0];
    if (S < 1 || S > (1 << 20)) return 0;
    for (size_t c = 1; c < idx.size(); ++c)
        if (idx[c] - idx[c - 1] != S) return 0;
    return S;
}

double cg_mass_frac(double p1, double p2) {   // schwarzschild_jacobian.jl:47-53
    if (approx_eq(p1, p2)) return 0.5;
    double al = std::log(p2 / p1);
    if (std::fabs(al) < 1e-3) return 0.5 + al / 12.0;
    double be = p2 / p1;
    return ((al - 1.0) * be + 1.0) / (al * (be - 1.0));
}

void cg_column_vmr_grad(double v1, double v2, double p1, double p2, double &d1, double &d2) {   // vmr_jacobian.jl:81-107
    if (v1 <= 0.0 || v2 <= 0.0) { d1 = 0.5; d2 = 0.5; return; }
    if (approx_eq(v1, v2)) { double fr = cg_mass_frac(p1, p2); d1 = 1.0 - fr; d2 = fr; return; }
    double Lg = std::log(p2 / p1), al = std::log(v2 / v1) / Lg, be = al + 1.0, dp = std::fabs(p1 - p2);
    if (std::fabs(be) < 1e-6) {
        double h1 = std::max(std::fabs(v1), 1e-300) * 1e-6, h2 = std::max(std::fabs(v2), 1e-300) * 1e-6;
        d1 = (cg_column_vmr(v1 + h1, v2, p1, p2) - cg_column_vmr(v1 - h1, v2, p1, p2)) / (2 * h1);
        d2 = (cg_column_vmr(v1, v2 + h2, p1, p2) - cg_column_vmr(v1, v2 - h2, p1, p2)) / (2 * h2);
        return;
    }
    double r = p2 / p1, rb = std::pow(r, be);
    double c = v1 * p1 * (1.0 - rb) / (be * dp);
    double J = (v1 * p1 / Lg) * ((rb - 1.0) / (be * be) - rb * Lg / be);
    d1 = c / v1 - J / (v1 * dp);
    d2 = J / (v2 * dp);
}

double dbt_dr(double nu, double R) {   // vmr_jacobian.jl:120-125
    if (R <= 0.0) return 0.0;
    double A = 1.191042953e-5 * nu * nu * nu, Lg = std::log1p(A / R);
    return 1.4387769 * nu * A / (Lg * Lg * R * (R + A));
}

}  // namespace

// TIPS-2024 Q(T), linear in T between the 1 K nodes, T / 296 without a table row
// (src/HITRAN/partition.jl:10-26); the relaxation-matrix host code's Q ratios.
double relmat_partition(int mol, int iso, double T) {
    const int r = tips_row_of(mol, iso);
    if (r < 0) return T / 296.0;
    const double *v = irs_tips_q[r];
    long lo = (long)std::floor(T) - IRS_TIPS_T_MIN, hi = lo + 1;
    lo = std::min<long>(std::max<long>(lo, 0), IRS_TIPS_NT - 1);
    hi = std::min<long>(std::max<long>(hi, 0), IRS_TIPS_NT - 1);
    const double fr = T - std::floor(T);
    return v[lo] * (1.0 - fr) + v[hi] * fr;
}

hipStream_t ctx_stream(irs_ctx *ctx) { return ctx->stream; }
int ctx_device(irs_ctx *ctx) { return ctx->device; }
int ctx_fail(irs_ctx *ctx, int code, const std::string &msg) { return fail(ctx, code, msg); }

// =============================================================================
extern "C" {

int irs_version(void) { return IRS_ABI_VERSION; }

void irs_default_options(irs_fm_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->internal_dnu = 0.001;
    o->high_res_factor = 0;
    o->cutoff = 25.0;
    o->lm_cutoff = 25.0;
    o->dptmn = 1e-6;
    o->x_far = 122.0;
    o->with_ils = 1;
    o->apodization = IRS_APOD_GAUSSIAN;
    o->line_mixing = IRS_LM_NONE;
    o->T_method = IRS_TMETHOD_LOGP_AT_PCG;
    o->source_function = IRS_SRC_CIM;
    o->continua = IRS_CONT_ALL;      // apply_continuum = true, all five continua (sounder.jl:245-246)
    o->relmat = -1;
    o->lm_species = 2;
    o->relmat_band_mode = nullptr;
    o->keep_threshold = 1e-4;        // compute_vpw_band_perturbation default (line_mixing.jl:1069)
}

int irs_device_info(int device, char *buf, int buflen) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return IRS_ENODEV;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return IRS_EHIP;
    if (buf && buflen > 0)
        std::snprintf(buf, buflen, "%s;%s;%d;%d;%.1f", p.name, p.gcnArchName, p.multiProcessorCount,
                      p.clockRate / 1000, p.totalGlobalMem / 1073741824.0);
    return IRS_OK;
}

int irs_create(int device, irs_ctx **out) {
    if (!out) return IRS_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return IRS_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return IRS_ENODEV;
    irs_ctx *ctx = new irs_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return IRS_EHIP;
    }
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            irs_destroy(ctx);
            return IRS_EHIP;
        }
    const size_t tb = sizeof(double) * IRS_TIPS_NISO * IRS_TIPS_NT;
    if (ensure(ctx->d_tips, tb) != hipSuccess ||
        hipMemcpy(ctx->d_tips.p, irs_tips_q, tb, hipMemcpyHostToDevice) != hipSuccess) {
        irs_destroy(ctx);
        return IRS_ENOMEM;
    }
    {
        std::vector<double> ck((size_t)IRS_CKD_H2O_N * 4 + (size_t)IRS_CKD_CO2_N * 2);
        for (int r = 0; r < IRS_CKD_H2O_N; ++r)
            for (int c = 0; c < 4; ++c) ck[(size_t)r * 4 + c] = irs_ckd_h2o[r][c];
        double *c2 = ck.data() + (size_t)IRS_CKD_H2O_N * 4;
        for (int r = 0; r < IRS_CKD_CO2_N; ++r) {
            c2[2 * r] = irs_ckd_co2[r][1] * irs_ckd_co2[r][2];   // S x xfac (continuum.jl:90)
            c2[2 * r + 1] = irs_ckd_co2[r][3];
        }
        if (ensure(ctx->d_ckd, sizeof(double) * ck.size()) != hipSuccess ||
            hipMemcpy(ctx->d_ckd.p, ck.data(), sizeof(double) * ck.size(), hipMemcpyHostToDevice) != hipSuccess) {
            irs_destroy(ctx);
            return IRS_ENOMEM;
        }
    }
    // the device address of the no-allocation scratch array on this device
    if (hipGetSymbolAddress((void **)&ctx->scratch, HIP_SYMBOL(scratch)) != hipSuccess || !ctx->scratch) {
        irs_destroy(ctx);
        return IRS_EHIP;
    }
    if (const char *e = std::getenv("IRS_LS_P")) {
        int v = std::atoi(e);
        if (v == 4 || v == 8 || v == 16) ctx->sched.tile_points = v;
    }
    *out = ctx;
    return IRS_OK;
}

void irs_destroy(irs_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &l : ctx->lists)
        if (l.block) (void)hipFree(l.block);
    DevBuf *bufs[] = {&ctx->d_lists, &ctx->d_tips, &ctx->d_items, &ctx->d_rec0, &ctx->d_rec1, &ctx->d_rec2, &ctx->d_rec3,
                      &ctx->d_keys, &ctx->d_ktab, &ctx->d_tau, &ctx->d_rhi, &ctx->d_counters, &ctx->d_tmp0, &ctx->d_tmp1,
                      &ctx->d_tmp2, &ctx->d_tmp3, &ctx->d_tmp4, &ctx->d_tmp5, &ctx->d_chidx, &ctx->d_chw,
                      &ctx->d_chnu, &ctx->d_rch, &ctx->d_bt, &ctx->j_items, &ctx->j_g, &ctx->j_tau,
                      &ctx->j_dt, &ctx->j_v, &ctx->j_rte, &ctx->j_cols, &ctx->j_desc, &ctx->j_w, &ctx->j_K,
                      &ctx->j_fac};
    for (DevBuf *b : bufs) b->release();
    for (auto &kv : ctx->taps) kv.second.release();
    for (auto &kv : ctx->orders) kv.second.release();
    for (auto &g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->d_queue.release();
    ctx->d_prof.release();
    ctx->d_chod.release();
    ctx->d_pre.release();
    ctx->d_chlev.release();
    ctx->d_chg.release();
    ctx->d_chrows.release();
    ctx->d_chsums.release();
    ctx->d_chmom.release();
    ctx->d_chmsums.release();
    ctx->d_ckd.release();
    ctx->d_pblk.release();
    ctx->d_rec64.release();
    ctx->d_tau64.release();
    for (DevBuf *b : {&ctx->d_items_lm, &ctx->d_tau_lm, &ctx->d_lm_terms, &ctx->d_lm_off, &ctx->d_lm_coef, &ctx->d_lm_out})
        b->release();
    for (auto &b : ctx->d_cia) b.release();
    ctx->h_items.release();
    ctx->h_prof.release();
    ctx->h_out.release();
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *irs_last_error(const irs_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int irs_last_stats(const irs_ctx *ctx, irs_stats *out) {
    if (!ctx || !out) return IRS_EINVAL;
    *out = ctx->stats;
    return IRS_OK;
}

int irs_linelist_upload(irs_ctx *ctx, const irs_linelist *ll, int32_t *handle) {
    if (!ctx || !ll || !handle) return fail(ctx, IRS_EINVAL, "null argument");
    (void)hipSetDevice(ctx->device);
    const int64_t n = ll->n;
    if (n < 0 || n > 2000000000) return fail(ctx, IRS_EINVAL, "bad line count");
    if (n > 0 && (!ll->mol || !ll->iso || !ll->nu || !ll->S || !ll->gamma_air || !ll->gamma_self ||
                  !ll->elower || !ll->n_air || !ll->delta_air))
        return fail(ctx, IRS_EINVAL, "missing line-list field");
    for (int64_t j = 1; j < n; ++j)
        if (ll->nu[j] < ll->nu[j - 1])
            return fail(ctx, IRS_EINVAL, "line list must be sorted by wavenumber (linelist.jl:52)");
    ListHost lh;
    lh.used = true;
    lh.n = n;
    std::vector<int16_t> rows(n);
    double dmax = 0.0, dop = 0.0;
    bool lm = false;
    for (int64_t j = 0; j < n; ++j) {
        rows[j] = (int16_t)tips_row_of(ll->mol[j], ll->iso[j]);
        dmax = std::max(dmax, std::fabs((double)ll->delta_air[j]));
        dop = std::max(dop, 3.58126e-7 / std::sqrt(mol_mass_h(ll->mol[j])));
        if (ll->y_lm && ll->y_lm[j] != 0.0) {
            lm = true;
            dop = std::max(dop, 3.5812e-7 / std::sqrt(lm_mass_h(ll->mol[j], ll->iso[j])));
        }
    }
    lh.max_abs_dair = dmax;
    lh.dop_coef = dop;
    lh.has_lm = lm;
    lh.nu_min = n ? ll->nu[0] : 0.0;
    lh.nu_max = n ? ll->nu[n - 1] : 0.0;
    // one device block: [nu S elower ylm](f64) [gair gself nair dair](f32) [tips_row](i16) [mol iso](i8)
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t bytes = nn * (4 * 8 + 4 * 4 + 2 + 2) + 256;
    if (hipMalloc(&lh.block, bytes) != hipSuccess) return fail(ctx, IRS_ENOMEM, "line-list allocation failed");
    char *b = (char *)lh.block;
    double *d_nu = (double *)b, *d_S = d_nu + nn, *d_el = d_S + nn, *d_y = d_el + nn;
    float *d_ga = (float *)(d_y + nn), *d_gs = d_ga + nn, *d_na = d_gs + nn, *d_da = d_na + nn;
    int16_t *d_row = (int16_t *)(d_da + nn);
    int8_t *d_mol = (int8_t *)(d_row + nn), *d_iso = d_mol + nn;
    std::vector<double> ybuf;
    const double *ysrc = ll->y_lm;
    if (!ysrc) {
        ybuf.assign(nn, 0.0);
        ysrc = ybuf.data();
    }
    if (n > 0) {
        HIPCHK(hipMemcpy(d_nu, ll->nu, 8 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_S, ll->S, 8 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_el, ll->elower, 8 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_y, ysrc, 8 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_ga, ll->gamma_air, 4 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_gs, ll->gamma_self, 4 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_na, ll->n_air, 4 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_da, ll->delta_air, 4 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_row, rows.data(), 2 * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_mol, ll->mol, n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_iso, ll->iso, n, hipMemcpyHostToDevice));
    }
    lh.dev.n = n;
    lh.dev.nu = d_nu;
    lh.dev.S = d_S;
    lh.dev.elower = d_el;
    lh.dev.ylm = lm ? d_y : nullptr;
    lh.dev.gair = d_ga;
    lh.dev.gself = d_gs;
    lh.dev.nair = d_na;
    lh.dev.dair = d_da;
    lh.dev.tips_row = d_row;
    lh.dev.mol = d_mol;
    lh.dev.iso = d_iso;
    lh.nu_host.assign(ll->nu, ll->nu + n);
    int32_t h = -1;
    for (size_t i = 0; i < ctx->lists.size(); ++i)
        if (!ctx->lists[i].used) { h = (int32_t)i; break; }
    if (h < 0) {
        h = (int32_t)ctx->lists.size();
        ctx->lists.push_back(ListHost());
    }
    ctx->lists[h] = std::move(lh);
    *handle = h;
    return IRS_OK;
}

int irs_linelist_free(irs_ctx *ctx, int32_t handle) {
    if (!ctx) return IRS_EINVAL;
    int rc = check_list(ctx, handle);
    if (rc) return rc;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->lists[handle].block) (void)hipFree(ctx->lists[handle].block);
    ctx->lists[handle] = ListHost();
    return IRS_OK;
}

int irs_relmat_upload(irs_ctx *ctx, const irs_relmat *rm, int32_t *handle) {
    if (!ctx || !rm || !handle) return fail(ctx, IRS_EINVAL, "null argument");
    std::unique_ptr<RelmatHost> h(new RelmatHost);
    std::string e = relmat_build(rm, *h);
    if (!e.empty()) return fail(ctx, IRS_EINVAL, e);
    for (size_t i = 0; i < ctx->relmats.size(); ++i)
        if (!ctx->relmats[i]) {
            ctx->relmats[i] = std::move(h);
            *handle = (int32_t)i;
            return IRS_OK;
        }
    ctx->relmats.push_back(std::move(h));
    *handle = (int32_t)ctx->relmats.size() - 1;
    return IRS_OK;
}

int irs_relmat_free(irs_ctx *ctx, int32_t handle) {
    if (!ctx) return IRS_EINVAL;
    if (handle < 0 || handle >= (int32_t)ctx->relmats.size() || !ctx->relmats[handle])
        return fail(ctx, IRS_EINVAL, "bad relaxation-matrix handle");
    ctx->relmats[handle].reset();
    return IRS_OK;
}

// Uploads per-unit term lists (units in order) and their offsets.
static int upload_lm_terms(irs_ctx *ctx, const std::vector<std::vector<LmTerm>> &units) {
    std::vector<int64_t> off(units.size() + 1, 0);
    for (size_t u = 0; u < units.size(); ++u) off[u + 1] = off[u] + (int64_t)units[u].size();
    std::vector<LmTerm> flat((size_t)std::max<int64_t>(off.back(), 1));
    for (size_t u = 0; u < units.size(); ++u)
        std::copy(units[u].begin(), units[u].end(), flat.begin() + off[u]);
    HIPCHK(ensure(ctx->d_lm_terms, sizeof(LmTerm) * flat.size()));
    HIPCHK(ensure(ctx->d_lm_off, sizeof(int64_t) * off.size()));
    HIPCHK(hipMemcpyAsync(ctx->d_lm_terms.p, flat.data(), sizeof(LmTerm) * flat.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_lm_off.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));     // the host vectors go out of scope
    return IRS_OK;
}

static int check_relmat(irs_ctx *ctx, int32_t h) {
    if (h < 0 || h >= (int32_t)ctx->relmats.size() || !ctx->relmats[h])
        return fail(ctx, IRS_EINVAL, "bad relaxation-matrix handle");
    return IRS_OK;
}

int irs_lm_perturbation(irs_ctx *ctx, int32_t relmat, const uint8_t *band_mode, double nu0, double dnu,
                        int64_t n_nu, double T, double p_atm, double lm_cutoff, double x_far,
                        double keep_threshold, double *dsigma) {
    if (!ctx || !dsigma || n_nu <= 0 || !(dnu > 0.0)) return fail(ctx, IRS_EINVAL, "bad arguments");
    (void)hipSetDevice(ctx->device);
    int rc = check_relmat(ctx, relmat);
    if (rc) return rc;
    std::vector<std::vector<LmTerm>> units(1);
    std::string e = relmat_terms(*ctx->relmats[relmat], band_mode, T, p_atm, keep_threshold, units[0]);
    if (!e.empty()) return fail(ctx, IRS_EINVAL, e);
    rc = upload_lm_terms(ctx, units);
    if (rc) return rc;
    HIPCHK(ensure(ctx->d_lm_out, sizeof(double) * (size_t)n_nu));
    const Grid g = make_grid(nu0, dnu, nu0 + (double)(n_nu - 1) * dnu);
    LmTermArgs a{};
    a.terms = (const LmTerm *)ctx->d_lm_terms.p;
    a.off = (const int64_t *)ctx->d_lm_off.p;
    a.n = n_nu;
    a.n_units = 1;
    a.an = g.an;
    a.sn = g.sn;
    a.den = g.den;
    a.cutoff = lm_cutoff;
    a.x_far = x_far;
    a.out64 = (double *)ctx->d_lm_out.p;
    HIPCHK(launch_lm_terms(a, ctx->stream));
    HIPCHK(hipMemcpyAsync(dsigma, ctx->d_lm_out.p, sizeof(double) * (size_t)n_nu, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

int irs_faddeeva32(irs_ctx *ctx, const double *x, const double *y, int64_t n, int mode, double *w_re,
                   double *w_im) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (!x || !y || !w_re || !w_im || n < 0) return fail(ctx, IRS_EINVAL, "bad arguments");
    if (mode != 0 && mode != 1) return fail(ctx, IRS_EINVAL, "mode must be 0 (core rule) or 1 (pair rule)");
    for (int64_t i = 0; i < n; ++i)
        if (!(y[i] >= 0.0) || !std::isfinite(x[i]) || !std::isfinite(y[i]))
            return fail(ctx, IRS_EINVAL, "w(z) needs finite x and y >= 0");
    if (n == 0) return IRS_OK;
    const size_t b = sizeof(double) * (size_t)n;
    HIPCHK(ensure(ctx->d_tmp2, 4 * b));
    double *d = (double *)ctx->d_tmp2.p;
    HIPCHK(hipMemcpyAsync(d, x, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d + n, y, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_faddeeva32(d, d + n, n, mode, d + 2 * n, d + 3 * n, ctx->stream));
    HIPCHK(hipMemcpyAsync(w_re, d + 2 * n, b, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(w_im, d + 3 * n, b, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

int irs_xsec_layer(irs_ctx *ctx, int32_t handle, double nu0, double dnu, int64_t n_nu, double T, double p_atm,
                   double vmr_self, double coef, double dptmn, double cutoff, double lm_cutoff, double x_far,
                   int line_mixing, double *sigma) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    int rc = check_list(ctx, handle);
    if (rc) return rc;
    if (!sigma || n_nu <= 0 || !(dnu > 0.0)) return fail(ctx, IRS_EINVAL, "bad grid");
    if (!(T > 0.0) || !(p_atm >= 0.0)) return fail(ctx, IRS_EINVAL, "bad T or p");
    if (!(cutoff > 0.0) || !(lm_cutoff > 0.0)) return fail(ctx, IRS_EINVAL, "cutoff must be > 0");
    if (line_mixing != IRS_LM_NONE && line_mixing != IRS_LM_VPY) return fail(ctx, IRS_EINVAL, "bad line_mixing");
    auto t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    const ListHost &lh = ctx->lists[handle];
    const bool lm = line_mixing == IRS_LM_VPY && lh.has_lm;
    // The VP_Y path evaluates the Voigt part without self terms and without
    // DPTMIN (line_mixing.jl:1351-1369, sounder.jl:295-299).
    const double vs = lm ? 0.0 : vmr_self;
    // records carry coef; a power-of-two scale keeps FP32 values far from denormals
    const double scale = 1152921504606846976.0;   // 2^60
    if (!(coef > 0.0)) return fail(ctx, IRS_EINVAL, "coef must be > 0");
    // the DPTMIN test uses the caller's coef; the scale is applied to the records only
    ItemDev it = make_item(lh, handle, T, p_atm, vs, coef * scale, dptmn, cutoff, lm_cutoff, x_far, lm, dnu);
    it.dptmn = dptmn * scale;
    std::vector<ItemDev> items{it};
    rc = ensure_common(ctx);
    if (rc) return rc;
    std::map<int32_t, int32_t> koff;
    std::map<int32_t, int64_t> ktoff;
    rc = build_keys(ctx, {handle}, nu0, dnu, koff, n_nu, std::max(cutoff, lm_cutoff), &ktoff);
    if (rc) return rc;
    items[0].key_off = koff[handle];
    items[0].ktab_off = ktoff.count(handle) ? ktoff[handle] : 0;
    HIPCHK(ensure(ctx->d_tau, sizeof(float) * n_nu));
    rc = run_linesum(ctx, items, 1, 1, 1, nu0, dnu, n_nu, (float *)ctx->d_tau.p);
    if (rc) return rc;
    std::vector<float> h(n_nu);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->d_tau.p, sizeof(float) * n_nu, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double inv = 1.0 / (coef * scale);
    for (int64_t i = 0; i < n_nu; ++i) sigma[i] = (double)h[i] * inv;
    read_counters(ctx);
    ctx->stats.ms_prep = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->stats.ms_linesum = elapsed_ms(ctx->ev[1], ctx->ev[2]);
    ctx->stats.n_hi = n_nu;
    ctx->stats.n_layers = 1;
    ctx->stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return IRS_OK;
}

int irs_xsec_layer_noalloc(irs_ctx *ctx, int32_t handle, double nu0, double dnu, int64_t n_nu, double T, double p_atm,
                           double vmr_self, double coef, double dptmn, double cutoff, double lm_cutoff, double x_far,
                           int line_mixing, double *sigma) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    int rc = check_list(ctx, handle);
    if (rc) return rc;
    if (!sigma || n_nu <= 0 || !(dnu > 0.0)) return fail(ctx, IRS_EINVAL, "bad grid");
    if (!(T > 0.0) || !(p_atm >= 0.0)) return fail(ctx, IRS_EINVAL, "bad T or p");
    if (!(cutoff > 0.0) || !(lm_cutoff > 0.0)) return fail(ctx, IRS_EINVAL, "cutoff must be > 0");
    if (line_mixing != IRS_LM_NONE && line_mixing != IRS_LM_VPY) return fail(ctx, IRS_EINVAL, "bad line_mixing");
    if (!(coef > 0.0)) return fail(ctx, IRS_EINVAL, "coef must be > 0");
    if (!ctx->scratch) return fail(ctx, IRS_EHIP, "no scratch array on this device");
    const ListHost &lh = ctx->lists[handle];
    const int64_t n = lh.n;
    // scratch layout, every part 256-byte aligned
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_list = 0, o_item = o_list + al(sizeof(LineListDev)), o_cnt = o_item + al(sizeof(ItemDev));
    const size_t o_queue = o_cnt + al(sizeof(unsigned long long) * IRS_NCNT), o_keys = o_queue + 256;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t o_r0 = o_keys + al(sizeof(int32_t) * nn), o_r1 = o_r0 + al(sizeof(float4) * nn);
    const size_t o_r2 = o_r1 + al(sizeof(float4) * nn), o_r3 = o_r2 + al(sizeof(float4) * nn);
    const size_t o_pre = o_r3 + al(sizeof(float4) * 6 * nn), o_pblk = o_pre + al(sizeof(double) * 3 * nn);
    const size_t o_tau = o_pblk + al(sizeof(double) * 3 * ((nn + 255) / 256));
    const size_t need = o_tau + al(sizeof(float) * (size_t)n_nu);
    if (need > (size_t)IRS_NOALLOC_SCRATCH_BYTES)
        return fail(ctx, IRS_EINVAL, "irs_xsec_layer_noalloc: " + std::to_string(n) + " lines and " +
                                         std::to_string(n_nu) + " grid points need " + std::to_string(need) +
                                         " bytes of scratch, more than its " +
                                         std::to_string((long long)IRS_NOALLOC_SCRATCH_BYTES) + "; use irs_xsec_layer");
    unsigned char *S = ctx->scratch;
    hipStream_t st = ctx->stream;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    // the same item as irs_xsec_layer (records carry coef x 2^60)
    const bool lm = line_mixing == IRS_LM_VPY && lh.has_lm;
    const double vs = lm ? 0.0 : vmr_self;
    const double scale = 1152921504606846976.0;   // 2^60
    ItemDev it = make_item(lh, 0, T, p_atm, vs, coef * scale, dptmn, cutoff, lm_cutoff, x_far, lm, dnu);
    it.dptmn = dptmn * scale;
    it.key_off = 0;
    it.ktab_off = 0;
    it.rec_off = 0;
    it.pb_off = 0;
    HIPCHK(hipMemcpyAsync(S + o_list, &lh.dev, sizeof(LineListDev), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(S + o_item, &it, sizeof(ItemDev), hipMemcpyHostToDevice, st));
    HIPCHK(launch_reset_counters((unsigned *)(S + o_queue), (unsigned long long *)(S + o_cnt), IRS_NCNT, st));
    HIPCHK(launch_keys(lh.dev.nu, n, nu0, dnu, (int32_t *)(S + o_keys), st));
    PrepArgs pa{};
    pa.items = (const ItemDev *)(S + o_item);
    pa.lists = (const LineListDev *)(S + o_list);
    pa.tips_q = (const double *)ctx->d_tips.p;
    pa.n_items = 1;
    pa.nu_lo = nu0;
    pa.dnu = dnu;
    pa.rec0 = (float4 *)(S + o_r0);
    pa.rec1 = (float4 *)(S + o_r1);
    pa.rec2 = (float4 *)(S + o_r2);
    pa.rec3 = (float4 *)(S + o_r3);
    pa.counters = (unsigned long long *)(S + o_cnt);
    pa.n_hi = n_nu;
    pa.pre = (double *)(S + o_pre);
    pa.pblk = (double *)(S + o_pblk);
    pa.n_rec = (int64_t)nn;
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(launch_prep(pa, (int)n, st));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    LinesumArgs la{};
    la.items = pa.items;
    la.rec0 = pa.rec0;
    la.rec1 = pa.rec1;
    la.rec2 = pa.rec2;
    la.rec3 = pa.rec3;
    la.keys = (const int32_t *)(S + o_keys);
    la.ktab = nullptr;                     // window bounds by binary search in the keys
    la.out = (float *)(S + o_tau);
    la.n_hi = n_nu;
    la.n_layers = 1;
    la.n_species = 1;
    la.nu_lo = nu0;
    la.dnu = dnu;
    la.n_prof = 1;
    const int P = ctx->sched.tile_points;
    la.tiles = (int)((n_nu + (int64_t)IRS_LS_THREADS * P - 1) / ((int64_t)IRS_LS_THREADS * P));
    la.tile_order = nullptr;
    la.queue = (unsigned *)(S + o_queue);
    far_params(ctx, dnu, P, la.far_d, la.far_hc);
    la.work = pa.counters + 3;
    la.pre = pa.pre;
    la.n_rec = pa.n_rec;
    HIPCHK(launch_linesum(la, P, ctx->sched.persistent ? linesum_persistent_capacity(P) : 0, st));
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    // FP32 output into the first half of the caller's FP64 buffer, widened in
    // place from the end (element i's double covers floats 2i and 2i+1, both >= i)
    HIPCHK(hipMemcpyAsync(sigma, S + o_tau, sizeof(float) * (size_t)n_nu, hipMemcpyDeviceToHost, st));
    unsigned long long cs[IRS_NCNT];
    HIPCHK(hipMemcpyAsync(cs, S + o_cnt, sizeof(cs), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double inv = 1.0 / (coef * scale);
    unsigned char *bytes = (unsigned char *)sigma;
    for (int64_t i = n_nu - 1; i >= 0; --i) {
        float v;
        std::memcpy(&v, bytes + 4 * i, sizeof(v));
        const double d = (double)v * inv;
        std::memcpy(bytes + 8 * i, &d, sizeof(d));
    }
    unsigned long long c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < IRS_NCNT; ++i) c[i % 8] += cs[i];
    ctx->stats.active_lines = (int64_t)c[0];
    ctx->stats.active_pairs = (int64_t)c[1];
    ctx->stats.dispersive_pairs = (int64_t)c[2];
    ctx->stats.expanded_line_slices = (int64_t)c[3];
    ctx->stats.fine_pairs = (int64_t)c[4];
    ctx->stats.coarse_pairs = (int64_t)c[5];
    ctx->stats.fine_pairs_lm = (int64_t)c[6];
    ctx->stats.coarse_pairs_lm = (int64_t)c[7];
    ctx->stats.ms_prep = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->stats.ms_linesum = elapsed_ms(ctx->ev[1], ctx->ev[2]);
    ctx->stats.n_hi = n_nu;
    ctx->stats.n_layers = 1;
    ctx->stats.linesum_launches = 1;
    ctx->stats.tile_points = P;
    ctx->stats.persistent = ctx->sched.persistent;
    return IRS_OK;
}

int irs_rte(irs_ctx *ctx, double nu0, double dnu, int64_t n_nu, const double *tau, int n_layers,
            const double *T_lev, double T_sfc, double mu, double eps_sfc, int source_function, const double *T_ave,
            double *radiance) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (!tau || !T_lev || !radiance || n_nu <= 0 || n_layers <= 0)
        return fail(ctx, IRS_EINVAL, "bad RTE arguments");
    if (!(mu > 0.0 && mu <= 1.0)) return fail(ctx, IRS_EINVAL, "μ must be in (0, 1]");
    if (source_function != IRS_SRC_TOON && source_function != IRS_SRC_CIM)
        return fail(ctx, IRS_EINVAL, "source_function must be :toon or :cim");
    if (source_function == IRS_SRC_CIM && !T_ave) return fail(ctx, IRS_EINVAL, ":cim source_function requires T_ave");
    Grid g = make_grid(nu0, dnu, nu0 + dnu * (double)(n_nu - 1));
    const size_t tb = sizeof(double) * (size_t)n_nu * n_layers;
    HIPCHK(ensure(ctx->d_tmp2, tb));
    HIPCHK(ensure(ctx->d_tmp3, sizeof(double) * (n_layers + 1)));
    HIPCHK(ensure(ctx->d_tmp4, sizeof(double) * (n_layers + 1)));
    HIPCHK(ensure(ctx->d_rhi, sizeof(double) * n_nu));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp2.p, tau, tb, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp3.p, T_lev, sizeof(double) * (n_layers + 1), hipMemcpyHostToDevice, ctx->stream));
    if (T_ave)
        HIPCHK(hipMemcpyAsync(ctx->d_tmp4.p, T_ave, sizeof(double) * n_layers, hipMemcpyHostToDevice, ctx->stream));
    RteArgs ra{};
    ra.tau64 = (const double *)ctx->d_tmp2.p;
    ra.n = n_nu;
    ra.n_layers = n_layers;
    ra.T_lev = (const double *)ctx->d_tmp3.p;
    ra.T_ave = (const double *)ctx->d_tmp4.p;
    ra.T_sfc = T_sfc;
    ra.mu = mu;
    ra.eps = eps_sfc;
    ra.source_function = source_function;
    ra.an = g.an;
    ra.sn = g.sn;
    ra.den = g.den;
    ra.out = (double *)ctx->d_rhi.p;
    HIPCHK(launch_rte(ra, ctx->stream));
    HIPCHK(hipMemcpyAsync(radiance, ctx->d_rhi.p, sizeof(double) * n_nu, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

int irs_ils_half(double dnu, int64_t *half) {
    if (!half || !(dnu > 0.0)) return IRS_EINVAL;
    *half = (int64_t)std::ceil(16.0 / dnu);
    return IRS_OK;
}

int irs_ils_kernel(irs_ctx *ctx, double dnu, double opd_max, double fwhm, int apodization, int n_pts,
                   double *taps) {
    if (!ctx || !taps) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    double *d = nullptr;
    int64_t half = 0;
    int rc = get_taps(ctx, dnu, opd_max, fwhm, apodization, n_pts, &d, &half);
    if (rc) return rc;
    HIPCHK(hipMemcpy(taps, d, sizeof(double) * (2 * half + 1), hipMemcpyDeviceToHost));
    return IRS_OK;
}

int irs_ils_channels(irs_ctx *ctx, const double *spectrum, double nu0, double dnu, int64_t n_nu,
                     const irs_sounder *snd, int apodization, double *r_ch) {
    if (!ctx || !spectrum || !snd || !r_ch || n_nu <= 0) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    double *taps = nullptr;
    int64_t half = 0;
    int rc = get_taps(ctx, dnu, snd->opd_max, snd->fwhm_gauss, apodization, 2048, &taps, &half);
    if (rc) return rc;
    Grid gi = make_grid(nu0, dnu, nu0 + dnu * (double)(n_nu - 1));
    gi.n = n_nu;
    Grid gc = make_grid(snd->nu_min, snd->dnu, snd->nu_max);
    std::vector<int64_t> idx;
    std::vector<double> w, nu;
    channel_positions(gi, gc, idx, w, nu);
    HIPCHK(ensure(ctx->d_tmp5, sizeof(double) * n_nu));
    HIPCHK(ensure(ctx->d_chidx, sizeof(int64_t) * gc.n));
    HIPCHK(ensure(ctx->d_chw, sizeof(double) * gc.n));
    HIPCHK(ensure(ctx->d_chnu, sizeof(double) * gc.n));
    HIPCHK(ensure(ctx->d_rch, sizeof(double) * gc.n));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp5.p, spectrum, sizeof(double) * n_nu, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chidx.p, idx.data(), sizeof(int64_t) * gc.n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chw.p, w.data(), sizeof(double) * gc.n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chnu.p, nu.data(), sizeof(double) * gc.n, hipMemcpyHostToDevice, ctx->stream));
    IlsArgs ia{};
    ia.spec = (const double *)ctx->d_tmp5.p;
    ia.n = n_nu;
    ia.taps = taps;
    ia.half = half;
    ia.dnu = dnu;
    ia.ch_idx = (const int64_t *)ctx->d_chidx.p;
    ia.ch_w = (const double *)ctx->d_chw.p;
    ia.ch_nu = (const double *)ctx->d_chnu.p;
    ia.n_ch = gc.n;
    ia.r_out = (double *)ctx->d_rch.p;
    ia.bt_out = nullptr;
    HIPCHK(launch_ils_channels(ia, ctx->stream));
    HIPCHK(hipMemcpyAsync(r_ch, ctx->d_rch.p, sizeof(double) * gc.n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

int irs_brightness_temperature(const double *nu, const double *radiance, int64_t n, double *bt) {
    if (!nu || !radiance || !bt || n < 0) return IRS_EINVAL;
    for (int64_t i = 0; i < n; ++i) {   // planck.jl:43-46
        const double R = radiance[i];
        bt[i] = R <= 0.0 ? 0.0 : 1.4387769 * nu[i] / std::log1p(1.191042953e-5 * nu[i] * nu[i] * nu[i] / R);
    }
    return IRS_OK;
}

int irs_forward_dims(const irs_sounder *snd, const irs_fm_options *opt, double *nu_lo, double *dnu_hi,
                     int64_t *n_hi, int64_t *n_channels) {
    if (!snd || !opt) return IRS_EINVAL;
    if (!(snd->dnu > 0.0) || !(snd->nu_max >= snd->nu_min)) return IRS_EINVAL;
    Grid g = internal_grid(snd, opt, nullptr);
    Grid gc = make_grid(snd->nu_min, snd->dnu, snd->nu_max);
    if (nu_lo) *nu_lo = g.at(0);
    if (dnu_hi) *dnu_hi = g.step;
    if (n_hi) *n_hi = g.n;
    if (n_channels) *n_channels = gc.n;
    return IRS_OK;
}

// Device sequence of one profile group: staging copies, counter reset, prep,
// line sum, RTE, channel ILS, copy-back.  Issues no allocation and no memset,
// so it can be captured in a graph.  Events are recorded only when timing.
static hipError_t enqueue_group(irs_ctx *ctx, const GroupPlan &pl, bool timing) {
    hipStream_t st = ctx->stream;
    hipError_t e;
    const int L = pl.L;
    const size_t NL = (size_t)pl.gn;
#define ENQ(x)                      \
    do {                            \
        e = (x);                    \
        if (e != hipSuccess) return e; \
    } while (0)
    ENQ(hipMemcpyAsync((void *)pl.items, pl.h_items, (size_t)pl.items_bytes, hipMemcpyHostToDevice, st));
    ENQ(hipMemcpyAsync((void *)pl.prof, pl.h_prof, (size_t)pl.prof_bytes, hipMemcpyHostToDevice, st));
    ENQ(launch_reset_counters((unsigned *)pl.queue, (unsigned long long *)pl.counters, IRS_NCNT, st));
    PrepArgs pa{};
    pa.items = (const ItemDev *)pl.items;
    pa.lists = (const LineListDev *)pl.lists;
    pa.tips_q = (const double *)pl.tips;
    pa.n_items = (int)pl.n_items;
    pa.nu_lo = pl.nu_lo;
    pa.dnu = pl.dnu;
    pa.rec0 = (float4 *)pl.rec0;
    pa.rec1 = (float4 *)pl.rec1;
    pa.rec2 = (float4 *)pl.rec2;
    pa.rec3 = (float4 *)pl.rec3;
    pa.counters = (unsigned long long *)pl.counters;
    pa.n_hi = pl.n;
    pa.pre = (double *)pl.pre;
    pa.pblk = (double *)pl.pblk;
    pa.n_rec = pl.n_rec;
    if (timing) ENQ(hipEventRecord(ctx->ev[0], st));
    ENQ(launch_prep(pa, (int)pl.max_lines, st));
    if (timing) ENQ(hipEventRecord(ctx->ev[1], st));
    LinesumArgs la{};
    la.items = pa.items;
    la.rec0 = pa.rec0;
    la.rec1 = pa.rec1;
    la.rec2 = pa.rec2;
    la.keys = (const int32_t *)pl.keys;
    la.ktab = (const int32_t *)pl.ktab;
    la.ktab_lo = pl.ktab_lo;
    la.ktab_n = pl.ktab_n;
    la.out = (float *)pl.tau;
    la.n_hi = pl.n;
    la.n_layers = L;
    la.n_species = pl.ns;
    la.nu_lo = pl.nu_lo;
    la.dnu = pl.dnu;
    la.tiles = pl.tiles;
    la.n_prof = pl.gn;
    la.tile_order = (const int32_t *)pl.order;
    la.queue = (unsigned *)pl.queue;
    la.far_d = pl.far_d;
    la.far_hc = pl.far_hc;
    la.work = (unsigned long long *)pl.counters + 3;
    la.pre = pa.pre;
    la.n_rec = pa.n_rec;
    la.rec3 = pa.rec3;
    if (pl.items_lm1) la.items = (const ItemDev *)pl.items_lm1;
    ENQ(launch_linesum(la, pl.P, pl.pblocks, st));
    if (pl.items_lm1) {
        ENQ(launch_reset_counters((unsigned *)pl.queue, nullptr, 0, st));
        la.items = (const ItemDev *)pl.items_lm2;
        la.out = (float *)pl.tau_lm;
        ENQ(launch_linesum(la, pl.P, pl.pblocks, st));
        LmTermArgs lt{};
        lt.terms = (const LmTerm *)pl.lm_terms;
        lt.off = (const int64_t *)pl.lm_off;
        lt.n = pl.n;
        lt.n_units = pl.gn * L;
        lt.an = pl.an;
        lt.sn = pl.sn;
        lt.den = pl.den;
        lt.cutoff = pl.lm_cutoff;
        lt.x_far = pl.x_far;
        lt.tau = (float *)pl.tau;
        lt.tau_lm = (const float *)pl.tau_lm;
        lt.coef = (const double *)pl.lm_coef;
        ENQ(launch_lm_terms(lt, st));
    }
    if (timing) ENQ(hipEventRecord(ctx->ev[2], st));
    if (pl.continua) {
        ContArgs co = cont_args(ctx, pl.continua);
        co.tau = (float *)pl.tau;
        co.n = pl.n;
        co.n_layers = L;
        co.n_prof = pl.gn;
        co.an = pl.an;
        co.sn = pl.sn;
        co.den = pl.den;
        co.lay = (const ContLayer *)((const double *)pl.prof + NL * (2 * L + 4));
        ENQ(launch_continuum(co, st));
    }
    if (timing) ENQ(hipEventRecord(ctx->ev[7], st));
    if (pl.mode >= 1) {
        ChodArgs ca{};
        ca.tau = (const float *)pl.tau;
        ca.n = pl.n;
        ca.n_layers = L;
        ca.taps = (const double *)pl.taps;
        ca.half = pl.with_ils ? pl.half : 0;
        ca.dnu = pl.with_ils ? pl.dnu : 1.0;
        ca.ch_idx = (const int64_t *)pl.chidx;
        ca.n_ch = pl.nch;
        ca.floor = pl.chod_floor;
        ca.od_max = 50.0f;
        ca.scal = (const double *)pl.prof + NL * (2 * L + 1);
        ca.out = (float *)pl.chod;
        if (pl.mode >= 2) {
            ca.lev_out = (float *)pl.chlev;
            ca.lev_in = ca.lev_out;
            ca.g_out = (float *)pl.chg;
            ca.raw_G = pl.mode == 3;
        }
        ca.m_out = (float *)pl.chmom;
        if (pl.chrows)
            ENQ(launch_chod_banded(ca, pl.gn, pl.ch_idx0, (int)pl.ch_stride, (float *)pl.chrows, (double *)pl.chsums,
                                   (double *)pl.chmsums, st));
        else
            ENQ(launch_chod(ca, pl.gn, st));
        if (timing) {
            ENQ(hipEventRecord(ctx->ev[3], st));
            ENQ(hipEventRecord(ctx->ev[4], st));
        }
        if (pl.mode == 3) {
            ENQ(hipMemcpyAsync((void *)pl.h_out, pl.chlev, sizeof(float) * pl.nch * (L + 1) * NL,
                               hipMemcpyDeviceToHost, st));
            ENQ(hipMemcpyAsync((float *)pl.h_out + pl.nch * (L + 1) * NL, pl.chg, sizeof(float) * pl.nch * L * NL,
                               hipMemcpyDeviceToHost, st));
            return hipSuccess;
        }
        ENQ(hipMemcpyAsync((void *)pl.h_out, pl.chod, sizeof(float) * pl.nch * L * NL, hipMemcpyDeviceToHost, st));
        if (pl.mode == 2 || pl.mode == 4)
            ENQ(hipMemcpyAsync((float *)pl.h_out + pl.nch * L * NL, pl.chg, sizeof(float) * pl.nch * L * NL,
                               hipMemcpyDeviceToHost, st));
        if (pl.mode == 4)
            ENQ(hipMemcpyAsync((float *)pl.h_out + 2 * pl.nch * L * NL, pl.chmom, sizeof(float) * pl.nch * L * NL,
                               hipMemcpyDeviceToHost, st));
        return hipSuccess;
    }
    const double *prof = (const double *)pl.prof;
    RteArgs ra{};
    ra.tau32 = (const float *)pl.tau;
    ra.n = pl.n;
    ra.n_layers = L;
    ra.T_lev = prof;
    ra.T_ave = prof + NL * (L + 1);
    ra.scal = prof + NL * (2 * L + 1);
    ra.source_function = pl.source_function;
    ra.an = pl.an;
    ra.sn = pl.sn;
    ra.den = pl.den;
    ra.out = (double *)pl.rhi;
    ra.n_prof = pl.gn;
    ENQ(launch_rte(ra, st));
    if (timing) ENQ(hipEventRecord(ctx->ev[3], st));
    if (pl.ch_stride > 0) {
        IlsBlockArgs ib{};
        ib.spec = (const double *)pl.rhi;
        ib.n = pl.n;
        ib.taps = (const double *)pl.taps;
        ib.half = pl.with_ils ? pl.half : 0;
        ib.dnu = pl.with_ils ? pl.dnu : 1.0;
        ib.idx0 = pl.ch_idx0;
        ib.stride = (int)pl.ch_stride;
        ib.ch_nu = (const double *)pl.chnu;
        ib.n_ch = pl.nch;
        ib.r_out = (double *)pl.rch;
        ib.bt_out = (double *)pl.bt;
        ENQ(launch_ils_block(ib, pl.gn, st));
    } else {
        IlsArgs ia{};
        ia.spec = (const double *)pl.rhi;
        ia.n = pl.n;
        ia.taps = (const double *)pl.taps;
        ia.half = pl.with_ils ? pl.half : 0;
        ia.dnu = pl.with_ils ? pl.dnu : 1.0;
        ia.ch_idx = (const int64_t *)pl.chidx;
        ia.ch_w = (const double *)pl.chw;
        ia.ch_nu = (const double *)pl.chnu;
        ia.n_ch = pl.nch;
        ia.r_out = (double *)pl.rch;
        ia.bt_out = (double *)pl.bt;
        ia.n_prof = pl.gn;
        ENQ(launch_ils_channels(ia, st));
    }
    if (timing) ENQ(hipEventRecord(ctx->ev[4], st));
    double *hout = (double *)pl.h_out;
    ENQ(hipMemcpyAsync(hout, pl.rch, sizeof(double) * pl.nch * NL, hipMemcpyDeviceToHost, st));
    ENQ(hipMemcpyAsync(hout + pl.nch * NL, pl.bt, sizeof(double) * pl.nch * NL, hipMemcpyDeviceToHost, st));
#undef ENQ
    return hipSuccess;
}

// Runs the group from a cached graph, capturing it on the first use of a plan.
static int launch_group_graph(irs_ctx *ctx, const GroupPlan &pl) {
    GraphEntry *hit = nullptr;
    for (auto &g : ctx->graphs)
        if (std::memcmp(&g.plan, &pl, sizeof(GroupPlan)) == 0) hit = &g;
    if (!hit) {
        if (ctx->graphs.size() >= 4) {
            (void)hipGraphExecDestroy(ctx->graphs.front().exec);
            ctx->graphs.erase(ctx->graphs.begin());
        }
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_group(ctx, pl, false);
        hipGraph_t graph = nullptr;
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(ctx, IRS_EHIP, std::string("graph capture: ") +
                                           hipGetErrorString(e != hipSuccess ? e : e2));
        }
        GraphEntry ge;
        ge.plan = pl;
        e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(ctx, IRS_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        ctx->graphs.push_back(ge);
        hit = &ctx->graphs.back();
    }
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    HIPCHK(hipGraphLaunch(hit->exec, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
    ctx->stats.graph_launches += 1;
    return IRS_OK;
}

// Times each (tile_points, persistent) candidate on the first profile of the
// group (records already prepared by a prep launch) and keeps the fastest.
static int autotune_group(irs_ctx *ctx, const GroupPlan &base, const std::vector<const ItemDev *> &per_species,
                   const std::string &key) {
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync((void *)base.items, base.h_items, (size_t)base.items_bytes, hipMemcpyHostToDevice, st));
    HIPCHK(launch_reset_counters((unsigned *)base.queue, (unsigned long long *)base.counters, IRS_NCNT, st));
    PrepArgs pa{};
    pa.items = (const ItemDev *)base.items;
    pa.lists = (const LineListDev *)base.lists;
    pa.tips_q = (const double *)base.tips;
    pa.n_items = (int)(base.L * base.ns);     // first profile only
    pa.nu_lo = base.nu_lo;
    pa.dnu = base.dnu;
    pa.rec0 = (float4 *)base.rec0;
    pa.rec1 = (float4 *)base.rec1;
    pa.rec2 = (float4 *)base.rec2;
    pa.rec3 = (float4 *)base.rec3;
    pa.counters = (unsigned long long *)base.counters;
    pa.n_hi = base.n;
    pa.pre = (double *)base.pre;
    pa.pblk = (double *)base.pblk;
    pa.n_rec = base.n_rec;
    HIPCHK(launch_prep(pa, (int)base.max_lines, st));
    const int cand_P[3] = {4, 8, 16};
    double best = 1e300;
    int bestP = ctx->sched.tile_points, bestPers = ctx->sched.persistent;
    for (int ci = 0; ci < 3; ++ci)
        for (int pers = 0; pers < 2; ++pers) {
            const int P = cand_P[ci];
            const int32_t *ord = nullptr;
            int rc = tile_order(ctx, per_species, base.n, base.nu_lo, base.dnu, P, &ord);
            if (rc) return rc;
            LinesumArgs la{};
            la.items = pa.items;
            la.rec0 = pa.rec0;
            la.rec1 = pa.rec1;
            la.rec2 = pa.rec2;
            la.keys = (const int32_t *)base.keys;
            la.ktab = (const int32_t *)base.ktab;
            la.ktab_lo = base.ktab_lo;
            la.ktab_n = base.ktab_n;
            la.out = (float *)base.tau;
            la.n_hi = base.n;
            la.n_layers = base.L;
            la.n_species = base.ns;
            la.nu_lo = base.nu_lo;
            la.dnu = base.dnu;
            la.tiles = (int)((base.n + (int64_t)IRS_LS_THREADS * P - 1) / ((int64_t)IRS_LS_THREADS * P));
            la.n_prof = 1;
            la.tile_order = ord;
            la.queue = (unsigned *)base.queue;
            far_params(ctx, base.dnu, P, la.far_d, la.far_hc);
            la.pre = pa.pre;
            la.n_rec = pa.n_rec;
            la.rec3 = pa.rec3;
            const int pb = pers ? persistent_blocks(ctx, P) : 0;
            if (pers && pb <= 0) continue;
            double t_min = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
                HIPCHK(launch_reset_counters((unsigned *)base.queue, nullptr, 0, st));
                HIPCHK(hipEventRecord(ctx->ev[5], st));
                HIPCHK(launch_linesum(la, P, pb, st));
                HIPCHK(hipEventRecord(ctx->ev[6], st));
                HIPCHK(hipEventSynchronize(ctx->ev[6]));
                t_min = std::min(t_min, elapsed_ms(ctx->ev[5], ctx->ev[6]));
            }
            if (t_min < best) {
                best = t_min;
                bestP = P;
                bestPers = pers;
            }
        }
    ctx->tuned[key] = {bestP, bestPers};
    return IRS_OK;
}

// One profile through the every-pair FP64 line sum (k_ls64.hip), FP64 optical
// depths, continuum, the FP64 RTE and the channel ILS: the reference kernel's
// arithmetic on the device.  Eager, on the context stream.
static int run_profile_fp64(irs_ctx *ctx, const ItemDev *items_h, int L, int ns, int64_t n, const Grid &g,
                            int max_lines, int64_t n_rec, const double *h_prof, const double *taps, int64_t half,
                            double ils_dnu, int source_function, int continua, int64_t ch_stride, int64_t ch_idx0,
                            int64_t nch, double *R, double *BT, double &ms_prep, double &ms_ls, double &ms_cont,
                            double &ms_rte, double &ms_ils) {
    hipStream_t st = ctx->stream;
    const size_t n_items = (size_t)L * ns;
    const size_t prof_doubles = (size_t)(2 * L + 4) + (size_t)L * 8;
    HIPCHK(ensure(ctx->d_rec64, sizeof(double) * 8 * (size_t)std::max<int64_t>(n_rec, 1)));
    HIPCHK(ensure(ctx->d_tau64, sizeof(double) * (size_t)n * L));
    HIPCHK(hipMemcpyAsync(ctx->d_items.p, items_h, sizeof(ItemDev) * n_items, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->d_prof.p, h_prof, sizeof(double) * prof_doubles, hipMemcpyHostToDevice, st));
    HIPCHK(launch_reset_counters((unsigned *)ctx->d_queue.p, (unsigned long long *)ctx->d_counters.p, IRS_NCNT, st));
    PrepArgs pa{};
    pa.items = (const ItemDev *)ctx->d_items.p;
    pa.lists = (const LineListDev *)ctx->d_lists.p;
    pa.tips_q = (const double *)ctx->d_tips.p;
    pa.n_items = (int)n_items;
    pa.nu_lo = g.at(0);
    pa.dnu = g.step;
    pa.rec0 = (float4 *)ctx->d_rec0.p;
    pa.rec1 = (float4 *)ctx->d_rec1.p;
    pa.rec2 = (float4 *)ctx->d_rec2.p;
    pa.counters = (unsigned long long *)ctx->d_counters.p;
    pa.n_hi = n;
    pa.n_rec = std::max<int64_t>(n_rec, 1);
    pa.rec64 = (double *)ctx->d_rec64.p;
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(launch_prep(pa, max_lines, st));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    Linesum64Args la{};
    la.items = pa.items;
    la.rec64 = pa.rec64;
    la.keys = (const int32_t *)ctx->d_keys.p;
    la.tau = (double *)ctx->d_tau64.p;
    la.n_hi = n;
    la.n_layers = L;
    la.n_species = ns;
    la.dnu = g.step;
    la.an = g.an;
    la.sn = g.sn;
    la.den = g.den;
    HIPCHK(launch_linesum64(la, st));
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    const double *prof = (const double *)ctx->d_prof.p;
    if (continua) {
        ContArgs co = cont_args(ctx, continua);
        co.tau64 = (double *)ctx->d_tau64.p;
        co.add64 = 1;
        co.n = n;
        co.n_layers = L;
        co.n_prof = 1;
        co.an = g.an;
        co.sn = g.sn;
        co.den = g.den;
        co.lay = (const ContLayer *)(prof + (2 * L + 4));
        HIPCHK(launch_continuum(co, st));
    }
    HIPCHK(hipEventRecord(ctx->ev[7], st));
    RteArgs ra{};
    ra.tau64 = (const double *)ctx->d_tau64.p;
    ra.n = n;
    ra.n_layers = L;
    ra.T_lev = prof;
    ra.T_ave = prof + (L + 1);
    ra.scal = prof + (2 * L + 1);
    ra.source_function = source_function;
    ra.an = g.an;
    ra.sn = g.sn;
    ra.den = g.den;
    ra.out = (double *)ctx->d_rhi.p;
    ra.n_prof = 1;
    HIPCHK(launch_rte(ra, st));
    HIPCHK(hipEventRecord(ctx->ev[3], st));
    if (ch_stride > 0) {
        IlsBlockArgs ib{};
        ib.spec = (const double *)ctx->d_rhi.p;
        ib.n = n;
        ib.taps = taps;
        ib.half = half;
        ib.dnu = ils_dnu;
        ib.idx0 = ch_idx0;
        ib.stride = (int)ch_stride;
        ib.ch_nu = (const double *)ctx->d_chnu.p;
        ib.n_ch = nch;
        ib.r_out = (double *)ctx->d_rch.p;
        ib.bt_out = (double *)ctx->d_bt.p;
        HIPCHK(launch_ils_block(ib, 1, st));
    } else {
        IlsArgs ia{};
        ia.spec = (const double *)ctx->d_rhi.p;
        ia.n = n;
        ia.taps = taps;
        ia.half = half;
        ia.dnu = ils_dnu;
        ia.ch_idx = (const int64_t *)ctx->d_chidx.p;
        ia.ch_w = (const double *)ctx->d_chw.p;
        ia.ch_nu = (const double *)ctx->d_chnu.p;
        ia.n_ch = nch;
        ia.r_out = (double *)ctx->d_rch.p;
        ia.bt_out = (double *)ctx->d_bt.p;
        ia.n_prof = 1;
        HIPCHK(launch_ils_channels(ia, st));
    }
    HIPCHK(hipEventRecord(ctx->ev[4], st));
    HIPCHK(hipMemcpyAsync(R, ctx->d_rch.p, sizeof(double) * nch, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(BT, ctx->d_bt.p, sizeof(double) * nch, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ms_prep += elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ms_ls += elapsed_ms(ctx->ev[1], ctx->ev[2]);
    ms_cont += elapsed_ms(ctx->ev[2], ctx->ev[7]);
    ms_rte += elapsed_ms(ctx->ev[7], ctx->ev[3]);
    ms_ils += elapsed_ms(ctx->ev[3], ctx->ev[4]);
    return IRS_OK;
}

// Line-mixing terms of every (profile, layer) of a group at the line-mixing
// species' T_cg, p_cg (built on host threads), their upload, the per-unit coef and
// the two item arrays of the split line sum (items_lm1: every species but s_lm,
// items_lm2: s_lm alone).
static int prepare_relmat(irs_ctx *ctx, const ItemDev *items, int gn, int L, int ns, int s_lm,
                          const irs_fm_options &opt) {
    const RelmatHost &rm = *ctx->relmats[opt.relmat];
    const int nu_units = gn * L;
    std::vector<std::vector<LmTerm>> units((size_t)nu_units);
    std::vector<double> coef((size_t)nu_units, 0.0);
    std::vector<std::string> errs((size_t)nu_units);
    const int nth = std::max(1, std::min<int>(nu_units, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::vector<std::thread> pool;
    for (int w = 0; w < nth; ++w)
        pool.emplace_back([&, w]() {
            for (int u = w; u < nu_units; u += nth) {
                const ItemDev &it = items[(size_t)u * ns + s_lm];
                if (!it.active) continue;
                coef[u] = it.coef;
                errs[u] = relmat_terms(rm, opt.relmat_band_mode, it.T, it.p_atm, opt.keep_threshold, units[u]);
            }
        });
    for (auto &t : pool) t.join();
    for (const auto &e : errs)
        if (!e.empty()) return fail(ctx, IRS_EINVAL, e);
    int rc = upload_lm_terms(ctx, units);
    if (rc) return rc;
    const size_t ni = (size_t)gn * L * ns;
    std::vector<ItemDev> two(2 * ni);
    for (size_t i = 0; i < ni; ++i) {
        two[i] = items[i];
        two[ni + i] = items[i];
        if ((int)(i % ns) == s_lm) two[i].active = 0;
        else two[ni + i].active = 0;
    }
    HIPCHK(hipMemcpyAsync(ctx->d_items_lm.p, two.data(), sizeof(ItemDev) * two.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_lm_coef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

static int run_forward(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                       const irs_fm_options *opt_in, int mode, double chod_floor, double *R, double *BT,
                       float *tau_ch, float *cim_g, float *ils_mom = nullptr) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (!profiles || n_prof <= 0 || !snd || (mode == 0 && (!R || !BT)) || (mode >= 1 && !tau_ch) ||
        (mode >= 2 && !cim_g) || (mode == 4 && !ils_mom))
        return fail(ctx, IRS_EINVAL, "null argument");
    irs_fm_options opt;
    if (opt_in) opt = *opt_in; else irs_default_options(&opt);
    if (!(snd->dnu > 0.0)) return fail(ctx, IRS_EINVAL, "sounder dnu must be > 0");
    if (!(opt.cutoff > 0.0) || !(opt.lm_cutoff > 0.0)) return fail(ctx, IRS_EINVAL, "cutoff must be > 0");
    if (opt.source_function != IRS_SRC_TOON && opt.source_function != IRS_SRC_CIM)
        return fail(ctx, IRS_EINVAL, "source_function must be :toon or :cim");
    if (opt.apodization < 0 || opt.apodization > 3) return fail(ctx, IRS_EINVAL, "Unknown apodization");
    if (opt.line_mixing != IRS_LM_NONE && opt.line_mixing != IRS_LM_VPY && opt.line_mixing != IRS_LM_RELMAT)
        return fail(ctx, IRS_EINVAL, "bad line_mixing");
    if (opt.continua < 0 || opt.continua > IRS_CONT_ALL) return fail(ctx, IRS_EINVAL, "bad continua mask");
    const bool relmat = opt.line_mixing == IRS_LM_RELMAT;
    if (relmat) {
        if (opt.relmat < 0 || opt.relmat >= (int32_t)ctx->relmats.size() || !ctx->relmats[opt.relmat])
            return fail(ctx, IRS_EINVAL, "bad relaxation-matrix handle");
        if (ctx->sched.fp64_every_pair && mode == 0)
            return fail(ctx, IRS_EINVAL, "the every-pair FP64 path has no relaxation-matrix line mixing");
    }
    auto t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));

    const Grid g = internal_grid(snd, &opt, nullptr);
    const Grid gc = make_grid(snd->nu_min, snd->dnu, snd->nu_max);
    const int64_t n = g.n;
    const double nu_lo = g.at(0), dnu = g.step;
    const int L = profiles[0].n_levels - 1;
    const int ns = profiles[0].n_species;
    for (int q = 0; q < n_prof; ++q) {
        const irs_profile &pr = profiles[q];
        if (pr.n_levels - 1 != L || pr.n_species != ns)
            return fail(ctx, IRS_EINVAL, "all profiles of a batch need the same levels and species");
        if (!(pr.mu > 0.0 && pr.mu <= 1.0)) return fail(ctx, IRS_EINVAL, "μ must be in (0, 1]");
        for (int s = 0; s < ns; ++s) {
            if (pr.species[s] != profiles[0].species[s] || pr.linelist[s] != profiles[0].linelist[s])
                return fail(ctx, IRS_EINVAL, "all profiles of a batch need the same species and line lists");
            if (pr.linelist[s] >= 0) {
                int rc = check_list(ctx, pr.linelist[s]);
                if (rc) return rc;
            }
        }
    }
    if (L < 1) return fail(ctx, IRS_EINVAL, "need at least two levels");
    int rc = ensure_common(ctx);
    if (rc) return rc;
    std::vector<int32_t> used;
    for (int s = 0; s < ns; ++s)
        if (profiles[0].linelist[s] >= 0) used.push_back(profiles[0].linelist[s]);
    std::map<int32_t, int32_t> koff;
    std::map<int32_t, int64_t> ktoff;
    rc = build_keys(ctx, used, nu_lo, dnu, koff, n, std::max(opt.cutoff, opt.lm_cutoff), &ktoff);
    if (rc) return rc;

    double *taps = nullptr;
    int64_t half = 0;
    if (opt.with_ils) {
        rc = get_taps(ctx, dnu, snd->opd_max, snd->fwhm_gauss, opt.apodization, 2048, &taps, &half);
        if (rc) return rc;
    } else {
        // identity "ILS": a single unit tap, kept in the context
        TapsKey key{0.0, 0.0, 0.0, -1, 1};
        DevBuf &b = ctx->taps[key];
        if (!b.p) {
            const double one = 1.0;
            HIPCHK(ensure(b, sizeof(double)));
            HIPCHK(hipMemcpy(b.p, &one, sizeof(double), hipMemcpyHostToDevice));
        }
        taps = (double *)b.p;
    }
    std::vector<int64_t> cidx;
    std::vector<double> cw, cnu;
    channel_positions(g, gc, cidx, cw, cnu);
    const int64_t nch = gc.n;
    if (mode >= 1)
        for (int64_t c = 0; c < nch; ++c)
            if (cw[c] != 0.0)
                return fail(ctx, IRS_EINVAL, "channel optical depths need channels on internal grid nodes");
    HIPCHK(ensure(ctx->d_chidx, sizeof(int64_t) * nch));
    HIPCHK(ensure(ctx->d_chw, sizeof(double) * nch));
    HIPCHK(ensure(ctx->d_chnu, sizeof(double) * nch));
    HIPCHK(hipMemcpyAsync(ctx->d_chidx.p, cidx.data(), sizeof(int64_t) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chw.p, cw.data(), sizeof(double) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chnu.p, cnu.data(), sizeof(double) * nch, hipMemcpyHostToDevice, ctx->stream));

    // ---- profile groups: records + tau + R_hi per profile within a memory budget
    int64_t lines_per_prof = 0;
    int max_lines = 0;
    for (int s = 0; s < ns; ++s)
        if (profiles[0].linelist[s] >= 0) {
            lines_per_prof += ctx->lists[profiles[0].linelist[s]].n;
            max_lines = std::max<int>(max_lines, (int)ctx->lists[profiles[0].linelist[s]].n);
        }
    const double bytes_per_prof = (double)L * (lines_per_prof * 168.0 + n * (relmat ? 8.0 : 4.0)) + n * 8.0;
    double budget = 32.0e9;
    if (const char *e = std::getenv("IRS_GROUP_BYTES")) budget = std::max(1.0, std::atof(e));
    int gmax = (int)std::max(1.0, std::min((double)n_prof, budget / std::max(bytes_per_prof, 1.0)));
    if (const char *e = std::getenv("IRS_PROFILE_GROUP")) gmax = std::max(1, std::min(n_prof, std::atoi(e)));
    const bool fp64 = ctx->sched.fp64_every_pair && mode == 0;
    if (fp64) gmax = 1;                    // the every-pair FP64 path runs one profile at a time
    const int n_groups = (n_prof + gmax - 1) / gmax;
    const int group = (n_prof + n_groups - 1) / n_groups;

    // ---- schedule (tile points, persistent kernel), autotuned per shape
    std::vector<ItemDev> proto(ns);
    std::vector<const ItemDev *> per_species;
    int P = ctx->sched.tile_points, persistent = ctx->sched.persistent;
    double ms_prep = 0, ms_ls = 0, ms_rte = 0, ms_ils = 0, ms_graph = 0, ms_cont = 0;
    bool tuned_now = false;

    // per profile: T_lev [L+1], T_ave [L], {T_sfc, mu, eps}, then the continuum layer
    // states [L] (8 doubles each) at offset group * (2L + 4)
    const size_t prof_doubles = (size_t)group * (2 * L + 1 + 3) + (size_t)group * L * 8;
    HIPCHK(ensure(ctx->d_items, sizeof(ItemDev) * (size_t)group * L * ns));
    HIPCHK(ensure(ctx->d_rec0, sizeof(float4) * std::max<int64_t>((int64_t)group * L * lines_per_prof, 1)));
    HIPCHK(ensure(ctx->d_rec1, sizeof(float4) * std::max<int64_t>((int64_t)group * L * lines_per_prof, 1)));
    HIPCHK(ensure(ctx->d_rec2, sizeof(float4) * std::max<int64_t>((int64_t)group * L * lines_per_prof, 1)));
    HIPCHK(ensure(ctx->d_rec3, sizeof(float4) * 6 * std::max<int64_t>((int64_t)group * L * lines_per_prof, 1)));
    const int64_t n_rec_cap = std::max<int64_t>((int64_t)group * L * lines_per_prof, 1);
    HIPCHK(ensure(ctx->d_pre, sizeof(double) * 3 * n_rec_cap));
    HIPCHK(ensure(ctx->d_pblk, sizeof(double) * 3 * (n_rec_cap / 256 + (int64_t)group * L * ns + 1)));
    HIPCHK(ensure(ctx->d_tau, sizeof(float) * (size_t)n * L * group));
    HIPCHK(ensure(ctx->d_rhi, sizeof(double) * (size_t)n * group));
    HIPCHK(ensure(ctx->d_rch, sizeof(double) * nch * group));
    HIPCHK(ensure(ctx->d_bt, sizeof(double) * nch * group));
    HIPCHK(ensure(ctx->d_counters, sizeof(unsigned long long) * IRS_NCNT));
    HIPCHK(ensure(ctx->d_queue, 64));
    HIPCHK(ensure(ctx->d_prof, sizeof(double) * prof_doubles));
    HIPCHK(ensure_pinned(ctx->h_items, sizeof(ItemDev) * (size_t)group * L * ns));
    HIPCHK(ensure_pinned(ctx->h_prof, sizeof(double) * prof_doubles));
    HIPCHK(ensure_pinned(ctx->h_out, std::max(sizeof(double) * 2 * nch * group,
                                              sizeof(float) * (3 * L + 1) * nch * group)));
    if (mode >= 1) HIPCHK(ensure(ctx->d_chod, sizeof(float) * nch * L * group));
    // relaxation-matrix line mixing: the species it applies to, its second line-sum output
    int s_lm = -1;
    if (relmat) {
        for (int s = 0; s < ns; ++s)
            if (profiles[0].species[s] == opt.lm_species) { s_lm = s; break; }
        if (s_lm >= 0) {
            HIPCHK(ensure(ctx->d_tau_lm, sizeof(float) * (size_t)n * L * group));
            HIPCHK(ensure(ctx->d_items_lm, 2 * sizeof(ItemDev) * (size_t)group * L * ns));
            HIPCHK(ensure(ctx->d_lm_coef, sizeof(double) * (size_t)group * L));
        }
    }
    if (mode >= 2) {
        HIPCHK(ensure(ctx->d_chlev, sizeof(float) * nch * (L + 1) * group));
        HIPCHK(ensure(ctx->d_chg, sizeof(float) * nch * L * group));
    }

    for (int g0 = 0; g0 < n_prof; g0 += group) {
        const int gn = std::min(group, n_prof - g0);
        ItemDev *items = (ItemDev *)ctx->h_items.p;
        double *hp = (double *)ctx->h_prof.p;
        double *h_tlev = hp, *h_tave = hp + (size_t)gn * (L + 1), *h_scal = hp + (size_t)gn * (2 * L + 1);
        ContLayer *h_cont = (ContLayer *)(hp + (size_t)gn * (2 * L + 4));
        int64_t rec = 0, pblk = 0;
        for (int q = 0; q < gn; ++q) {
            const irs_profile &pr = profiles[g0 + q];
            LayerProps lp;
            rc = layer_properties(ctx, &pr, opt.T_method, lp);
            if (rc) return rc;
            for (int k = 0; k <= L; ++k) h_tlev[(size_t)q * (L + 1) + k] = pr.T[k];
            for (int k = 0; k < L; ++k) h_tave[(size_t)q * L + k] = lp.T_ave[k];
            h_scal[3 * q] = pr.T_sfc > 0.0 ? pr.T_sfc : pr.T[0];   // sounder.jl:337
            h_scal[3 * q + 1] = pr.mu;
            h_scal[3 * q + 2] = pr.eps_sfc;
            if (opt.continua) cont_layers(&pr, lp, h_cont + (size_t)q * L);
            for (int k = 0; k < L; ++k)
                for (int s = 0; s < ns; ++s) {
                    ItemDev &it = items[((size_t)q * L + k) * ns + s];
                    it = ItemDev{};
                    const int32_t h = pr.linelist[s];
                    const double vmr = lp.vmr_cg[s][k];
                    if (h >= 0 && vmr != 0.0 && ctx->lists[h].n > 0) {   // sounder.jl:285-286
                        const int sp = pr.species[s];
                        const double p_atm = lp.p_cg[s][k] / 1013.25;
                        const double T = lp.T_cg[s][k];
                        const bool lm = opt.line_mixing == IRS_LM_VPY && sp == 2 && ctx->lists[h].has_lm;
                        const double vs = (sp == 1 && !lm) ? vmr : 0.0;
                        const double coef = vmr * lp.dp[k] * 2.1209e22;        // sounder.jl:280,290
                        it = make_item(ctx->lists[h], h, T, p_atm, vs, coef, opt.dptmn, opt.cutoff,
                                       opt.lm_cutoff, opt.x_far, lm, dnu);
                        // DPTMIN is skipped on the line-mixing CO2 path (sounder.jl:295-299)
                        if (sp == 2 && opt.line_mixing != IRS_LM_NONE) it.dptmn = 0.0;
                        it.key_off = koff[h];
                        it.ktab_off = ktoff.count(h) ? ktoff[h] : 0;
                    }
                    it.rec_off = rec;
                    rec += it.n;
                    it.pb_off = pblk;
                    pblk += (it.n + 255) / 256;
                    if (q == 0 && k == L / 2 && it.active) proto[s] = it;
                }
        }
        if (g0 == 0) {
            for (int s = 0; s < ns; ++s)
                if (proto[s].active) per_species.push_back(&proto[s]);
            if (ctx->sched.autotune) {
                const std::string key = shape_key(per_species, n, nu_lo, dnu, 0);
                std::string pkey;
                if (!ctx->tuned.count(key) && std::getenv("IRS_TUNE_FILE")) {
                    pkey = persistent_tune_key(ctx, per_species, n, nu_lo, dnu);
                    int fP = 0, fPers = 0;
                    if (tune_file_lookup(pkey, fP, fPers)) ctx->tuned[key] = {fP, fPers};
                }
                if (!ctx->tuned.count(key)) {
                    GroupPlan tp;
                    std::memset(&tp, 0, sizeof tp);
                    tp.items = ctx->d_items.p; tp.rec0 = ctx->d_rec0.p; tp.rec1 = ctx->d_rec1.p;
                    tp.rec2 = ctx->d_rec2.p; tp.rec3 = ctx->d_rec3.p; tp.keys = ctx->d_keys.p;
                    tp.ktab = ctx->ktab_n > 0 ? ctx->d_ktab.p : nullptr; tp.ktab_lo = ctx->ktab_lo;
                    tp.ktab_n = ctx->ktab_n; tp.tau = ctx->d_tau.p;
                    tp.counters = ctx->d_counters.p; tp.queue = ctx->d_queue.p; tp.lists = ctx->d_lists.p;
                    tp.tips = ctx->d_tips.p; tp.h_items = ctx->h_items.p; tp.pre = ctx->d_pre.p;
                    tp.pblk = ctx->d_pblk.p;
                    tp.n_rec = n_rec_cap;
                    tp.items_bytes = (int64_t)(sizeof(ItemDev) * (size_t)L * ns);
                    tp.n = n; tp.L = L; tp.ns = ns; tp.max_lines = max_lines; tp.nu_lo = nu_lo; tp.dnu = dnu;
                    rc = autotune_group(ctx, tp, per_species, key);
                    if (rc) return rc;
                    tuned_now = true;
                    if (!pkey.empty()) tune_file_store(pkey, ctx->tuned[key].first, ctx->tuned[key].second);
                }
                P = ctx->tuned[key].first;
                persistent = ctx->tuned[key].second;
            }
        }
        if (s_lm >= 0) {
            rc = prepare_relmat(ctx, items, gn, L, ns, s_lm, opt);
            if (rc) return rc;
        }
        if (fp64) {
            const int64_t ch_stride = channel_stride(cidx, cw);
            rc = run_profile_fp64(ctx, items, L, ns, n, g, max_lines, rec, (const double *)ctx->h_prof.p, taps,
                                  opt.with_ils ? half : 0, opt.with_ils ? dnu : 1.0, opt.source_function,
                                  opt.continua, ch_stride, ch_stride > 0 ? cidx[0] : 0, nch, R + (size_t)g0 * nch,
                                  BT + (size_t)g0 * nch, ms_prep, ms_ls, ms_cont, ms_rte, ms_ils);
            if (rc) return rc;
            read_counters(ctx);
            ctx->stats.linesum_launches += 1;
            continue;
        }
        const int32_t *order = nullptr;
        rc = tile_order(ctx, per_species, n, nu_lo, dnu, P, &order);
        if (rc) return rc;
        GroupPlan pl;
        std::memset(&pl, 0, sizeof pl);
        pl.items = ctx->d_items.p; pl.rec0 = ctx->d_rec0.p; pl.rec1 = ctx->d_rec1.p; pl.rec2 = ctx->d_rec2.p; pl.rec3 = ctx->d_rec3.p;
        pl.keys = ctx->d_keys.p; pl.tau = ctx->d_tau.p;
        pl.ktab = ctx->ktab_n > 0 ? ctx->d_ktab.p : nullptr; pl.ktab_lo = ctx->ktab_lo; pl.ktab_n = ctx->ktab_n; pl.rhi = ctx->d_rhi.p; pl.counters = ctx->d_counters.p;
        pl.queue = ctx->d_queue.p; pl.lists = ctx->d_lists.p; pl.tips = ctx->d_tips.p; pl.order = order;
        pl.prof = ctx->d_prof.p; pl.rch = ctx->d_rch.p; pl.bt = ctx->d_bt.p; pl.chidx = ctx->d_chidx.p;
        pl.chw = ctx->d_chw.p; pl.chnu = ctx->d_chnu.p; pl.taps = taps; pl.h_items = ctx->h_items.p;
        pl.ch_stride = channel_stride(cidx, cw);
        pl.ch_idx0 = pl.ch_stride > 0 ? cidx[0] : 0;
        pl.pre = ctx->d_pre.p; pl.pblk = ctx->d_pblk.p; pl.n_rec = n_rec_cap;
        pl.h_prof = ctx->h_prof.p; pl.h_out = ctx->h_out.p;
        pl.n = n; pl.nch = nch; pl.half = half; pl.n_items = (int64_t)gn * L * ns; pl.max_lines = max_lines;
        pl.items_bytes = (int64_t)(sizeof(ItemDev) * (size_t)gn * L * ns);
        pl.prof_bytes = (int64_t)(sizeof(double) * ((size_t)gn * (2 * L + 1 + 3) + (opt.continua ? (size_t)gn * L * 8 : 0)));
        pl.gn = gn; pl.L = L; pl.ns = ns; pl.P = P;
        pl.pblocks = persistent ? persistent_blocks(ctx, P) : 0;
        pl.tiles = (int)((n + (int64_t)IRS_LS_THREADS * P - 1) / ((int64_t)IRS_LS_THREADS * P));
        pl.source_function = opt.source_function; pl.with_ils = opt.with_ils;
        far_params(ctx, dnu, P, pl.far_d, pl.far_hc);
        pl.nu_lo = nu_lo; pl.dnu = dnu; pl.an = g.an; pl.sn = g.sn; pl.den = g.den;
        pl.mode = mode;
        pl.chod = mode >= 1 ? ctx->d_chod.p : nullptr;
        pl.chlev = mode >= 2 ? ctx->d_chlev.p : nullptr;
        pl.chg = mode >= 2 ? ctx->d_chg.p : nullptr;
        pl.chod_floor = chod_floor;
        if (mode >= 1 && pl.ch_stride > 0) {
            const int64_t nr = chod_banded_rows(L, mode >= 2);
            HIPCHK(ensure(ctx->d_chrows, sizeof(float) * (size_t)nr * n));
            HIPCHK(ensure(ctx->d_chsums, sizeof(double) * (size_t)nr * nch));
            pl.chrows = ctx->d_chrows.p;
            pl.chsums = ctx->d_chsums.p;
            if (mode == 4) {
                HIPCHK(ensure(ctx->d_chmom, sizeof(float) * (size_t)nch * L * group));
                HIPCHK(ensure(ctx->d_chmsums, sizeof(double) * ((size_t)(L + 1) * nch + 2 * half + 1)));
                pl.chmom = ctx->d_chmom.p;
                pl.chmsums = ctx->d_chmsums.p;
            }
        } else if (mode == 4) {
            return fail(ctx, IRS_EINVAL, "ILS moments need equally spaced channels on internal grid nodes");
        }
        pl.continua = opt.continua;
        pl.cia_gen = opt.continua ? ctx->cia_gen : 0;
        if (s_lm >= 0) {
            pl.items_lm1 = ctx->d_items_lm.p;
            pl.items_lm2 = (const ItemDev *)ctx->d_items_lm.p + (size_t)gn * L * ns;
            pl.tau_lm = ctx->d_tau_lm.p;
            pl.lm_terms = ctx->d_lm_terms.p;
            pl.lm_off = ctx->d_lm_off.p;
            pl.lm_coef = ctx->d_lm_coef.p;
            pl.lm_cutoff = opt.lm_cutoff;
            pl.x_far = opt.x_far;
        }
        const bool graph = ctx->sched.use_graph && s_lm < 0;
        if (graph) {
            rc = launch_group_graph(ctx, pl);
            if (rc) return rc;
        } else {
            hipError_t e = enqueue_group(ctx, pl, true);
            if (e != hipSuccess) return fail(ctx, IRS_EHIP, std::string("forward: ") + hipGetErrorString(e));
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (graph) {
            ms_graph += elapsed_ms(ctx->ev[5], ctx->ev[6]);
        } else {
            ms_prep += elapsed_ms(ctx->ev[0], ctx->ev[1]);
            ms_ls += elapsed_ms(ctx->ev[1], ctx->ev[2]);
            ms_cont += elapsed_ms(ctx->ev[2], ctx->ev[7]);
            ms_rte += elapsed_ms(ctx->ev[7], ctx->ev[3]);
            ms_ils += elapsed_ms(ctx->ev[3], ctx->ev[4]);
        }
        if (mode == 3) {
            std::memcpy(tau_ch + (size_t)g0 * (L + 1) * nch, ctx->h_out.p, sizeof(float) * nch * (L + 1) * gn);
            std::memcpy(cim_g + (size_t)g0 * L * nch, (const float *)ctx->h_out.p + (size_t)nch * (L + 1) * gn,
                        sizeof(float) * nch * L * gn);
        } else if (mode >= 1) {
            std::memcpy(tau_ch + (size_t)g0 * L * nch, ctx->h_out.p, sizeof(float) * nch * L * gn);
            if (mode == 2 || mode == 4)
                std::memcpy(cim_g + (size_t)g0 * L * nch, (const float *)ctx->h_out.p + (size_t)nch * L * gn,
                            sizeof(float) * nch * L * gn);
            if (mode == 4)
                std::memcpy(ils_mom + (size_t)g0 * L * nch, (const float *)ctx->h_out.p + (size_t)2 * nch * L * gn,
                            sizeof(float) * nch * L * gn);
        } else {
            const double *hout = (const double *)ctx->h_out.p;
            std::memcpy(R + (size_t)g0 * nch, hout, sizeof(double) * nch * gn);
            std::memcpy(BT + (size_t)g0 * nch, hout + (size_t)nch * gn, sizeof(double) * nch * gn);
        }
        read_counters(ctx);
        ctx->stats.linesum_launches += 1;
    }
    ctx->stats.ms_prep = ms_prep;
    ctx->stats.ms_linesum = ms_ls;
    ctx->stats.ms_rte = ms_rte;
    ctx->stats.ms_ils = ms_ils;
    ctx->stats.ms_continuum = ms_cont;
    ctx->stats.ms_device = (ctx->sched.use_graph && !fp64 && !relmat) ? ms_graph
                                                                         : ms_prep + ms_ls + ms_cont + ms_rte + ms_ils;
    ctx->stats.n_hi = n;
    ctx->stats.n_layers = L;
    ctx->stats.tile_points = P;
    ctx->stats.persistent = persistent;
    ctx->stats.autotuned = tuned_now ? 1 : 0;
    ctx->stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return IRS_OK;
}

int irs_forward(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                const irs_fm_options *opt, double *R, double *BT) {
    return run_forward(ctx, profiles, n_prof, snd, opt, 0, 0.0, R, BT, nullptr, nullptr);
}

int irs_channel_od(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                   const irs_fm_options *opt, double floor, float *tau_ch, float *cim_g) {
    if (ctx && !(floor > 0.0 && floor < 1.0)) return fail(ctx, IRS_EINVAL, "floor must be in (0, 1)");
    return run_forward(ctx, profiles, n_prof, snd, opt, cim_g ? 2 : 1, floor, nullptr, nullptr, tau_ch, cim_g);
}

int irs_channel_od_moments(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                           const irs_fm_options *opt, double floor, float *tau_ch, float *cim_g, float *ils_mom) {
    if (ctx && !(floor > 0.0 && floor < 1.0)) return fail(ctx, IRS_EINVAL, "floor must be in (0, 1)");
    return run_forward(ctx, profiles, n_prof, snd, opt, 4, floor, nullptr, nullptr, tau_ch, cim_g, ils_mom);
}

int irs_layer_properties(const irs_profile *pr, int T_method, double *p_mid, double *dp, double *T_ave,
                         double *vmr_cg, double *p_cg, double *T_cg) {
    if (!pr || !p_mid || !dp || !T_ave || !vmr_cg || !p_cg || !T_cg || !pr->p_hpa || !pr->T ||
        (pr->n_species > 0 && !pr->vmr))
        return IRS_EINVAL;
    LayerProps lp;
    int rc = layer_properties(nullptr, pr, T_method, lp);
    if (rc) return rc;
    const int L = pr->n_levels - 1;
    for (int k = 0; k < L; ++k) {
        p_mid[k] = lp.p_mid[k];
        dp[k] = lp.dp[k];
        T_ave[k] = lp.T_ave[k];
    }
    for (int s = 0; s < pr->n_species; ++s)
        for (int k = 0; k < L; ++k) {
            vmr_cg[(size_t)s * L + k] = lp.vmr_cg[s][k];
            p_cg[(size_t)s * L + k] = lp.p_cg[s][k];
            T_cg[(size_t)s * L + k] = lp.T_cg[s][k];
        }
    return IRS_OK;
}

int irs_channel_transmittance(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                              const irs_fm_options *opt, float *T_lev, float *cim_G) {
    return run_forward(ctx, profiles, n_prof, snd, opt, 3, 1e-4, nullptr, nullptr, T_lev, cim_G);
}

int irs_set_schedule(irs_ctx *ctx, const irs_schedule *sc) {
    if (!ctx || !sc) return IRS_EINVAL;
    if (sc->tile_points != 4 && sc->tile_points != 8 && sc->tile_points != 16)
        return fail(ctx, IRS_EINVAL, "tile_points must be 4, 8 or 16");
    ctx->sched = *sc;
    ctx->sched.persistent = sc->persistent ? 1 : 0;
    ctx->sched.use_graph = sc->use_graph ? 1 : 0;
    ctx->sched.autotune = sc->autotune ? 1 : 0;
    ctx->sched.far_wing_cm = sc->far_wing_cm > 0.0 ? sc->far_wing_cm : 0.0;
    ctx->sched.fp64_every_pair = sc->fp64_every_pair ? 1 : 0;
    ctx->sched.reserved = 0;
    ctx->tuned.clear();
    return IRS_OK;
}

int irs_get_schedule(const irs_ctx *ctx, irs_schedule *sc) {
    if (!ctx || !sc) return IRS_EINVAL;
    *sc = ctx->sched;
    return IRS_OK;
}

int irs_cia_set(irs_ctx *ctx, int pair, int n_blocks, const double *T, const int64_t *off, const double *nu,
                const double *sigma) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (pair < 0 || pair > 2) return fail(ctx, IRS_EINVAL, "CIA pair must be 0 (CO2-CO2), 1 (N2-N2) or 2 (O2-O2)");
    if (n_blocks < 0 || n_blocks > IRS_CIA_MAXB)
        return fail(ctx, IRS_EINVAL, "CIA table needs 0.." + std::to_string(IRS_CIA_MAXB) + " blocks");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto &ge : ctx->graphs)
        if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
    ctx->graphs.clear();
    ctx->cia_gen += 1;
    ctx->cia_nb[pair] = 0;
    ctx->cia_nk[pair] = 0;
    if (n_blocks == 0) return IRS_OK;
    if (!T || !off || !nu || !sigma) return fail(ctx, IRS_EINVAL, "null argument");
    if (off[0] != 0) return fail(ctx, IRS_EINVAL, "CIA block offsets must start at 0");
    for (int b = 0; b < n_blocks; ++b) {
        if (off[b + 1] <= off[b]) return fail(ctx, IRS_EINVAL, "every CIA block needs at least one knot");
        for (int64_t i = off[b] + 1; i < off[b + 1]; ++i)
            if (!(nu[i] > nu[i - 1])) return fail(ctx, IRS_EINVAL, "CIA block wavenumbers must increase");
        if (!std::isfinite(T[b])) return fail(ctx, IRS_EINVAL, "bad CIA block temperature");
    }
    // blocks in ascending temperature, stable so that equal temperatures keep
    // their file order (the first one wins the bracketing, continuum.jl:346-350)
    std::vector<int> ord(n_blocks);
    for (int b = 0; b < n_blocks; ++b) ord[b] = b;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return T[x] < T[y]; });
    const int64_t nk = off[n_blocks];
    std::vector<double> buf((size_t)4 * n_blocks + 1 + 2 * (size_t)nk);
    int64_t *boff = (int64_t *)(buf.data() + 3 * n_blocks);
    double *bnu = buf.data() + 4 * n_blocks + 1, *bsig = bnu + nk;
    int64_t o = 0;
    for (int b = 0; b < n_blocks; ++b) {
        const int src = ord[b];
        buf[b] = T[src];
        buf[n_blocks + b] = nu[off[src]];
        buf[2 * n_blocks + b] = nu[off[src + 1] - 1];
        boff[b] = o;
        for (int64_t i = off[src]; i < off[src + 1]; ++i, ++o) {
            bnu[o] = nu[i];
            bsig[o] = sigma[i] > 0.0 ? sigma[i] : 0.0;
        }
    }
    boff[n_blocks] = o;
    HIPCHK(ensure(ctx->d_cia[pair], sizeof(double) * buf.size()));
    HIPCHK(hipMemcpy(ctx->d_cia[pair].p, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
    ctx->cia_nb[pair] = n_blocks;
    ctx->cia_nk[pair] = nk;
    return IRS_OK;
}

int irs_cia_load_file(irs_ctx *ctx, int pair, const char *path) {
    if (!ctx) return IRS_EINVAL;
    if (!path) return fail(ctx, IRS_EINVAL, "null path");
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(ctx, IRS_ENOTFOUND, std::string("cannot open CIA file ") + path);
    std::vector<double> T, nu, sig;
    std::vector<int64_t> off;
    std::vector<std::pair<double, double>> cur;
    auto flush = [&]() {
        if (cur.empty()) return;
        // HITRAN sometimes stores a block in descending wavenumber (continuum.jl:116-121)
        std::stable_sort(cur.begin(), cur.end(),
                         [](const std::pair<double, double> &a, const std::pair<double, double> &b) { return a.first < b.first; });
        off.push_back((int64_t)nu.size());
        for (auto &c : cur) { nu.push_back(c.first); sig.push_back(c.second); }
        cur.clear();
    };
    char line[1024];
    double curT = NAN;
    int bad = 0;
    while (std::fgets(line, sizeof line, f)) {
        const char *c = line;
        while (*c == ' ' || *c == '\t') ++c;
        if (*c == '\0' || *c == '\n' || *c == '\r') continue;
        if (std::isalpha((unsigned char)*c)) {           // header: pair numin numax n T sigma_max ...
            flush();
            char name[256];
            double a0, a1, tT;
            long nn;
            if (std::sscanf(c, "%255s %lf %lf %ld %lf", name, &a0, &a1, &nn, &tT) != 5) { bad = 1; break; }
            curT = tT;
            T.push_back(curT);
        } else {
            double x, y;
            if (T.empty() || std::sscanf(c, "%lf %lf", &x, &y) != 2) { bad = 1; break; }
            cur.push_back({x, std::max(0.0, y)});
        }
    }
    std::fclose(f);
    if (bad) return fail(ctx, IRS_EINVAL, std::string("malformed CIA file ") + path);
    flush();
    // headers without data rows do not make a block
    if (off.size() != T.size()) return fail(ctx, IRS_EINVAL, std::string("CIA block without data in ") + path);
    off.push_back((int64_t)nu.size());
    if (T.empty()) return fail(ctx, IRS_EINVAL, std::string("no CIA blocks in ") + path);
    if ((int)T.size() > IRS_CIA_MAXB)
        return fail(ctx, IRS_EINVAL, "CIA table needs 0.." + std::to_string(IRS_CIA_MAXB) + " blocks");
    return irs_cia_set(ctx, pair, (int)T.size(), T.data(), off.data(), nu.data(), sig.data());
}

int irs_continuum_layer(irs_ctx *ctx, double nu0, double dnu, int64_t n_nu, int continua, double vmr_h2o,
                        double vmr_co2, double vmr_n2, double vmr_o2, double p_hpa, double T, double dz_cm,
                        double *tau_c) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (!tau_c || n_nu <= 0 || !(dnu > 0.0)) return fail(ctx, IRS_EINVAL, "bad grid");
    if (continua < 0 || continua > IRS_CONT_ALL) return fail(ctx, IRS_EINVAL, "bad continua mask");
    if (!(p_hpa > 0.0) || !(T > 0.0)) return fail(ctx, IRS_EINVAL, "non-positive pressure or temperature");
    ContLayer cl{};
    cl.vh2o = vmr_h2o;
    cl.vco2 = vmr_co2;
    cl.vn2 = vmr_n2;
    cl.vo2 = vmr_o2;
    cl.p_mid = p_hpa;
    cl.T_mid = T;
    cl.dz = dz_cm;
    HIPCHK(ensure(ctx->d_tmp4, sizeof(ContLayer)));
    HIPCHK(ensure(ctx->d_tmp5, sizeof(double) * n_nu));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp4.p, &cl, sizeof(cl), hipMemcpyHostToDevice, ctx->stream));
    const Grid g = make_grid(nu0, dnu, nu0 + (double)(n_nu - 1) * dnu);
    ContArgs co = cont_args(ctx, continua);
    co.n = n_nu;
    co.n_layers = 1;
    co.n_prof = 1;
    co.an = g.an;
    co.sn = g.sn;
    co.den = g.den;
    co.lay = (const ContLayer *)ctx->d_tmp4.p;
    co.tau64 = (double *)ctx->d_tmp5.p;
    if (continua) {
        HIPCHK(launch_continuum(co, ctx->stream));
        HIPCHK(hipMemcpyAsync(tau_c, ctx->d_tmp5.p, sizeof(double) * n_nu, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        std::memset(tau_c, 0, sizeof(double) * n_nu);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return IRS_OK;
}

int irs_state_size(const irs_state_spec *sp, int n_levels) {
    if (!sp || n_levels <= 0) return -1;
    int n = sp->include_temperature ? n_levels : 0;
    for (int r = 0; r < sp->n_retrieved; ++r) n += (sp->column_scale && sp->column_scale[r]) ? 1 : n_levels;
    return n + (sp->include_tsfc ? 1 : 0) + (sp->include_emissivity ? 1 : 0);
}

int irs_jacobian(irs_ctx *ctx, const irs_profile *pr, const irs_sounder *snd, const irs_fm_options *opt_in,
                 const irs_state_spec *sp, double *K, double *y0, double *x0) {
    if (!ctx) return IRS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (!pr || !snd || !sp || !K || !y0 || !x0) return fail(ctx, IRS_EINVAL, "null argument");
    irs_fm_options opt;
    if (opt_in) opt = *opt_in; else irs_default_options(&opt);
    if (sp->observable != 0 && sp->observable != 1) return fail(ctx, IRS_EINVAL, "observable must be :bt or :radiance");
    if (sp->n_retrieved < 0 || sp->n_retrieved > IRS_JAC_MAXR) return fail(ctx, IRS_EINVAL, "at most 4 retrieved species");
    if (!(pr->mu > 0.0 && pr->mu <= 1.0)) return fail(ctx, IRS_EINVAL, "μ must be in (0, 1]");
    if (opt.source_function != IRS_SRC_TOON && opt.source_function != IRS_SRC_CIM)
        return fail(ctx, IRS_EINVAL, "source_function must be :toon or :cim");
    if (opt.continua < 0 || opt.continua > IRS_CONT_ALL) return fail(ctx, IRS_EINVAL, "bad continua mask");
    if (opt.line_mixing == IRS_LM_RELMAT &&
        (opt.relmat < 0 || opt.relmat >= (int32_t)ctx->relmats.size() || !ctx->relmats[opt.relmat]))
        return fail(ctx, IRS_EINVAL, "bad relaxation-matrix handle");
    auto t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    const Grid g = internal_grid(snd, &opt, nullptr);
    const Grid gc = make_grid(snd->nu_min, snd->dnu, snd->nu_max);
    const int64_t n = g.n, nch = gc.n;
    const double nu_lo = g.at(0), dnu = g.step;
    const int nlev = pr->n_levels, L = nlev - 1, ns = pr->n_species, nr = sp->n_retrieved;
    std::vector<int> retr_sidx(nr, -1);
    for (int r = 0; r < nr; ++r)
        for (int s = 0; s < ns; ++s)
            if (pr->species[s] == sp->retrieved[r]) retr_sidx[r] = s;
    for (int r = 0; r < nr; ++r)
        if (retr_sidx[r] < 0) return fail(ctx, IRS_EINVAL, "profile is missing a retrieved species");
    for (int s = 0; s < ns; ++s)
        if (pr->linelist[s] >= 0) {
            int rc = check_list(ctx, pr->linelist[s]);
            if (rc) return rc;
        }
    LayerProps lp;
    int rc = layer_properties(ctx, pr, opt.T_method, lp);
    if (rc) return rc;
    rc = ensure_common(ctx);
    if (rc) return rc;
    std::vector<int32_t> used;
    for (int s = 0; s < ns; ++s) if (pr->linelist[s] >= 0) used.push_back(pr->linelist[s]);
    std::map<int32_t, int32_t> koff;
    rc = build_keys(ctx, used, nu_lo, dnu, koff);
    if (rc) return rc;
    const double *P = pr->p_hpa, *Tl = pr->T;
    // ---- work items (full line set: the cutoff-freezing policy of jacobian_fd.jl:10-24)
    std::vector<JacItemDev> items((size_t)L * ns);
    int64_t rec = 0;
    int max_lines = 0;
    for (int k = 0; k < L; ++k)
        for (int s = 0; s < ns; ++s) {
            JacItemDev &it = items[(size_t)k * ns + s];
            std::memset(&it, 0, sizeof(it));
            it.retr = -1;
            const int32_t h = pr->linelist[s];
            const double vmr = lp.vmr_cg[s][k];
            if (h < 0 || vmr == 0.0 || ctx->lists[h].n == 0) continue;
            const ListHost &lh = ctx->lists[h];
            const int spid = pr->species[s];
            const bool lm = opt.line_mixing == IRS_LM_VPY && spid == 2 && lh.has_lm;
            const double p_atm = lp.p_cg[s][k] / 1013.25, T = lp.T_cg[s][k];
            const double vs = (spid == 1 && !lm) ? vmr : 0.0;
            ItemDev base = make_item(lh, h, T, p_atm, vs, vmr * lp.dp[k] * 2.1209e22, 0.0, opt.cutoff,
                                     opt.lm_cutoff, opt.x_far, lm, dnu);
            it.rec_off = rec;
            it.n = base.n; it.key_off = koff[h]; it.list = h; it.lm = base.lm; it.active = 1; it.M = base.M;
            it.C = base.C; it.Clm = base.Clm;
            it.T = T; it.p_atm = p_atm; it.vmr_self = vs; it.coef = base.coef; it.cutoff = opt.cutoff;
            it.x_far = opt.x_far; it.lm_cutoff = opt.lm_cutoff; it.kcore_per_nu = base.kcore_per_nu;
            const double p1 = P[k], p2 = P[k + 1];
            const double wT = opt.T_method == IRS_TMETHOD_MASS_WEIGHTED ? cg_mass_frac(p1, p2)
                                                                         : std::log(lp.p_cg[s][k] / p1) / std::log(p2 / p1);
            it.wT_lo = 1.0 - wT;
            it.wT_hi = wT;
            for (int r = 0; r < nr; ++r) {
                if (retr_sidx[r] != s) continue;
                it.retr = r;
                const double *vl = pr->vmr + (size_t)s * nlev;
                const double v1 = vl[k], v2 = vl[k + 1];
                double gV1, gV2, gP1 = 0, gP2 = 0, gT1 = 0, gT2 = 0;
                cg_column_vmr_grad(v1, v2, p1, p2, gV1, gV2);
                if (sp->vmr_coupling) {
                    const double h1 = std::max(std::fabs(v1), 1e-300) * 1e-6, h2 = std::max(std::fabs(v2), 1e-300) * 1e-6;
                    gP1 = (cg_pressure(v1 + h1, v2, p1, p2) - cg_pressure(v1 - h1, v2, p1, p2)) / (2 * h1);
                    gP2 = (cg_pressure(v1, v2 + h2, p1, p2) - cg_pressure(v1, v2 - h2, p1, p2)) / (2 * h2);
                    if (opt.T_method != IRS_TMETHOD_MASS_WEIGHTED) {
                        const double dtdp = (Tl[k + 1] - Tl[k]) / (lp.p_cg[s][k] * std::log(p2 / p1));
                        gT1 = dtdp * gP1;
                        gT2 = dtdp * gP2;
                    }
                    if (spid == 1) { it.gs_lo = gV1; it.gs_hi = gV2; }
                }
                it.gv_lo = gV1 / vmr; it.gv_hi = gV2 / vmr;
                it.gp_lo = gP1 / 1013.25; it.gp_hi = gP2 / 1013.25;
                it.gt_lo = gT1; it.gt_hi = gT2;
            }
            rec += it.n;
            max_lines = std::max(max_lines, it.n);
        }
    HIPCHK(ensure(ctx->j_items, sizeof(JacItemDev) * items.size()));
    HIPCHK(ensure(ctx->j_g, sizeof(float4) * 6 * std::max<int64_t>(rec, 1)));
    HIPCHK(hipMemcpyAsync(ctx->j_items.p, items.data(), sizeof(JacItemDev) * items.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    float4 *gb = (float4 *)ctx->j_g.p;
    const int64_t R6 = std::max<int64_t>(rec, 1);
    JacPrepArgs pa{};
    pa.items = (const JacItemDev *)ctx->j_items.p;
    pa.lists = (const LineListDev *)ctx->d_lists.p;
    pa.tips_q = (const double *)ctx->d_tips.p;
    pa.n_items = (int)items.size();
    pa.nu_lo = nu_lo;
    pa.dnu = dnu;
    pa.g0 = gb; pa.g1 = gb + R6; pa.g2 = gb + 2 * R6; pa.g3 = gb + 3 * R6; pa.g4 = gb + 4 * R6; pa.g5 = gb + 5 * R6;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    HIPCHK(launch_prep_grad(pa, max_lines, ctx->stream));
    const size_t NL = (size_t)n * L;
    HIPCHK(ensure(ctx->j_tau, sizeof(float) * NL));
    HIPCHK(ensure(ctx->j_dt, sizeof(float) * NL * 2));
    HIPCHK(ensure(ctx->j_v, sizeof(float) * NL * 2 * std::max(nr, 1)));
    JacSumArgs sa{};
    sa.items = pa.items;
    sa.g0 = pa.g0; sa.g1 = pa.g1; sa.g2 = pa.g2; sa.g3 = pa.g3; sa.g4 = pa.g4; sa.g5 = pa.g5;
    sa.keys = (const int32_t *)ctx->d_keys.p;
    sa.n_hi = n;
    sa.n_layers = L;
    sa.n_species = ns;
    sa.n_retr = nr;
    sa.dnu = dnu;
    sa.nu_lo = nu_lo;
    far_params(ctx, dnu, IRS_JAC_TILE_P, sa.far_d, sa.far_hc);
    // the gradient sum's per-pair lines cost seven sums each: half the far-wing
    // distance and half the node spacing (the same HC / D, so the same interpolation
    // bound) move half of them onto coarse nodes (149.5 -> 139.7 ms on full IASI)
    if (sa.far_hc >= 4 && (256 * IRS_JAC_TILE_P) / (sa.far_hc / 2) + 3 <= IRS_JAC_FAR_MAXNC) {
        sa.far_hc /= 2;
        sa.far_d = (sa.far_d + 1) / 2;
    }
    sa.tau = (float *)ctx->j_tau.p;
    sa.dtlo = (float *)ctx->j_dt.p;
    sa.dthi = sa.dtlo + NL;
    sa.vlo = (float *)ctx->j_v.p;
    sa.vhi = sa.vlo + NL * std::max(nr, 1);
    if (opt.line_mixing == IRS_LM_RELMAT) {
        // the line-mixing perturbation of the relaxation matrix's species and its central
        // differences in T (0.02 K) and p (1e-4 relative) per layer, the terms in FP64
        // (cross_section_jacobian.jl:256-257,283-306)
        int s_lm = -1;
        for (int s = 0; s < ns; ++s)
            if (pr->species[s] == opt.lm_species) { s_lm = s; break; }
        if (s_lm >= 0) {
            const RelmatHost &rmh = *ctx->relmats[opt.relmat];
            const double hT = 0.02;
            std::vector<std::vector<LmTerm>> units((size_t)5 * L);
            std::vector<std::string> errs((size_t)L);
            const int nth = std::max(1, std::min<int>(L, std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
            std::vector<std::thread> pool;
            for (int w = 0; w < nth; ++w)
                pool.emplace_back([&, w]() {
                    for (int k = w; k < L; k += nth) {
                        const JacItemDev &it = items[(size_t)k * ns + s_lm];
                        if (!it.active) continue;
                        const double T = it.T, pa0 = it.p_atm, hp = std::max(pa0, 1e-6) * 1e-4;
                        const double Ts[5] = {T, T + hT, T - hT, T, T};
                        const double ps[5] = {pa0, pa0, pa0, pa0 + hp, pa0 - hp};
                        for (int v = 0; v < 5 && errs[k].empty(); ++v) {
                            errs[k] = relmat_terms(rmh, opt.relmat_band_mode, Ts[v], ps[v], opt.keep_threshold,
                                                   units[(size_t)5 * k + v]);
                            for (auto &tm : units[(size_t)5 * k + v]) tm.fp64 = 1;
                        }
                    }
                });
            for (auto &t : pool) t.join();
            for (const auto &e : errs)
                if (!e.empty()) return fail(ctx, IRS_EINVAL, e);
            HIPCHK(ensure(ctx->d_tau_lm, sizeof(float) * 3 * NL));
            HIPCHK(ensure(ctx->d_lm_out, sizeof(double) * 5 * (size_t)n));
            HIPCHK(hipMemsetAsync(ctx->d_tau_lm.p, 0, sizeof(float) * 3 * NL, ctx->stream));
            float *lm0 = (float *)ctx->d_tau_lm.p, *lmT = lm0 + NL, *lmP = lmT + NL;
            for (int k = 0; k < L; ++k) {
                const JacItemDev &it = items[(size_t)k * ns + s_lm];
                if (!it.active) continue;
                std::vector<std::vector<LmTerm>> five(units.begin() + 5 * k, units.begin() + 5 * k + 5);
                int rc2 = upload_lm_terms(ctx, five);
                if (rc2) return rc2;
                LmTermArgs la{};
                la.terms = (const LmTerm *)ctx->d_lm_terms.p;
                la.off = (const int64_t *)ctx->d_lm_off.p;
                la.n = n;
                la.n_units = 5;
                la.an = g.an;
                la.sn = g.sn;
                la.den = g.den;
                la.cutoff = opt.lm_cutoff;
                la.x_far = opt.x_far;
                la.out64 = (double *)ctx->d_lm_out.p;
                HIPCHK(launch_lm_terms(la, ctx->stream));
                HIPCHK(launch_lm_fd((const double *)ctx->d_lm_out.p, n, it.coef, hT, std::max(it.p_atm, 1e-6) * 1e-4,
                                    lm0 + (size_t)k * n, lmT + (size_t)k * n, lmP + (size_t)k * n, ctx->stream));
            }
            sa.lm0 = lm0;
            sa.lmT = lmT;
            sa.lmP = lmP;
            sa.lm_s = s_lm;
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    HIPCHK(launch_grad_linesum(sa, ctx->stream));
    if (opt.continua) {
        // continuum optical depth and its analytic T / H2O / CO2 derivatives
        // (vmr_jacobian.jl:288-349: the VMR terms only with vmr_coupling, as there)
        std::vector<ContLayer> cl(L);
        cont_layers(pr, lp, cl.data());
        HIPCHK(ensure(ctx->d_tmp4, sizeof(ContLayer) * L));
        HIPCHK(hipMemcpyAsync(ctx->d_tmp4.p, cl.data(), sizeof(ContLayer) * L, hipMemcpyHostToDevice, ctx->stream));
        ContArgs co = cont_args(ctx, opt.continua);
        co.tau = sa.tau;
        co.n = n;
        co.n_layers = L;
        co.n_prof = 1;
        co.an = g.an;
        co.sn = g.sn;
        co.den = g.den;
        co.lay = (const ContLayer *)ctx->d_tmp4.p;
        if (sp->include_temperature) {
            co.dtlo = sa.dtlo;
            co.dthi = sa.dthi;
        }
        for (int r = 0; r < nr && sp->vmr_coupling; ++r) {
            const int spid = sp->retrieved[r];
            float *lo = sa.vlo + (size_t)r * NL, *hi = sa.vhi + (size_t)r * NL;
            if (spid == 1) { co.vlo_h = lo; co.vhi_h = hi; }
            if (spid == 2) { co.vlo_c = lo; co.vhi_c = hi; }
        }
        HIPCHK(launch_continuum(co, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    // ---- RTE Jacobian
    std::vector<double> tlev(Tl, Tl + nlev), tave(L), frac(L);
    for (int k = 0; k < L; ++k) {
        tave[k] = cg_temperature_mass(Tl[k], Tl[k + 1], P[k], P[k + 1]);
        frac[k] = cg_mass_frac(P[k], P[k + 1]);
    }
    HIPCHK(ensure(ctx->d_tmp3, sizeof(double) * (3 * nlev + 8)));
    double *dT = (double *)ctx->d_tmp3.p;
    HIPCHK(hipMemcpyAsync(dT, tlev.data(), sizeof(double) * nlev, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dT + nlev, tave.data(), sizeof(double) * L, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dT + 2 * nlev, frac.data(), sizeof(double) * L, hipMemcpyHostToDevice, ctx->stream));
    const size_t rte_doubles = (size_t)n * (3 + L + (L + 1));
    HIPCHK(ensure(ctx->j_rte, sizeof(double) * rte_doubles));
    double *rb = (double *)ctx->j_rte.p;
    RteJacArgs ra{};
    ra.tau = sa.tau;
    ra.n = n;
    ra.n_layers = L;
    ra.T_lev = dT;
    ra.T_ave = dT + nlev;
    ra.frac = dT + 2 * nlev;
    ra.T_sfc = pr->T_sfc > 0.0 ? pr->T_sfc : Tl[0];
    ra.mu = pr->mu;
    ra.eps = pr->eps_sfc;
    ra.source_function = opt.source_function;
    ra.an = g.an; ra.sn = g.sn; ra.den = g.den;
    ra.I = rb; ra.dtsfc = rb + n; ra.deps = rb + 2 * n; ra.dtau = rb + 3 * n; ra.dTlev = rb + 3 * n + NL;
    HIPCHK(hipMemsetAsync(ra.dTlev, 0, sizeof(double) * (size_t)n * (L + 1), ctx->stream));
    HIPCHK(launch_rte_jac(ra, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    // ---- columns
    std::vector<JacColDesc> cols;
    std::vector<double> lw((size_t)std::max(nr, 1) * nlev, 0.0);
    int ci = 0;
    if (sp->include_temperature)
        for (int j = 0; j < nlev; ++j) { cols.push_back({0, j, j, 0}); x0[ci++] = Tl[j]; }
    for (int r = 0; r < nr; ++r) {
        const double *vl = pr->vmr + (size_t)retr_sidx[r] * nlev;
        const bool cs = sp->column_scale && sp->column_scale[r];
        for (int j = 0; j < nlev; ++j) lw[(size_t)r * nlev + j] = (cs || sp->log_vmr) ? vl[j] : 1.0;
        if (cs) { cols.push_back({1, 0, L, r}); x0[ci++] = 0.0; }
        else
            for (int j = 0; j < nlev; ++j) {
                cols.push_back({1, j, j, r});
                x0[ci++] = sp->log_vmr ? std::log(std::max(vl[j], 1e-30)) : vl[j];
            }
    }
    if (sp->include_tsfc) { cols.push_back({2, 0, 0, 0}); x0[ci++] = ra.T_sfc; }
    if (sp->include_emissivity) { cols.push_back({3, 0, 0, 0}); x0[ci++] = pr->eps_sfc; }
    const int ncol = (int)cols.size();
    if (ncol == 0) return fail(ctx, IRS_EINVAL, "empty state vector");
    HIPCHK(ensure(ctx->j_desc, sizeof(JacColDesc) * ncol));
    HIPCHK(ensure(ctx->j_w, sizeof(double) * lw.size()));
    HIPCHK(ensure(ctx->j_cols, sizeof(double) * (size_t)n * ncol));
    HIPCHK(hipMemcpyAsync(ctx->j_desc.p, cols.data(), sizeof(JacColDesc) * ncol, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->j_w.p, lw.data(), sizeof(double) * lw.size(), hipMemcpyHostToDevice, ctx->stream));
    JacColArgs ca{};
    ca.cols = (const JacColDesc *)ctx->j_desc.p;
    ca.n = n;
    ca.n_layers = L;
    ca.dtau = ra.dtau; ca.dTlev = ra.dTlev; ca.dtsfc = ra.dtsfc; ca.deps = ra.deps;
    ca.level_w = (const double *)ctx->j_w.p;
    ca.dtlo = sa.dtlo; ca.dthi = sa.dthi; ca.vlo = sa.vlo; ca.vhi = sa.vhi;
    ca.out = (double *)ctx->j_cols.p;
    // level_w is indexed [r][L+1]
    HIPCHK(launch_jac_columns(ca, ncol, ctx->stream));
    // ---- channel tail: I -> R at channels (for y0 and dBT/dR), then every column
    double *taps = nullptr;
    int64_t half = 0;
    std::vector<double> one(1, 1.0);
    if (opt.with_ils) {
        rc = get_taps(ctx, dnu, snd->opd_max, snd->fwhm_gauss, opt.apodization, 2048, &taps, &half);
        if (rc) return rc;
    } else {
        HIPCHK(ensure(ctx->d_tmp0, sizeof(double)));
        HIPCHK(hipMemcpyAsync(ctx->d_tmp0.p, one.data(), sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        taps = (double *)ctx->d_tmp0.p;
    }
    std::vector<int64_t> cidx;
    std::vector<double> cw, cnu;
    channel_positions(g, gc, cidx, cw, cnu);
    HIPCHK(ensure(ctx->d_chidx, sizeof(int64_t) * nch));
    HIPCHK(ensure(ctx->d_chw, sizeof(double) * nch));
    HIPCHK(ensure(ctx->d_chnu, sizeof(double) * nch));
    HIPCHK(ensure(ctx->d_rch, sizeof(double) * nch));
    HIPCHK(ensure(ctx->d_bt, sizeof(double) * nch));
    HIPCHK(hipMemcpyAsync(ctx->d_chidx.p, cidx.data(), sizeof(int64_t) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chw.p, cw.data(), sizeof(double) * nch, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_chnu.p, cnu.data(), sizeof(double) * nch, hipMemcpyHostToDevice, ctx->stream));
    const int64_t ch_stride = channel_stride(cidx, cw);
    IlsArgs ia{};
    ia.spec = ra.I;
    ia.n = n;
    ia.taps = taps;
    ia.half = opt.with_ils ? half : 0;
    ia.dnu = opt.with_ils ? dnu : 1.0;
    ia.ch_idx = (const int64_t *)ctx->d_chidx.p;
    ia.ch_w = (const double *)ctx->d_chw.p;
    ia.ch_nu = (const double *)ctx->d_chnu.p;
    ia.n_ch = nch;
    ia.r_out = (double *)ctx->d_rch.p;
    ia.bt_out = (double *)ctx->d_bt.p;
    IlsBlockArgs ib{};
    ib.n = n;
    ib.taps = taps;
    ib.half = ia.half;
    ib.dnu = ia.dnu;
    ib.idx0 = ch_stride > 0 ? cidx[0] : 0;
    ib.stride = (int)ch_stride;
    ib.ch_nu = ia.ch_nu;
    ib.n_ch = nch;
    if (ch_stride > 0) {
        ib.spec = ra.I;
        ib.r_out = ia.r_out;
        ib.bt_out = ia.bt_out;
        HIPCHK(launch_ils_block(ib, 1, ctx->stream));
    } else {
        HIPCHK(launch_ils_channels(ia, ctx->stream));
    }
    std::vector<double> Rc(nch), BTc(nch), fac(nch);
    HIPCHK(hipMemcpyAsync(Rc.data(), ctx->d_rch.p, sizeof(double) * nch, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(BTc.data(), ctx->d_bt.p, sizeof(double) * nch, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (int64_t c = 0; c < nch; ++c) {
        fac[c] = sp->observable == 0 ? dbt_dr(cnu[c], Rc[c]) : 1.0;
        y0[c] = sp->observable == 0 ? BTc[c] : Rc[c];
    }
    HIPCHK(ensure(ctx->j_fac, sizeof(double) * nch));
    HIPCHK(ensure(ctx->j_K, sizeof(double) * (size_t)nch * ncol));
    HIPCHK(hipMemcpyAsync(ctx->j_fac.p, fac.data(), sizeof(double) * nch, hipMemcpyHostToDevice, ctx->stream));
    JacIlsArgs ja{};
    ja.cols = ca.out;
    ja.n = n;
    ja.taps = taps;
    ja.half = ia.half;
    ja.dnu = ia.dnu;
    ja.ch_idx = ia.ch_idx;
    ja.ch_w = ia.ch_w;
    ja.factor = (const double *)ctx->j_fac.p;
    ja.n_ch = nch;
    ja.K = (double *)ctx->j_K.p;
    if (ch_stride > 0) {
        ib.spec = ca.out;
        ib.factor = (const double *)ctx->j_fac.p;
        ib.r_out = (double *)ctx->j_K.p;
        ib.bt_out = nullptr;
        HIPCHK(launch_ils_block(ib, ncol, ctx->stream));
    } else {
        HIPCHK(launch_jac_ils(ja, ncol, ctx->stream));
    }
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    HIPCHK(hipMemcpyAsync(K, ctx->j_K.p, sizeof(double) * (size_t)nch * ncol, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->stats.ms_prep = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->stats.ms_linesum = elapsed_ms(ctx->ev[1], ctx->ev[2]);
    ctx->stats.ms_rte = elapsed_ms(ctx->ev[2], ctx->ev[3]);
    ctx->stats.ms_ils = elapsed_ms(ctx->ev[3], ctx->ev[4]);
    ctx->stats.n_hi = n;
    ctx->stats.n_layers = L;
    ctx->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return IRS_OK;
}

}  // extern "C"
