// Argument blocks of the Jacobian kernels (k_jac.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "irs_internal.h"

#define IRS_JAC_MAXR 4   // retrieved species handled in one gradient pass
#define IRS_JAC_TILE_P 4 // grid points per thread of the gradient line sum (tile = 256 * P)
#define IRS_JAC_FAR_MAXNC 128   // coarse far-wing nodes per gradient tile

// Per (layer, species) gradient work item.
struct JacItemDev {
    int64_t rec_off;
    int32_t n, key_off, list, lm, active, M;
    float C, Clm;
    double T, p_atm, vmr_self, coef, cutoff, x_far;
    double lm_cutoff;             // cm^-1, the VP_Y window (Clm = lm_cutoff / dnu)
    double kcore_per_nu;          // grid steps per cm^-1 of wavenumber bounding |x| <= x_far
    double wT_lo, wT_hi;          // dT_cg/dT_lev at the lower / upper level
    int32_t retr, pad;            // retrieved-species index or -1
    double gv_lo, gv_hi;          // d vmr_cg / d v_level / vmr_cg   (number-density term)
    double gp_lo, gp_hi;          // d p_cg / d v_level / 1013.25   (p in atm)
    double gt_lo, gt_hi;          // d T_cg / d v_level
    double gs_lo, gs_hi;          // d vmr_self / d v_level (H2O only, else 0)
};

struct JacPrepArgs {
    const JacItemDev *items;
    const LineListDev *lists;
    const double *tips_q;
    int n_items;
    double nu_lo, dnu;
    float4 *g0, *g1, *g2, *g3, *g4, *g5;
};

struct JacSumArgs {
    const JacItemDev *items;      // [n_layers][n_species]
    const float4 *g0, *g1, *g2, *g3, *g4, *g5;
    const int32_t *keys;
    int64_t n_hi;
    int n_layers, n_species, n_retr;
    float *tau, *dtlo, *dthi, *vlo, *vhi;   // [L][n]; vlo/vhi [r][L][n]
    double dnu;                   // internal grid step (cm^-1)
    double nu_lo;                 // wavenumber of grid point 0
    int far_d, far_hc;            // coarse far-wing split (grid steps); far_hc = 0 turns it off
    // optional [L][n]: relaxation-matrix line-mixing perturbation of species lm_s and its
    // T and p derivatives (coef folded), added before that species' final clamp
    const float *lm0, *lmT, *lmP;
    int lm_s;
};

struct RteJacArgs {
    const float *tau;             // [L][n]
    int64_t n;
    int n_layers;
    const double *T_lev, *T_ave, *frac;
    double T_sfc, mu, eps;
    int source_function;
    double an, sn, den;
    double *I, *dtau, *dTlev, *dtsfc, *deps;   // dtau [L][n], dTlev [L+1][n] (zeroed)
};

struct JacColDesc {
    int kind;                     // 0 T level, 1 VMR, 2 T_sfc, 3 emissivity
    int level, level_end, r;
};

struct JacColArgs {
    const JacColDesc *cols;
    int64_t n;
    int n_layers;
    const double *dtau, *dTlev, *dtsfc, *deps, *level_w;   // level_w [r][L+1]
    const float *dtlo, *dthi, *vlo, *vhi;
    double *out;                  // [n_cols][n]
};

struct JacIlsArgs {
    const double *cols;           // [n_cols][n]
    int64_t n;
    const double *taps;
    int64_t half;
    double dnu;
    const int64_t *ch_idx;
    const double *ch_w, *factor;
    int64_t n_ch;
    double *K;                    // [n_cols][n_ch]
};

hipError_t launch_prep_grad(const JacPrepArgs &a, int max_lines, hipStream_t s);
hipError_t launch_grad_linesum(const JacSumArgs &a, hipStream_t s);
hipError_t launch_lm_fd(const double *d, int64_t n, double coef, double hT, double hp, float *lm0, float *lmT,
                        float *lmP, hipStream_t s);
hipError_t launch_rte_jac(const RteJacArgs &a, hipStream_t s);
hipError_t launch_jac_columns(const JacColArgs &a, int n_cols, hipStream_t s);
hipError_t launch_jac_ils(const JacIlsArgs &a, int n_cols, hipStream_t s);
