// Internal declarations shared by the HIP kernels and the C-ABI host code.
#pragma once

#include <string>

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IRS_SQRT_LN2    0.83255461115769775635
#define IRS_INV_SQRT_PI 0.56418958354775628695
#define IRS_PI          3.14159265358979323846

// Work counters: IRS_CNT_SLOTS copies of 8 counters (the host sums the copies);
// each workgroup adds into the copy of its slot, so that concurrent atomics are
// spread over many addresses instead of serialising on one.
//   [0] active lines, [1] Voigt pairs, [2] dispersive pairs (prep)
//   [3] line x 64-point slice expansions (EDGE and far FAST lines), [4] fine pairs, [5] coarse-node pairs,
//   [6], [7] the same for VP_Y items (line sum)
#define IRS_CNT_SLOTS 16
#define IRS_NCNT (8 * IRS_CNT_SLOTS)

// Line-sum tile: 256 threads x P points, thread t owns points i0 + t + 256*p.
#define IRS_LS_THREADS 256

// |z| from which the line sum evaluates w(z) by its asymptotic series (the
// mid-core records written by the prep kernel); exact FP32 w(z) below it.
#define IRS_Z1 8.0

// One uploaded line list on the device (structure of arrays, sorted by nu).
struct LineListDev {
    int64_t n;
    const int8_t *mol, *iso;
    const double *nu, *S, *elower, *ylm;
    const float *gair, *gself, *nair, *dair;
    const int16_t *tips_row;       // row of the partition table per line, -1: none
};

// Per (profile, layer, species) work item.  Written by the host, read by the
// prep and line-sum kernels with wave-uniform indices (scalar loads).
struct ItemDev {
    int64_t rec_off;        // first record of this item in rec0/rec1/rec2
    int32_t n;              // lines in the list
    int32_t key_off;        // offset of the list's key array
    int32_t list;           // index into the line-list table
    int32_t lm;             // 1: VP_Y dispersive term active
    int32_t active;         // 0: skip (no lines or vmr == 0)
    int32_t M;              // max |pressure shift| in grid steps, rounded up, + 2
    float   C, Clm;         // cutoff / lm_cutoff in grid steps
    double  kcore_per_nu;   // x_far * max Doppler coefficient * sqrt(T) / (sqrt(ln2) * dnu)
    double  T, p_atm, vmr_self, coef, dptmn, cutoff, x_far;
    double  lm_cutoff;      // cm^-1 (the VP_Y window; lm_cutoff / dnu = Clm)
    int64_t ktab_off;       // offset of the list's key table (LinesumArgs::ktab)
    int64_t pb_off;         // first 256-line block of this item in PrepArgs::pblk
};

struct PrepArgs {
    const ItemDev *items;
    const LineListDev *lists;
    const double *tips_q;          // [n_rows][331]
    int n_items;
    double nu_lo, dnu;
    float4 *rec0, *rec1, *rec2;
    int32_t *keys;                 // per list, floor((nu - nu_lo)/dnu)
    unsigned long long *counters;  // [IRS_CNT_SLOTS][8] work counters (slots 0..2 of each copy)
    int64_t n_hi;
    double *pre;                   // optional [3][n_rec]: per-item inclusive prefix sums of the
    int64_t n_rec;                 // pedestal amplitude P, P c and P c^2 (c: centre, grid steps)
    float4 *rec3;                  // optional [n_rec][6]: mid-core records (Voigt, then dispersive)
    double *pblk;                  // with pre: [sum over items of ceil(n/256)][3] block sums of the scan
    double *rec64;                 // optional [n_rec][8] FP64 line records of the every-pair FP64 sum:
                                   // {shifted centre, f, y, Snorm (0 if rejected), Hc, YSn, fd, yd}
};

// Every-pair FP64 line sum (k_ls64.hip), one profile: tau [n_layers][n_hi] FP64.
struct Linesum64Args {
    const ItemDev *items;          // [n_layers][n_species]
    const double *rec64;           // PrepArgs::rec64
    const int32_t *keys;
    double *tau;
    int64_t n_hi;
    int n_layers, n_species;
    double dnu, an, sn, den;       // grid node i = (an + i*sn)/den
};

struct LinesumArgs {
    const ItemDev *items;          // [n_prof][n_layers][n_species]
    const float4 *rec0, *rec1, *rec2;
    const int32_t *keys;
    float *out;                    // [n_prof][n_layers][n_hi]
    int64_t n_hi;
    int n_layers, n_species;
    double nu_lo, dnu;
    int tiles, n_prof;
    const int32_t *tile_order;     // optional tile permutation (heavy tiles first)
    unsigned *queue;               // work counter of the persistent kernel (zero at launch)
    int far_d, far_hc;             // coarse far-wing split: distance and node spacing in grid steps (0: off)
    const double *pre;             // [3][n_rec] pedestal prefix sums (PrepArgs::pre)
    int64_t n_rec;
    const float4 *rec3;            // [n_rec][6] mid-core records (PrepArgs::rec3)
    unsigned long long *work;      // optional: counters + 3; per copy [0] line-slice expansions,
                                   // [1] fine pairs evaluated, [2] coarse-node pairs, [3], [4] the
                                   // same for VP_Y items
    const int32_t *ktab;           // optional key tables: per list, ktab[off + v - ktab_lo] = first line
    int64_t ktab_lo, ktab_n;       // with key >= v, for v in [ktab_lo, ktab_lo + ktab_n)
};

struct RteArgs {
    const float *tau32;            // [n_prof][n_layers][n] (one of tau32/tau64)
    const double *tau64;
    int64_t n;
    int n_layers;
    const double *T_lev, *T_ave;   // device, [n_prof][L+1], [n_prof][L]
    double T_sfc, mu, eps;         // used when scal is null
    const double *scal;            // optional [n_prof][3] = {T_sfc, mu, eps}
    int source_function;
    double an, sn, den;            // grid node i = (an + i*sn)/den
    double *out;                   // [n_prof][n]
    int n_prof;                    // blockIdx.y = profile (0 is taken as 1)
};

struct IlsArgs {
    const double *spec;            // [n]
    int64_t n;
    const double *taps;            // [2*half+1]
    int64_t half;
    double dnu;
    const int64_t *ch_idx;         // left node of each channel
    const double *ch_w;            // interpolation weight of the right node (0: exact)
    const double *ch_nu;
    int64_t n_ch;
    double *r_out, *bt_out;        // [n_prof][n_ch]
    int n_prof;                    // blockIdx.y = profile; spec is [n_prof][n] (0 is taken as 1)
};

// Channel ILS on equally spaced exact-node channels (ils_block_kernel):
// channel c is the ILS correlation of column col at node idx0 + c * stride.
struct IlsBlockArgs {
    const double *spec;            // [n_cols][n]
    int64_t n;
    const double *taps;            // [2*half+1]
    int64_t half;
    double dnu;
    int64_t idx0;
    int stride;
    const double *ch_nu;           // for bt_out
    const double *factor;          // optional per-channel factor (dBT/dR of the Jacobian columns)
    int64_t n_ch;
    double *r_out, *bt_out;        // [n_cols][n_ch]; bt_out optional
};

// Channel optical depths (k_chod.hip)
struct ChodArgs {
    const float *tau;              // [n_prof][n_layers][n]
    int64_t n;
    int n_layers;
    const double *taps;            // [2*half+1], area-normalised with dnu
    int64_t half;
    double dnu;
    const int64_t *ch_idx;         // exact internal node of each channel
    int64_t n_ch;
    double floor;                  // level-to-space channel transmittance below which a layer
                                   // takes its own (layer-alone) channel optical depth
    float od_max;
    const double *scal;            // [n_prof][3] = {T_sfc, mu, eps}: the path is tau / mu
    float *out;                    // [n_prof][n_layers][n_ch]
    float *lev_out;                // optional [n_prof][n_layers+1][n_ch] level-to-space channel transmittance
    const float *lev_in;           // = lev_out, read by the CIM pass
    float *g_out;                  // optional [n_prof][n_layers][n_ch] CIM channel weight
    int raw_G;                     // 1: store G_k unnormalised
    float *m_out;                  // optional [n_prof][n_layers][n_ch] normalised ILS first moment
                                   // ILS (x) [(nu - nu_c)(T_{k+1} - T_k)] / T^ch_{k+1} (exact-node channels only)
};

// launchers (k_*.hip)
hipError_t launch_prep(const PrepArgs &a, int max_lines, hipStream_t s);
hipError_t launch_keytab(const int32_t *keys, int n, int64_t lo, int64_t count, int32_t *out, hipStream_t s);
hipError_t launch_faddeeva32(const double *x, const double *y, int64_t n, int mode, double *re, double *im,
                             hipStream_t s);
hipError_t launch_linesum(const LinesumArgs &a, int P, int persistent_blocks, hipStream_t s);
hipError_t launch_reset_counters(unsigned *queue, unsigned long long *counters, int n_counters, hipStream_t s);
int linesum_persistent_capacity(int P);
hipError_t launch_chod(const ChodArgs &a, int n_prof, hipStream_t s);
// exact-node channels idx_c = idx0 + c stride: rows [chod_banded_rows][n] FP32 and
// sums [chod_banded_rows][n_ch] FP64 scratch of one profile
int64_t chod_banded_rows(int n_layers, bool cim);
// msums (needed with m_out): [(n_layers + 1) n_ch + 2 half + 1] FP64 scratch
hipError_t launch_chod_banded(const ChodArgs &a, int n_prof, int64_t idx0, int stride, float *rows, double *sums,
                              double *msums, hipStream_t s);
hipError_t launch_rte(const RteArgs &a, hipStream_t s);
hipError_t launch_ils_taps(const double *A, const double *L, int n_pts, double dL, double dnu,
                           int64_t half, double *taps, hipStream_t s);
hipError_t launch_ils_channels(const IlsArgs &a, hipStream_t s);
hipError_t launch_ils_block(const IlsBlockArgs &a, int n_cols, hipStream_t s);
hipError_t launch_linesum64(const Linesum64Args &a, hipStream_t s);

// ---- HITRAN .par ingest (k_par.hip, irs_par.hip) ------------------------------
struct ParFields {                 // device SoA of parsed records
    int8_t *mol, *iso;
    double *nu, *S, *A, *elower;
    float *gamma_air, *gamma_self, *n_air, *delta_air;
};
hipError_t launch_par_count(const unsigned char *b, int64_t n, int64_t *blk, int64_t n_blk, hipStream_t s);
hipError_t launch_par_newlines(const unsigned char *b, int64_t n, const int64_t *blk_off, int64_t n_blk, int64_t *nl,
                               hipStream_t s);
hipError_t launch_par_parse(const unsigned char *b, int64_t n, const int64_t *nl, int64_t n_nl, int64_t n_rec,
                            const ParFields &f, uint8_t *status, unsigned long long *n_host, hipStream_t s);
hipError_t launch_par_keep_count(const uint8_t *status, const double *nu, int64_t n_rec, double nu_min, double nu_max,
                                 int64_t *blk, hipStream_t s);
hipError_t launch_par_compact(const uint8_t *status, const ParFields &f, int64_t n_rec, double nu_min, double nu_max,
                              const int64_t *blk_off, const ParFields &g, hipStream_t s);
hipError_t launch_par_sort(const ParFields &g, int64_t n, int64_t n_pad, uint64_t *key, uint32_t *idx,
                           unsigned long long *descents, const ParFields &h, int *sorted_out, hipStream_t s);

// context accessors for code outside irs_api.hip
struct irs_ctx;
hipStream_t ctx_stream(irs_ctx *ctx);
int ctx_device(irs_ctx *ctx);
int ctx_fail(irs_ctx *ctx, int code, const std::string &msg);

