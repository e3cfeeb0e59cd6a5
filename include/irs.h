/*
 * irs.h — C ABI of the MI355X line-by-line engine (libirs.so).
 *
 * Plain C99: pointers, sizes, and POD structs; no torch or HIP types.  Every
 * entry point returns an int status (IRS_OK = 0, negative on error) and never
 * lets a C++ exception cross the boundary; the message of the last failure on
 * a context is available from irs_last_error().  Host arrays are owned by the
 * caller; the context owns all device memory, its HIP stream and its caches.
 * Calls on one context must be serialised by the caller; distinct contexts
 * may be used from distinct threads.
 *
 * Layout conventions are the reference's: uniform FP64 wavenumber grids,
 * layer arrays column-major with wavenumber fastest ([n_nu x n_layers]),
 * levels surface-first, pressures in hPa at profile level and atm at
 * line-shape level, radiance in mW m^-2 sr^-1 (cm^-1)^-1.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   irs_forward            forward_model            src/Sensor/sounder.jl:233-360
 *   irs_faddeeva32         faddeeva_voigt / _voigt_H (Re w) and the Im w of lm_dispersive_kernel!
 *                                                   src/CrossSections/voigt.jl:158-172,
 *                                                   src/CrossSections/line_mixing.jl:1166-1186
 *   irs_xsec_layer         compute_voigt_cross_sections (+ _reject_weak_lines,
 *                          compute_voigt_lm_cross_sections)
 *                                                   src/CrossSections/voigt.jl:333-462,
 *                                                   src/CrossSections/line_mixing.jl:1351-1369
 *   irs_rte                schwarzschild_rte + level_transmittances + add_reflected_downwelling!
 *                                                   src/Solver/schwarzschild.jl:130-209,228-274,
 *                                                   src/Solver/transmittance.jl:26-43
 *   irs_ils_kernel         ils_kernel               src/Sensor/ils.jl:80-114
 *   irs_ils_channels       apply_ils + _resample_to_channels
 *                                                   src/Sensor/ils.jl:131-153,
 *                                                   src/Sensor/sounder.jl:372-377
 *   irs_brightness_temperature  brightness_temperature  src/Solver/planck.jl:43-57
 *   irs_layer_properties   layer_properties         src/Atmosphere/layers.jl:10-178
 *   irs_jacobian           analytic_jacobian        src/Estimation/vmr_jacobian.jl:152-505
 *   irs_state_size         StateVectorSpec length   src/Estimation/state_vector.jl:253-300
 *   irs_linelist_upload    HITRANLinelist(lines)    src/HITRAN/linelist.jl:52-56
 *   irs_set_schedule / irs_get_schedule   detect_backend / ka_backend  src/Parallel/strategy.jl:21-90
 */
#ifndef IRS_H
#define IRS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRS_ABI_VERSION 1

/* status codes */
#define IRS_OK         0
#define IRS_EINVAL    -1   /* a precondition of the reference operator failed */
#define IRS_EHIP      -2   /* HIP runtime error */
#define IRS_ENOMEM    -3   /* device allocation failed */
#define IRS_ENOTFOUND -4   /* unknown handle */
#define IRS_ENODEV    -5   /* no usable gfx950 device */

/* enums mirroring the reference's keyword symbols */
#define IRS_APOD_GAUSSIAN   0   /* :gaussian            ils.jl:90-92  */
#define IRS_APOD_NB_WEAK    1   /* :norton_beer_weak    ils.jl:26-30  */
#define IRS_APOD_NB_MEDIUM  2   /* :norton_beer_medium */
#define IRS_APOD_NB_STRONG  3   /* :norton_beer_strong */

#define IRS_SRC_TOON 0          /* source_function=:toon  schwarzschild.jl:43-47,161-173 */
#define IRS_SRC_CIM  1          /* source_function=:cim   schwarzschild.jl:59-64,174-193 */

#define IRS_TMETHOD_LOGP_AT_PCG   0   /* T_method=:logp_at_pcg   layers.jl:133-178 */
#define IRS_TMETHOD_MASS_WEIGHTED 1   /* T_method=:mass_weighted */

/* continua (sounder.jl:245-246,310-322; continuum.jl:163-327).  The CIA
 * tables are not bundled (HITRAN does not allow redistribution): until one is
 * loaded with irs_cia_set / irs_cia_load_file its pair adds zero, as the
 * reference does when the data file is absent (continuum.jl:254-258). */
#define IRS_CONT_H2O    1       /* :h2o     MT-CKD 4.3 self + foreign (bundled table) */
#define IRS_CONT_CO2    2       /* :co2     MT-CKD CO2 line-coupling continuum (bundled table) */
#define IRS_CONT_CO2CIA 4       /* :co2_cia HITRAN CO2-CO2 CIA, scaled by n_CO2 n_air */
#define IRS_CONT_N2     8       /* :n2      HITRAN N2-N2 CIA */
#define IRS_CONT_O2    16       /* :o2      HITRAN O2-O2 CIA */
#define IRS_CONT_ALL   31

#define IRS_CIA_CO2CO2 0
#define IRS_CIA_N2N2   1
#define IRS_CIA_O2O2   2

#define IRS_LM_NONE 0           /* line_mixing=nothing */
#define IRS_LM_VPY  1           /* first-order (VP_Y) mixing on the CO2 list, Y per line */
#define IRS_LM_RELMAT 2         /* line_mixing=VPYLineMixing(relmat) or VPWLineMixing(relmat):
                                   the perturbation of irs_fm_options.relmat, per band by
                                   relmat_band_mode (line_mixing.jl:144-206) */

/* Per-band treatment of a relaxation matrix (irs_fm_options.relmat_band_mode). */
#define IRS_RELMAT_SKIP 0       /* not evaluated (filtered, or li > 8, or no W0/B0 table) */
#define IRS_RELMAT_VPY  1       /* dispersive term with Y(T) (line_mixing.jl:1193-1349) */
#define IRS_RELMAT_VPW  2       /* eigenmodes of D + i p W minus the no-coupling limit
                                   (line_mixing.jl:892-1091) */

typedef struct irs_ctx irs_ctx;

/* Line list, structure of arrays, sorted by unshifted wavenumber
 * (HITRANLine fields, linelist.jl:20-31).  y_lm may be NULL. */
typedef struct {
    int64_t       n;
    const int8_t *mol;
    const int8_t *iso;
    const double *nu;          /* cm^-1 */
    const double *S;           /* cm^-1/(molec cm^-2) at 296 K */
    const float  *gamma_air;   /* cm^-1/atm */
    const float  *gamma_self;  /* cm^-1/atm */
    const double *elower;      /* cm^-1 */
    const float  *n_air;
    const float  *delta_air;   /* cm^-1/atm */
    const double *y_lm;        /* atm^-1, first-order line-mixing coefficient, or NULL */
} irs_linelist;

/* HITRAN relaxation-matrix data (HITRANRelmatData, line_mixing.jl:35-117): bands
 * of S-file lines, each band's lines contiguous, and the W0/B0 tables keyed by
 * (min(li, lf), max(li, lf)), one entry per (sub-table, Ji, Jip); sub-table 0..8 =
 * pp pq pr qp qq qr rp rq rr (branch types of the initial and the coupled line). */
typedef struct {
    int32_t        n_bands;
    const int8_t  *li, *lf, *isot, *molecule;   /* [n_bands]; isot 10 for HITRAN's '0' */
    const int64_t *line_off;                    /* [n_bands + 1] */
    const int16_t *Ji;                          /* per line */
    const int8_t  *branch;                      /* -1 P, 0 Q, +1 R */
    const double  *nu, *gV_air, *E_lower, *Dipo0, *PopuT0, *DipoT;
    const float   *n_air, *shift;
    int32_t        n_tables;
    const int8_t  *table_lli, *table_llf;       /* [n_tables] */
    const int64_t *entry_off;                   /* [n_tables + 1] */
    const int8_t  *entry_sub;
    const int32_t *entry_ji, *entry_jip;
    const double  *entry_w0, *entry_b0;
} irs_relmat;

/* Sounder (sounder.jl:33-40) */
typedef struct {
    double  nu_min, nu_max, dnu;
    double  opd_max, fwhm_gauss;
    int64_t n_channels;
} irs_sounder;

/* One profile: levels surface first (profiles.jl:12-17) plus the scene
 * scalars forward_model takes as keywords. */
typedef struct {
    int            n_levels;
    const double  *p_hpa;       /* [n_levels] */
    const double  *T;           /* [n_levels] */
    int            n_species;
    const int     *species;     /* HITRAN molecule id of each species */
    const int32_t *linelist;    /* handle per species from irs_linelist_upload, -1 = none */
    const double  *vmr;         /* [n_species][n_levels] */
    double         T_sfc;       /* <= 0: use T[0] (sounder.jl:337) */
    double         eps_sfc;
    double         mu;          /* cos(viewing zenith angle) */
} irs_profile;

/* forward_model keywords (sounder.jl:233-252), same defaults via
 * irs_default_options(). */
typedef struct {
    double internal_dnu;        /* 0.001 */
    int    high_res_factor;     /* <= 0: derive from internal_dnu */
    double cutoff;              /* 25 */
    double lm_cutoff;           /* = cutoff */
    double dptmn;               /* 1e-6 */
    double x_far;               /* 122 */
    int    with_ils;            /* 1 */
    int    apodization;         /* IRS_APOD_GAUSSIAN */
    int    line_mixing;         /* IRS_LM_NONE */
    int    T_method;            /* IRS_TMETHOD_LOGP_AT_PCG */
    int    source_function;     /* IRS_SRC_CIM */
    int    continua;            /* apply_continuum + continua: bit mask of IRS_CONT_* added to
                                   every layer (0 = apply_continuum=false); default IRS_CONT_ALL */
    int32_t relmat;             /* handle from irs_relmat_upload (line_mixing = IRS_LM_RELMAT), -1 */
    int32_t lm_species;         /* HITRAN molecule id the relaxation matrix applies to (2) */
    const uint8_t *relmat_band_mode;  /* [n_bands] IRS_RELMAT_* (line_mixing = IRS_LM_RELMAT) */
    double keep_threshold;      /* 1e-4: VP_W poles below this fraction of the largest |A|
                                   are dropped (line_mixing.jl:1065-1073) */
} irs_fm_options;

/* Timing and work counters of the last irs_forward / irs_xsec_layer call. */
typedef struct {
    double  ms_total;           /* host wall time of the call */
    double  ms_prep;            /* line-parameter + DPTMIN kernels */
    double  ms_linesum;         /* line-sum kernel (HIP events on the context stream) */
    double  ms_rte;
    double  ms_ils;
    int64_t n_hi;               /* internal grid points */
    int64_t n_layers;
    int64_t linesum_launches;
    int64_t active_lines;       /* sum over (profile, layer, species) */
    int64_t active_pairs;       /* nominal pairs |nu - nu0| <= cutoff over active lines */
    int64_t dispersive_pairs;   /* nominal VP_Y pairs */
    int64_t core_pairs_bound;   /* pairs in core ranges (evaluated with w(z)) */
    double  ms_device;          /* device time of the whole call (graph or kernels) */
    int64_t graph_launches;     /* profile groups run from a captured graph */
    int32_t tile_points;        /* schedule used by the line sum */
    int32_t persistent;
    int32_t autotuned;          /* 1: this call ran the autotuner */
    int32_t reserved1;
    int64_t fine_pairs;         /* line-sum work executed: pairs on the internal grid */
    int64_t coarse_pairs;       /* pairs on the coarse far-wing nodes */
    int64_t fine_pairs_lm;      /* of which on line-mixing (VP_Y) items */
    int64_t coarse_pairs_lm;
    int64_t expanded_line_slices; /* lines x 64-point slices evaluated by local expansion
                                     (cutoff-boundary and far FAST lines) */
    double  ms_continuum;       /* continuum kernel (0 with apply_continuum = false) */
} irs_stats;

/* Line-sum schedule of irs_forward.  No counterpart in the reference: it
 * replaces the backend/thread choice of src/Parallel/strategy.jl:21-90. */
typedef struct {
    int tile_points;            /* grid points per thread per tile: 4, 8 or 16 */
    int persistent;             /* 1 (default): persistent work-queue kernel over heaviest-first
                                   tiles; 0: one workgroup per (tile, layer, profile) */
    int use_graph;              /* 1: capture each profile group's device sequence in a hipGraph
                                   and replay it on later calls of the same shape */
    int autotune;               /* 1: time the (tile_points, persistent) candidates on the first
                                   profile of each new shape and keep the fastest; with the
                                   environment variable IRS_TUNE_FILE naming a file, results are
                                   appended there (keyed by device, grid and line-list content)
                                   and later contexts and processes reuse them untimed */
    double far_wing_cm;         /* > 0: far-wing sums of lines at least this far (cm^-1) from a
                                   tile are evaluated on a coarse grid (<= 0.032 cm^-1) and
                                   interpolated (cubic), and lines far from a 64-point slice
                                   (interior lines beyond 500 grid steps, cutoff-boundary
                                   lines) by local expansion in the slice; 0: every pair on
                                   the internal grid */
    int fp64_every_pair;        /* 1: irs_forward runs the reference's every-pair line sum in
                                   FP64 (FP64 optical depths, one profile at a time, no graph):
                                   the plain FP64 figure beside the FP32 engine */
    int reserved;
} irs_schedule;

int         irs_version(void);
/* Fills buf with "name;gcnArch;CUs;clockMHz;memGB".  Returns IRS_ENODEV without a GPU. */
int         irs_device_info(int device, char *buf, int buflen);
int         irs_create(int device, irs_ctx **out);
void        irs_destroy(irs_ctx *ctx);
const char *irs_last_error(const irs_ctx *ctx);
void        irs_default_options(irs_fm_options *opt);
int         irs_last_stats(const irs_ctx *ctx, irs_stats *out);

int         irs_set_schedule(irs_ctx *ctx, const irs_schedule *schedule);
int         irs_get_schedule(const irs_ctx *ctx, irs_schedule *schedule);

/* Per-layer channel optical depths of a batch of profiles along each profile's
 * viewing path (mu): the emulator's training target (no counterpart in the
 * reference).  With T_k = ILS (x) exp(-sum_{j>=k} tau_j / mu) at the channel node
 * (level-to-space) and t_k = ILS (x) exp(-tau_k / mu) (layer alone):
 *   tau_ch[k] = ln(T_{k+1} / T_k)  where T_{k+1} >= floor,  else  -ln t_k,
 * clamped to [0, 50].  tau_ch is [n_prof][n_layers][n_channels] (float); the
 * channels must be internal grid nodes.  cim_g (optional, same shape) is the
 * channel weight of the CIM source correction, ILS (x) [(T_{k+1} - T_k) f_cim(tau_k/mu)]
 * / T^ch_{k+1} (layer-alone form under an opaque column), clamped to [0, 1]. */
int irs_channel_od(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                   const irs_fm_options *opt, double floor, float *tau_ch, float *cim_g);
/* irs_channel_od with cim_g and the normalised ILS first moment of each layer's
 * transmittance step (no counterpart in the reference; equally spaced node
 * channels only): ils_mom[k] = ILS (x) [(nu - nu_c)(T_{k+1} - T_k)] / T^ch_{k+1}
 * where T^ch_{k+1} >= floor, else 0, so that Planck's slope across the ILS enters
 * the channel RTE: B(T, nu) (T_{k+1} - T_k) -> B(T, nu_c) (T^ch_{k+1} - T^ch_k)
 * + dB/dnu(T, nu_c) ils_mom[k] T^ch_{k+1}. */
int irs_channel_od_moments(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                           const irs_fm_options *opt, double floor, float *tau_ch, float *cim_g,
                           float *ils_mom);
/* The emulator's target (no counterpart in the reference): ILS-convolved
 * level-to-space transmittances along each profile's path T_lev
 * [n_prof][n_layers+1][n_channels] (index n_layers = top, 1) and the CIM sums
 * cim_G [n_prof][n_layers][n_channels] = ILS (x) [(T_{k+1} - T_k) f_cim(tau_k/mu)],
 * so that  R_ch = eps B(T_sfc) T_0 + sum_k B(T_ave,k)(T_{k+1} - T_k)
 *                 + (B(T_{k+1}) - B(T_ave,k)) G_k  (Planck at the channel centre). */
int irs_channel_transmittance(irs_ctx *ctx, const irs_profile *profiles, int n_prof,
                              const irs_sounder *snd, const irs_fm_options *opt, float *T_lev,
                              float *cim_G);

/* HITRAN CIA table of one pair (IRS_CIA_*): n_blocks temperature blocks, block
 * b with knots nu[off[b] .. off[b+1]) and cross sections sigma (cm^5 molec^-2,
 * negative values are clamped to 0 as continuum.jl:131-134 does) at T[b].
 * n_blocks = 0 removes the table.  Replaces _load_cia_table + DataTable
 * (src/CrossSections/continuum.jl:106-148). */
int irs_cia_set(irs_ctx *ctx, int pair, int n_blocks, const double *T, const int64_t *off,
                const double *nu, const double *sigma);
/* Reads a HITRAN .cia text file (header line per block: pair nu_min nu_max n T ...,
 * then "nu sigma" rows) into the pair's table (continuum.jl:112-145). */
int irs_cia_load_file(irs_ctx *ctx, int pair, const char *path);
/* Continuum of one layer on the grid (nu0, dnu, n_nu), FP64 on the device:
 * tau_c[i] = dz_cm * the sum of the selected continua (absorption coefficients in
 * cm^-1) at the given VMRs, pressure and temperature.  With one bit and dz_cm = 1
 * this is h2o_continuum / co2_continuum / co2_cia / n2_cia / o2_cia
 * (continuum.jl:163-327); forward_model uses the mid-layer state and
 * dz = _dp_to_dz_cm(dp, p_mid, T_mid) (sounder.jl:310-322; optical_depth.jl:123-130). */
int irs_continuum_layer(irs_ctx *ctx, double nu0, double dnu, int64_t n_nu, int continua,
                        double vmr_h2o, double vmr_co2, double vmr_n2, double vmr_o2, double p_hpa,
                        double T, double dz_cm, double *tau_c);

/* Keeps a relaxation matrix in the context (host memory: the per-layer W(T),
 * Y(T) and eigenmodes are built on the host, their term sums run on the
 * device) and builds every band's coupling table (line_mixing.jl:530-578). */
int irs_relmat_upload(irs_ctx *ctx, const irs_relmat *rm, int32_t *handle);
int irs_relmat_free(irs_ctx *ctx, int32_t handle);

/* The line-mixing perturbation of a relaxation matrix at one (T, p_atm) on the
 * grid nodes of (nu0, dnu): dsigma[n_nu] in cm^2/molec, per band by band_mode
 * (IRS_RELMAT_*): _lm_perturbation of VPYLineMixing / VPWLineMixing
 * (line_mixing.jl:283-303), without the Voigt baseline and clamp. */
int irs_lm_perturbation(irs_ctx *ctx, int32_t relmat, const uint8_t *band_mode, double nu0, double dnu,
                        int64_t n_nu, double T, double p_atm, double lm_cutoff, double x_far,
                        double keep_threshold, double *dsigma);

/* Host-only: the terms irs_lm_perturbation sums, for a relaxation matrix given
 * directly (no context or device): up to `capacity` rows of {c, f, y, a, b} (the
 * term adds a Re w + b Im w of w((nu - c) f + i y) where |nu - c| <= cutoff),
 * sorted by c; *n_terms is the full count. */
int irs_relmat_terms(const irs_relmat *rm, const uint8_t *band_mode, double T, double p_atm,
                     double keep_threshold, int64_t capacity, double *terms, int64_t *n_terms);

/* Copies a sorted line list to the device; returns a handle. */
int irs_linelist_upload(irs_ctx *ctx, const irs_linelist *ll, int32_t *handle);
int irs_linelist_free(irs_ctx *ctx, int32_t handle);

/* w(x + iy) = exp(-z^2) erfc(-iz), y >= 0, for n host points, evaluated on the
 * device by the FP32 routines the line sums use: mode 0 follows the CORE range
 * (inside |z| < 8 the trapezoid rule for y < 6, else the continued fraction),
 * mode 1 the per-pair ranges and the gradient sum (trapezoid for |x| + y < 5.5).
 * Re w is the Voigt H function (faddeeva_voigt); Im w enters the VP_Y term. */
int irs_faddeeva32(irs_ctx *ctx, const double *x, const double *y, int64_t n, int mode, double *w_re,
                   double *w_im);

/* Cross sections of one species in one layer on nu_i = grid node i of
 * (nu0, dnu): sigma[n_nu] in cm^2/molec.  dptmn > 0 applies DPTMIN with the
 * given coef (= vmr * dp * N_air); line_mixing = IRS_LM_VPY adds the VP_Y
 * dispersive term of the lines with y_lm != 0 (and skips DPTMIN, as
 * sounder.jl:295-299 does on the line-mixing CO2 path). */
int irs_xsec_layer(irs_ctx *ctx, int32_t handle, double nu0, double dnu, int64_t n_nu,
                   double T, double p_atm, double vmr_self, double coef, double dptmn,
                   double cutoff, double lm_cutoff, double x_far, int line_mixing,
                   double *sigma);

/* Size of the fixed device scratch array irs_xsec_layer_noalloc works in: one
 * call needs about 172 bytes per line of the list plus 4 bytes per grid point
 * (each part rounded up to 256 bytes), e.g. a 300 000-line list on the
 * 2 147 001-point full IASI grid. */
#define IRS_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* irs_xsec_layer without any device allocation: the same arguments, kernels
 * and result (bit-identical sigma), with every device buffer taken from a
 * fixed IRS_NOALLOC_SCRATCH_BYTES array of the library that each GPU holds
 * until the process ends.  Returns IRS_EINVAL, naming the bytes needed, when
 * the list and grid do not fit.  The list must have been uploaded before (the
 * upload allocates).  All calls on one GPU share that GPU's array: calls that
 * overlap in time on one GPU (several threads, several contexts) race and must
 * be serialised by the caller.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
int irs_xsec_layer_noalloc(irs_ctx *ctx, int32_t handle, double nu0, double dnu, int64_t n_nu,
                           double T, double p_atm, double vmr_self, double coef, double dptmn,
                           double cutoff, double lm_cutoff, double x_far, int line_mixing,
                           double *sigma);

/* ---- HITRAN .par ingest ---------------------------------------------------
 * Output arrays of the parsed records (host memory, `capacity` entries each). */
typedef struct {
    int8_t *mol, *iso;
    double *nu;                /* cm^-1 */
    double *S;                 /* cm^-1/(molec cm^-2) at 296 K */
    double *A;                 /* s^-1 */
    double *elower;            /* cm^-1 */
    float  *gamma_air, *gamma_self, *n_air, *delta_air;
} irs_par_lines;

#define IRS_LINELIST_CACHE_VERSION 1

/* load_hitran_par over a buffer (src/HITRAN/fetch.jl:12-61, the record parser
 * parse_hitran_par, src/HITRAN/linelist.jl:90-103), parsed on the device: each
 * newline-terminated record (a trailing CR removed) of at least 67 characters
 * whose ten fixed-width fields parse is kept if nu_min <= nu <= nu_max (pass
 * -inf / +inf for no bound); records shorter than 67 characters or with a field
 * that does not parse are skipped.  The kept lines are returned sorted stably by
 * wavenumber (HITRANLinelist, linelist.jl:51-56).  capacity >= n_bytes / 67 + 1
 * always suffices.  n_skipped (may be NULL): records not parsed. */
int irs_par_parse(irs_ctx *ctx, const char *bytes, int64_t n_bytes, double nu_min, double nu_max,
                  int64_t capacity, const irs_par_lines *out, int64_t *n_lines, int64_t *n_skipped);

/* load_linelist (src/HITRAN/cache.jl:153-175): the .par file at par_path parsed
 * whole (irs_par_parse) or read from its cache file in cache_dir (NULL or "": no
 * cache), then the window [nu_min, nu_max] applied.  The cache file
 * (cache_dir/linelist_<key>.irsll, key = hash of the absolute path, size,
 * modification time and IRS_LINELIST_CACHE_VERSION) is written atomically when
 * missing or when rebuild != 0.  IRS_ENOTFOUND names a missing file. */
int irs_par_load(irs_ctx *ctx, const char *par_path, const char *cache_dir, int rebuild, double nu_min,
                 double nu_max, int64_t capacity, const irs_par_lines *out, int64_t *n_lines);

/* Schwarzschild RTE on a host tau [n_nu x n_layers]; T_ave required for CIM. */
int irs_rte(irs_ctx *ctx, double nu0, double dnu, int64_t n_nu, const double *tau, int n_layers,
            const double *T_lev, double T_sfc, double mu, double eps_sfc, int source_function,
            const double *T_ave, double *radiance);

/* ILS taps: out[2*half+1], half = ceil(16/dnu) (ils.jl:80-114). */
int irs_ils_half(double dnu, int64_t *half);
int irs_ils_kernel(irs_ctx *ctx, double dnu, double opd_max, double fwhm, int apodization,
                   int n_pts, double *taps);

/* ILS convolution of a host spectrum on (nu0, dnu, n_nu) sampled at the sounder
 * channels (exact-node or linear interpolation, Flat extrapolation). */
int irs_ils_channels(irs_ctx *ctx, const double *spectrum, double nu0, double dnu, int64_t n_nu,
                     const irs_sounder *snd, int apodization, double *r_ch);

int irs_brightness_temperature(const double *nu, const double *radiance, int64_t n, double *bt);

/* layer_properties (src/Atmosphere/layers.jl:10-178) of one profile, host only:
 * p_mid, dp, T_ave (TBAR) [n_layers]; vmr_cg, p_cg, T_cg [n_species][n_layers]. */
int irs_layer_properties(const irs_profile *profile, int T_method, double *p_mid, double *dp,
                         double *T_ave, double *vmr_cg, double *p_cg, double *T_cg);

/* Geometry of the internal grid for (snd, opt). */
int irs_forward_dims(const irs_sounder *snd, const irs_fm_options *opt, double *nu_lo,
                     double *dnu_hi, int64_t *n_hi, int64_t *n_channels);

/* Retrieval state layout (StateVectorSpec, src/Estimation/state_vector.jl:253-300):
 * [T levels][VMR block per retrieved species][T_sfc][emissivity]. */
typedef struct {
    int        include_temperature;
    int        n_retrieved;        /* <= 4 */
    const int *retrieved;          /* HITRAN ids, each must be one of the profile's species */
    const int *column_scale;       /* per retrieved species: 1 = ColumnScale, 0 = FullProfile (may be NULL) */
    int        include_tsfc;
    int        include_emissivity;
    int        log_vmr;
    int        vmr_coupling;       /* CG pressure / temperature / self-broadening coupling */
    int        observable;         /* 0 brightness temperature, 1 radiance */
} irs_state_spec;

int irs_state_size(const irs_state_spec *spec, int n_levels);

/* analytic_jacobian (src/Estimation/vmr_jacobian.jl:152-505) for one profile,
 * full line set (no DPTMIN).  K is column-major [n_state][n_channels]; y0 the
 * observable at the linearisation point [n_channels]; x0 the packed state. */
int irs_jacobian(irs_ctx *ctx, const irs_profile *profile, const irs_sounder *snd,
                 const irs_fm_options *opt, const irs_state_spec *spec, double *K, double *y0,
                 double *x0);

/* ---- Emulator (no counterpart in the reference) -------------------------
 * A device-resident MLP  n_in -> 3 x n_hidden (GELU) -> n_out, trained and
 * evaluated with hand-written BF16 MFMA kernels (FP32 accumulation, FP32
 * master weights, Adam).  Parameters are packed W1[n_in_pad][H], b1[H],
 * W2[H][H], b2, W3[H][H], b3, W4[H][n_out], b4 (rows of W1 beyond n_in are 0).
 * Batches are multiples of 64 rows; n_hidden and n_out multiples of 64. */
typedef struct irs_mlp irs_mlp;
int         irs_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, uint64_t seed,
                           irs_mlp **out);
void        irs_mlp_destroy(irs_mlp *mlp);
const char *irs_mlp_last_error(const irs_mlp *mlp);
int64_t     irs_mlp_n_params(const irs_mlp *mlp);
int         irs_mlp_input_pad(const irs_mlp *mlp);
int         irs_mlp_get_params(irs_mlp *mlp, float *params);
int         irs_mlp_set_params(irs_mlp *mlp, const float *params);      /* resets the Adam state */
/* y[n][n_out] = MLP(x[n][n_in]) */
int         irs_mlp_forward(irs_mlp *mlp, const float *x, int64_t n, float *y);
/* loss = mean((MLP(x) - t)^2) over B x n_out, grad = d loss / d params (packed) */
int         irs_mlp_loss_grad(irs_mlp *mlp, const float *x, const float *t, int B, float *grad, float *loss);
/* training set kept on the device: X[n][n_in], T[n][n_out] */
int         irs_mlp_set_data(irs_mlp *mlp, const float *X, const float *T, int64_t n);
/* steps of Adam (b1 0.9, b2 0.999, eps 1e-8) on shuffled minibatches; losses[steps] (may be NULL) */
int         irs_mlp_train(irs_mlp *mlp, int steps, int batch, float lr, uint64_t seed, float *losses);
/* J[n][n_out][n_in] = dy/dx (forward-mode tangents) */
int         irs_mlp_input_jacobian(irs_mlp *mlp, const float *x, int64_t n, float *J);

/* Emulated forward model.  Per layer k and channel c three targets, the layer
 * channel transmittance t_k, the CIM weight g_k and the normalised ILS first
 * moment m_k of irs_channel_od_moments, decoded in the layer's own basis:
 *   y_k(c) = mean[k][c] + inv_w[k][c] sum_m a_{k,m} basis[k][m][c]
 *   (mean, inv_w [n_layers][3 n_channels], basis [n_layers][n_pc][3 n_channels];
 *   t on columns 0..n_ch-1, g on n_ch..2 n_ch-1, m on 2 n_ch..3 n_ch-1), with
 *   a_k = A_k phi(z_k) + coef_scale MLP(z_k[0 .. n_in-1]),
 *   z_k = (features_k - x_mean_k) / x_std_k  (n_feat raw features per layer,
 *        x_mean and x_std [n_layers][n_feat]),
 *   phi = [1, z_i (i < n_local), z_a z_b (a <= b < n_local), z_i (n_local <= i < n_feat)],
 *   A [n_layers][n_phi][n_pc], n_phi = 1 + n_local + n_local (n_local + 1) / 2 + n_feat - n_local;
 * then the channel RTE with Planck linear in nu across the ILS:
 *   R = eps [B(T_sfc) T_0 + B'(T_sfc) M_s] + sum_k T_{k+1} [B(T_ave,k)(1 - t_k)
 *       + (B(T_{k+1}) - B(T_ave,k)) g_k + B'(T_ave,k) m_k],  M_s = -sum_k m_k T_{k+1},
 *   B' = dB/dnu at the channel centre, T_L = 1, T_k = t_k T_{k+1}
 * (CIM source only; for eps < 1 the reflected term is approximate).  n_feat <= 24. */
typedef struct irs_emu irs_emu;
int         irs_emu_create(irs_mlp *mlp, int n_layers, int n_pc, int n_channels, int n_feat, int n_local,
                           const float *x_mean, const float *x_std, const float *A, float coef_scale,
                           const float *basis, const float *mean, const float *inv_w,
                           const double *ch_nu, irs_emu **out);
void        irs_emu_destroy(irs_emu *emu);
const char *irs_emu_last_error(const irs_emu *emu);
/* features[n_prof][L][n_feat] (raw), T_lev[n_prof][L+1], T_ave[n_prof][L], scal[n_prof][3] =
 * {T_sfc, mu, eps}; R, BT [n_prof][n_channels]; t_out[n_prof][L][n_channels] (decoded T_k) optional */
int         irs_emu_forward(irs_emu *emu, const float *features, const double *T_lev, const double *T_ave,
                            const double *scal, int n_prof, int source_function, double *R, double *BT,
                            float *t_out);

/* Profile-context features of the emulator (n_ctx = n_feat - n_local): context_c =
 * sum_j proj[c][j] (v_j - mean_j), v = [T levels, ln q_H2O levels, ln q_O3 levels (floor
 * 1e-30), T_sfc, 1/mu] of the scene (n_vec = 3 n_levels + 2).  No counterpart in the
 * reference. */
int         irs_emu_set_context(irs_emu *emu, int n_ctx, int n_vec, const double *mean, const double *proj);

/* Emulator + per-scene LBL spot check + full-LBL fallback in one call (the learned
 * path behind forward_model's signature; no counterpart in the reference).  The
 * emulator answers every scene (n_local must be 13: per layer ln p_mid, T_ave, ln vmr
 * H2O, ln vmr O3, CO2 ppm, 1/mu, ln H2O and O3 columns above, T_ave of the layer
 * above, and the H2O-weighted ln p and T, O3-weighted T and mass-weighted T of the
 * layers above; the library computes them); block b of consecutive channels [block_c0[b],
 * block_c0[b] + block_n[b]) is recomputed by LBL on its own sub-grid with the line
 * lists block_lists[b][n_species] (the caller's lists filtered to the block
 * +- 42 cm^-1); a scene whose largest |BT_emulator - BT_LBL| over the spot channels
 * exceeds `threshold` is recomputed with full LBL (irs_forward with opt) and
 * fell_back[q] = 1.  With a correction set (irs_emu_set_correction, whose blocks these
 * must be) every channel of every scene is first corrected from the scene's spot errors,
 * the spot channels take their LBL values, and the fallback test compares the
 * correction's statistic with `threshold` instead.  R, BT [n_prof][n_channels];
 * spot_err [n_prof] (the tested statistic) may be NULL. */
int         irs_forward_emulated(irs_ctx *ctx, irs_emu *emu, const irs_profile *profiles, int n_prof,
                                 const irs_sounder *snd, const irs_fm_options *opt, int n_blocks,
                                 const int64_t *block_c0, const int32_t *block_n, const int32_t *block_lists,
                                 double threshold, double *R, double *BT, uint8_t *fell_back, double *spot_err);

/* Spot-check correction of the emulator (no counterpart in the reference): the scene's
 * errors at the spot channels, e = BT_emulator[S] - BT_LBL[S] over the n_blocks blocks
 * [block_c0[b], block_c0[b] + block_n[b]) (irs_forward_emulated must get the same blocks
 * in the same order; at most 1024 channels), move every channel by the linear map fitted on calibration scenes,
 * BT[c] -= sum_s e[s] W[s][c] (W [n_spot][n_channels]).  The fallback statistic is
 * max(w_raw max|e|, w_lbo r), r the largest leave-one-block-out residual
 * max_b max |e_b - e_{S without b} M_b| with lbo_W the maps M_b [n_spot - n_b][n_b],
 * concatenated over b (may be NULL when w_lbo = 0).  n_blocks = 0 removes the correction. */
int         irs_emu_set_correction(irs_emu *emu, int n_blocks, const int64_t *block_c0, const int32_t *block_n,
                                   const float *W, const float *lbo_W, double w_raw, double w_lbo);

/* Loads a trained and calibrated emulator from the file written by the Python package
 * (ChannelODEmulator.export_binary): creates its MLP on `device` (max_batch rows per
 * launch) and the emulator with its profile context and, when the file holds one, its
 * spot-check correction; returns the calibration's threshold and spot blocks (at most
 * max_blocks) for irs_forward_emulated.  Destroy the emulator, then the MLP.  No
 * counterpart in the reference (the learned path's loader for Julia and C callers). */
int         irs_emu_load(const char *path, int device, int max_batch, irs_mlp **mlp, irs_emu **emu,
                         double *threshold, int max_blocks, int64_t *block_c0, int32_t *block_n, int *n_blocks);

/* Emulator Jacobians (chain rule through the MLP input Jacobian, the regression,
 * the decode and the adjoint of the channel RTE; the reflected term of eps < 1 held fixed):
 * BT [n_prof][n_ch]; dfeat [n_prof][L][n_feat][n_ch] = dBT/d(raw feature);
 * dTave [n_prof][L][n_ch], dTlev [n_prof][L+1][n_ch], dTsfc [n_prof][n_ch]: the
 * Planck terms of dBT/dT. */
int         irs_emu_jacobian(irs_emu *emu, const float *features, const double *T_lev,
                             const double *T_ave, const double *scal, int n_prof, double *BT,
                             float *dfeat, float *dTave, float *dTlev, float *dTsfc);

/* forward_model over a batch of profiles: R, BT are [n_prof][n_channels]. */
int irs_forward(irs_ctx *ctx, const irs_profile *profiles, int n_prof, const irs_sounder *snd,
                const irs_fm_options *opt, double *R, double *BT);

#ifdef __cplusplus
}
#endif
#endif
